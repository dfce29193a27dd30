"""Pins the oracle's restatement of Jet's pressure solvers (system build,
markers, coarse levels, gradient) and level-set helpers.

Jet's pressure solvers cannot be compiled here (they pull in TriangleMesh3,
whose source needs the absent tinyobj submodule), so they are pinned by
Jet's own unit-test known answers
(src/tests/unit_tests/grid_single_phase_pressure_solver3_tests.cpp,
src/tests/unit_tests/grid_fractional_single_phase_pressure_solver3_tests.cpp),
by the discrete identity div(u') = b - A p that projection must satisfy, by
closed forms, and (for fractionInside*) bitwise against Jet's header
templates compiled into oracle/_ref.
"""
import numpy as np
import pytest

BIG = np.finfo(np.float64).max


def cc(fn, n, h=1.0):
    nx, ny, nz = n
    z, y, x = np.meshgrid(h * (np.arange(nz) + 0.5), h * (np.arange(ny) + 0.5),
                          h * (np.arange(nx) + 0.5), indexing="ij")
    return np.ascontiguousarray(fn(x, y, z), dtype=np.float64)


def vertices(fn, n, h=1.0):
    nx, ny, nz = n
    z, y, x = np.meshgrid(h * np.arange(nz + 1), h * np.arange(ny + 1), h * np.arange(nx + 1),
                          indexing="ij")
    return np.ascontiguousarray(fn(x, y, z), dtype=np.float64)


def kat_velocity():
    """Jet's test field: v = 1 on interior y-faces (j = 1, 2), 0 elsewhere."""
    u = np.zeros((3, 3, 4))
    v = np.zeros((3, 4, 3))
    w = np.zeros((4, 3, 3))
    v[:, 1:3, :] = 1.0
    return u, v, w


def solve_single_phase(oracle, u, v, w, bsdf, fsdf, h=(1.0, 1.0, 1.0)):
    m = oracle.sp_markers(bsdf, fsdf)
    A, b = oracle.sp_build_system(m, u, v, w, h)
    p, it, res, ok = oracle.cg_solve(A, b, 200, 1e-12)
    uo, vo, wo = oracle.sp_apply_gradient(m, p, u, v, w, h)
    return m, A, b, p, (uo, vo, wo)


def test_kat_single_phase(oracle):
    # GridSinglePhasePressureSolver3.SolveSinglePhase (:14-67)
    u, v, w = kat_velocity()
    n = (3, 3, 3)
    _, _, _, p, (uo, vo, wo) = solve_single_phase(
        oracle, u, v, w, cc(lambda x, y, z: BIG + 0 * x, n), cc(lambda x, y, z: -BIG + 0 * x, n))
    for a in (uo, vo, wo):
        assert np.abs(a).max() < 1e-6
    np.testing.assert_allclose(p[:, 1:, :] - p[:, :-1, :], -1.0, atol=1e-6)


def test_kat_single_phase_with_boundary(oracle):
    # ...WithBoundary (:69-130): solid wall on the right-most column.
    u, v, w = kat_velocity()
    n = (3, 3, 3)
    _, _, _, p, (uo, vo, wo) = solve_single_phase(
        oracle, u, v, w, cc(lambda x, y, z: -x + 2.0, n), cc(lambda x, y, z: -BIG + 0 * x, n))
    assert np.abs(uo).max() < 1e-6 and np.abs(wo).max() < 1e-6
    expect = np.zeros_like(vo)
    expect[:, 1:3, 2] = 1.0
    np.testing.assert_allclose(vo, expect, atol=1e-6)
    np.testing.assert_allclose(p[:, 1:, :2] - p[:, :-1, :2], -1.0, atol=1e-6)


def test_kat_free_surface(oracle):
    # ...SolveFreeSurface (:132-189): fluidSdf = y - 2 -> p = 2 - j.
    u, v, w = kat_velocity()
    n = (3, 3, 3)
    _, _, _, p, (uo, vo, wo) = solve_single_phase(
        oracle, u, v, w, cc(lambda x, y, z: BIG + 0 * x, n), cc(lambda x, y, z: y - 2.0, n))
    for a in (uo, vo, wo):
        assert np.abs(a).max() < 1e-6
    for j in range(2):
        np.testing.assert_allclose(p[:, j, :], 2.0 - j, atol=1e-6)


def test_kat_free_surface_with_boundary(oracle):
    # ...SolveFreeSurfaceWithBoundary (:250-314)
    u, v, w = kat_velocity()
    n = (3, 3, 3)
    _, _, _, p, (uo, vo, wo) = solve_single_phase(
        oracle, u, v, w, cc(lambda x, y, z: -x + 2.0, n), cc(lambda x, y, z: y - 2.0, n))
    assert np.abs(uo).max() < 1e-6 and np.abs(wo).max() < 1e-6
    expect = np.zeros_like(vo)
    expect[:, 1:3, 2] = 1.0
    np.testing.assert_allclose(vo, expect, atol=1e-6)
    for j in range(2):
        np.testing.assert_allclose(p[:, j, :2], 2.0 - j, atol=1e-6)


def test_kat_fractional_free_surface(oracle):
    # GridFractionalSinglePhasePressureSolver3.SolveFreeSurface (:17-74):
    # ghost-fluid theta = 0.5 at the surface -> p = 1.5 - j.
    u, v, w = kat_velocity()
    n = (3, 3, 3)
    fs = cc(lambda x, y, z: y - 2.0, n).astype(np.float32)
    uw, vw, ww = oracle.fr_face_weights(vertices(lambda x, y, z: BIG + 0 * x, n))
    assert uw.min() == vw.min() == ww.min() == 1.0
    A, b = oracle.fr_build_system(fs, uw, vw, ww, u, v, w, (1.0, 1.0, 1.0))
    p, it, res, ok = oracle.cg_solve(A, b, 200, 1e-12)
    uo, vo, wo = oracle.fr_apply_gradient(fs, uw, vw, ww, p, u, v, w, (1.0, 1.0, 1.0))
    for a in (uo, vo, wo):
        assert np.abs(a).max() < 1e-6
    for j in range(2):
        np.testing.assert_allclose(p[:, j, :], 1.5 - j, atol=1e-6)


def random_scene(n, rng, h=1.0):
    """Random solid spheres + tilted free surface + U(-1,1) faces with solid
    and wall faces zeroed (the blocked boundary condition)."""
    nx, ny, nz = n
    cs = [rng.uniform(0, nx * h, 3) for _ in range(3)]
    rs = [rng.uniform(1.0, 0.3 * nx * h) for _ in range(3)]

    def bsdf(x, y, z):
        d = np.full(np.broadcast(x, y, z).shape, BIG)
        for c, r in zip(cs, rs):
            d = np.minimum(d, np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - r)
        return d

    def fsdf(x, y, z):
        return y - (0.6 * ny * h + 0.1 * (x - 0.5 * nx * h))

    u = rng.uniform(-1, 1, (nz, ny, nx + 1))
    v = rng.uniform(-1, 1, (nz, ny + 1, nx))
    w = rng.uniform(-1, 1, (nz + 1, ny, nx))
    return bsdf, fsdf, u, v, w


def weighted_div(u, v, w, h, uw=None, vw=None, ww=None):
    """Discrete divergence; face weights (if given) multiply interior faces
    only, wall faces stay unweighted (Jet's fractional b)."""
    if uw is not None:
        u = u.copy() * 1.0
        v = v.copy() * 1.0
        w = w.copy() * 1.0
        u[:, :, 1:-1] *= uw[:, :, 1:-1]
        v[:, 1:-1, :] *= vw[:, 1:-1, :]
        w[1:-1, :, :] *= ww[1:-1, :, :]
    return ((u[:, :, 1:] - u[:, :, :-1]) / h + (v[:, 1:, :] - v[:, :-1, :]) / h
            + (w[1:, :, :] - w[:-1, :, :]) / h)


@pytest.mark.parametrize("seed,n,h", [(0, (12, 10, 8), 1.0), (1, (9, 9, 9), 0.5), (2, (16, 16, 16), 0.25)])
def test_single_phase_projection_identity(oracle, seed, n, h):
    # After projection, div(u') = b - A p on every fluid cell that has no Air
    # neighbour on its -x/-y/-z side (Jet updates only +side faces of fluid
    # cells, src/jet/grid_single_phase_pressure_solver3.cpp:415-429), for ANY
    # p -- so a random p exercises the build and the gradient together.
    rng = np.random.default_rng(seed)
    bsdf, fsdf, u, v, w = random_scene(n, rng, h)
    m = oracle.sp_markers(cc(bsdf, n, h), cc(fsdf, n, h))
    A, b = oracle.sp_build_system(m, u, v, w, (h, h, h))
    p = rng.uniform(-1, 1, b.shape)
    p[m != 0] = 0.0  # non-fluid rows are (1, 0): their solution is 0
    uo, vo, wo = oracle.sp_apply_gradient(m, p, u, v, w, (h, h, h))
    lhs = weighted_div(uo, vo, wo, h)
    rhs = b - oracle.mvm(A, p)
    air = m == 1
    ok = m == 0
    ok[:, :, 1:] &= ~air[:, :, :-1]
    ok[:, 1:, :] &= ~air[:, :-1, :]
    ok[1:, :, :] &= ~air[:-1, :, :]
    assert ok.sum() > 0.2 * ok.size
    np.testing.assert_allclose(lhs[ok], rhs[ok], rtol=1e-12, atol=1e-10 / h)
    # Non-fluid rows are identity rows with zero right-hand side.
    assert np.all(A[m != 0] == np.array([1.0, 0.0, 0.0, 0.0])) and np.all(b[m != 0] == 0.0)
    # Symmetric positive definite on the fluid block (air present).
    d = rng.uniform(-1, 1, b.shape) * (m == 0)
    assert np.sum(d * oracle.mvm(A, d)) > 0.0


@pytest.mark.parametrize("seed,n,h", [(3, (12, 10, 8), 1.0), (4, (16, 16, 16), 0.5)])
def test_fractional_projection_identity(oracle, seed, n, h):
    # Weighted divergence after the fractional projection equals b - A p on
    # every fluid row (zero collider velocity), for any p.
    rng = np.random.default_rng(seed)
    bsdf, fsdf, u, v, w = random_scene(n, rng, h)
    fs = cc(fsdf, n, h).astype(np.float32)
    uw, vw, ww = oracle.fr_face_weights(vertices(bsdf, n, h))
    A, b = oracle.fr_build_system(fs, uw, vw, ww, u, v, w, (h, h, h))
    p = rng.uniform(-1, 1, b.shape)
    fluid = fs < 0
    p[~fluid] = 0.0
    uo, vo, wo = oracle.fr_apply_gradient(fs, uw, vw, ww, p, u, v, w, (h, h, h))
    lhs = weighted_div(uo, vo, wo, h, uw.astype(np.float64), vw.astype(np.float64),
                       ww.astype(np.float64))
    rhs = b - oracle.mvm(A, p)
    rows = fluid & (A[..., 0] != 1.0)
    assert rows.sum() > 0.2 * rows.size
    np.testing.assert_allclose(lhs[rows], rhs[rows], rtol=1e-9, atol=1e-9 / h)
    assert 0.0 < uw.min() or uw.min() == 0.0
    assert uw.max() <= 1.0 and (uw[(uw > 0)] >= 0.01).all()


def test_fraction_inside_bitwise_vs_jet(oracle, ref):
    rng = np.random.default_rng(5)
    for _ in range(4000):
        phis = rng.uniform(-1, 1, 4) * rng.choice([1e-3, 1.0, 10.0])
        if rng.random() < 0.1:
            phis[rng.integers(0, 4)] = 0.0
        assert oracle.fraction_inside(*phis) == ref.fraction_inside(*phis)
        assert oracle.fraction_inside_sdf(*phis[:2]) == ref.fraction_inside_sdf(*phis[:2])


def test_fraction_inside_linear_interface_area(oracle):
    # For a linear phi the cut region is a polygon Batty's formula measures
    # exactly, except the saddle case (two diagonal corners inside), which a
    # linear function cannot produce.  Compare with a 2000^2 midpoint count.
    rng = np.random.default_rng(6)
    g = (np.arange(2000) + 0.5) / 2000
    X, Y = np.meshgrid(g, g, indexing="xy")
    for _ in range(25):
        a, b, c = rng.uniform(-1, 1, 3)
        phi = lambda x, y: a * x + b * y + c
        got = oracle.fraction_inside(phi(0, 0), phi(1, 0), phi(0, 1), phi(1, 1))
        want = np.mean(phi(X, Y) < 0)
        assert abs(got - want) < 2e-3


def test_fraction_inside_sdf_closed_form(oracle):
    assert oracle.fraction_inside_sdf(-1.0, -2.0) == 1.0
    assert oracle.fraction_inside_sdf(1.0, 2.0) == 0.0
    assert oracle.fraction_inside_sdf(-1.0, 3.0) == 0.25
    assert oracle.fraction_inside_sdf(3.0, -1.0) == 0.25


def test_coarsen_markers_special_cases(oracle):
    # Uniform 4^3 neighbourhoods keep their marker; ties follow argmax3
    # (include/jet/detail/math_utils-inl.h:107-113): the later code wins.
    for code in (0, 1, 2):
        f = np.full((4, 4, 4), code, np.int8)
        assert np.all(oracle.sp_coarsen_markers(f) == code)
    # 2x2x2 coarse grid: coarse cell (0,0,0) sees fine indices {0,0,1,2}^3.
    f = np.zeros((4, 4, 4), np.int8)
    f[:, :, 2:] = 1  # fine i = 2, 3 -> Air; cell 0 sees i in {0,0,1,2}: 3/4 fluid
    c = oracle.sp_coarsen_markers(f)
    assert c[0, 0, 0] == 0 and c[0, 0, 1] == 1
    # Exact fluid/air tie (32/32) -> Air.
    f = np.zeros((8, 8, 8), np.int8)
    f[:, :, 4:] = 1
    c = oracle.sp_coarsen_markers(f)
    # coarse i = 1 sees fine {1,2,3,4}: 3 fluid + 1 air per row; i = 2 sees {3,4,5,6}: 1+3
    assert c[1, 1, 1] == 0 and c[1, 1, 2] == 1
    f = np.zeros((8, 8, 8), np.int8)
    f[:, :, 5:] = 1  # coarse i = 2 sees {3,4,5,6} -> 2 fluid + 2 air: tie -> Air
    assert oracle.sp_coarsen_markers(f)[1, 1, 2] == 1
    f[:, :, 5:] = 2  # fluid/boundary tie -> Boundary
    assert oracle.sp_coarsen_markers(f)[1, 1, 2] == 2


def test_face_fill_coarse_linear_field_exact(oracle):
    # Trilinear sampling of a linear velocity field reproduces it at the
    # coarse face positions (away from clamped extrapolation at the edges).
    n, h = 16, 0.5
    origin = (0.25, -1.0, 2.0)
    a = np.array([0.3, -0.7, 1.1])

    def field(x, y, z):
        return 1.0 + a[0] * x + a[1] * y + a[2] * z

    def faces(nn, hh):
        ix = np.arange(nn + 1)
        ic = np.arange(nn) + 0.5
        zu, yu, xu = np.meshgrid(origin[2] + hh * ic, origin[1] + hh * ic, origin[0] + hh * ix, indexing="ij")
        zv, yv, xv = np.meshgrid(origin[2] + hh * ic, origin[1] + hh * ix, origin[0] + hh * ic, indexing="ij")
        zw, yw, xw = np.meshgrid(origin[2] + hh * ix, origin[1] + hh * ic, origin[0] + hh * ic, indexing="ij")
        return (np.ascontiguousarray(field(xu, yu, zu)), np.ascontiguousarray(field(xv, yv, zv)),
                np.ascontiguousarray(field(xw, yw, zw)))

    u, v, w = faces(n, h)
    cu, cv, cw, ch = oracle.face_fill_coarse(u, v, w, (h, h, h), origin)
    eu, ev, ew = faces(n // 2, 2 * h)
    np.testing.assert_allclose(cu[1:-1, 1:-1, :], eu[1:-1, 1:-1, :], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(cv[1:-1, :, 1:-1], ev[1:-1, :, 1:-1], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(cw[:, 1:-1, 1:-1], ew[:, 1:-1, 1:-1], rtol=1e-13, atol=1e-13)


def test_restrict_float_special_cases(oracle):
    # Centred kernel preserves constants; the staggered direction injects.
    f = np.full((8, 8, 8), 0.75, np.float32)
    np.testing.assert_array_equal(oracle.fr_restrict_float(f, (4, 4, 4)), np.float32(0.75))
    rng = np.random.default_rng(7)
    uwf = rng.uniform(0, 1, (8, 8, 9)).astype(np.float32)
    uwc = oracle.fr_restrict_float(uwf, (4, 4, 5))
    # x is staggered (9 -> 5): weights {0,1,0} pick fine i = 2i; y, z centred.
    k = np.array([0.125, 0.375, 0.375, 0.125])
    j, kk, i = 1, 2, 3
    idx = [2 * j - 1, 2 * j, 2 * j + 1, 2 * j + 2]
    kidx = [2 * kk - 1, 2 * kk, 2 * kk + 1, 2 * kk + 2]
    want = sum(k[a] * k[c] * float(uwf[kidx[c], idx[a], 2 * i]) for a in range(4) for c in range(4))
    assert abs(float(uwc[kk, j, i]) - want) < 1e-6


def test_single_phase_mg_hierarchy(oracle):
    # Coarse levels: markers by majority vote, A re-discretised at 2h; the
    # solid/air pattern is preserved on a scene aligned with the coarse grid.
    n = (16, 16, 16)
    u = np.zeros((16, 16, 17))
    v = np.zeros((16, 17, 16))
    w = np.zeros((17, 16, 16))
    markers, sys = oracle.sp_build_mg(u, v, w, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0),
                                      cc(lambda x, y, z: BIG + 0 * x, n),
                                      cc(lambda x, y, z: y - 8.0, n), 3)
    assert [m.shape for m in markers] == [(16, 16, 16), (8, 8, 8), (4, 4, 4)]
    for l, m in enumerate(markers):
        half = m.shape[1] // 2
        assert np.all(m[:, :half, :] == 0) and np.all(m[:, half:, :] == 1)
        A = sys.A[l]
        inv = 0.25 ** l
        # interior fluid cell below the surface: centre 6/h^2, couplings -1/h^2
        assert A[1, 1, 1, 0] == 6 * inv and A[1, 1, 1, 1] == -inv


def test_fractional_moving_collider_uniform_velocity(oracle):
    # The moving-boundary term (src/jet/grid_fractional_single_phase_pressure_
    # solver3.cpp:241-255): when every face velocity and the collider's velocity
    # are the same constant U, the weighted flux w U through a face and the
    # collider's share (1 - w) U add up to U on every face, so the divergence
    # cancels: b = 0 on every fluid row (center >= eps) whose faces are all
    # inside the domain (a wall face's flux is unweighted).  A sign error or
    # a dropped (1 - w) in the collider term breaks it wherever a cell's
    # opposite faces have different weights, which the sphere makes common.
    n = (20, 18, 16)
    h = (0.5, 0.5, 0.5)
    vs = vertices(lambda x, y, z: np.sqrt((x - 5.2) ** 2 + (y - 4.1) ** 2 + (z - 3.9) ** 2) - 2.3,
                  n, 0.5)
    uw, vw, ww = oracle.fr_face_weights(vs)
    fsdf = cc(lambda x, y, z: y - 7.0, n, 0.5).astype(np.float32)
    U = (0.7, -1.3, 0.45)
    nx, ny, nz = n
    u = np.full((nz, ny, nx + 1), U[0])
    v = np.full((nz, ny + 1, nx), U[1])
    w = np.full((nz + 1, ny, nx), U[2])
    A, b = oracle.fr_build_system(fsdf, uw, vw, ww, u, v, w, h, u.copy(), v.copy(), w.copy())
    fluid = (fsdf < 0) & (A[..., 0] >= 1e-12)
    inner = np.zeros_like(fluid)
    inner[1:-1, 1:-1, 1:-1] = True
    rows = fluid & inner
    # cells whose opposite faces differ in weight: where a wrong term would show
    dw = ((uw[:, :, 1:] != uw[:, :, :-1]) | (vw[:, 1:, :] != vw[:, :-1, :])
          | (ww[1:, :, :] != ww[:-1, :, :]))
    assert np.count_nonzero(rows & dw) >= 20
    assert np.max(np.abs(b[rows])) <= 1e-13 * max(abs(c) for c in U) / h[0]
    # without the collider the same rows see the weight jumps as divergence
    _, b0 = oracle.fr_build_system(fsdf, uw, vw, ww, u, v, w, h)
    assert np.max(np.abs(b0[rows & dw])) > 1e-3


def test_fractional_moving_collider_hand_computed(oracle):
    # One fluid cell (the middle of 3 x 1 x 1, fluid SDF < 0 everywhere) with
    # hand-set face weights; b by hand from the formula of buildSingleSystem:
    # sum over faces of +-(w u + (1 - w) u_b) / h, with the domain-wall faces
    # (outside the x range here only y and z) unweighted in their flux.
    fsdf = np.full((1, 1, 3), -1.0, np.float32)
    uw = np.array([[[1.0, 0.25, 0.75, 1.0]]], np.float32)
    vw = np.array([[[0.5, 0.5, 0.5], [0.2, 0.2, 0.2]]], np.float32)
    ww = np.array([[[0.6, 0.6, 0.6]], [[0.9, 0.9, 0.9]]], np.float32)
    u = np.array([[[0.0, 2.0, -3.0, 0.0]]])
    v = np.array([[[0.5, 0.5, 0.5], [1.5, 1.5, 1.5]]])
    w = np.array([[[-0.25, -0.25, -0.25]], [[0.75, 0.75, 0.75]]])
    bu = np.array([[[0.0, 4.0, 8.0, 0.0]]])
    bv = np.array([[[1.0, 1.0, 1.0], [-2.0, -2.0, -2.0]]])
    bw = np.array([[[3.0, 3.0, 3.0]], [[5.0, 5.0, 5.0]]])
    h = (0.5, 0.25, 2.0)
    A, b = oracle.fr_build_system(fsdf, uw, vw, ww, u, v, w, h, bu, bv, bw)
    ih = [1.0 / x for x in h]
    flux = (0.75 * -3.0 - 0.25 * 2.0) * ih[0]           # interior x faces, weighted
    flux += (1.5 - 0.5) * ih[1] + (0.75 - -0.25) * ih[2]  # wall faces in y and z
    f32 = lambda x: float(np.float32(x))  # the weights are stored as float
    coll = ((1 - 0.75) * 8.0 - (1 - 0.25) * 4.0) * ih[0]
    coll += ((1 - f32(0.2)) * -2.0 - (1 - 0.5) * 1.0) * ih[1]
    coll += ((1 - f32(0.9)) * 5.0 - (1 - f32(0.6)) * 3.0) * ih[2]
    assert b[0, 0, 1] == pytest.approx(flux + coll, rel=1e-12)
    # a wrong sign on the collider term would give flux - coll
    assert abs((flux - coll) - (flux + coll)) > 1.0
