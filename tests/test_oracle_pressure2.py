"""Pins the oracle's 2-D single-phase pressure build (GridSinglePhasePressureSolver2:
markers, coarse markers, system, gradient) by Jet's own 2-D unit-test known
answers (src/tests/unit_tests/grid_single_phase_pressure_solver2_tests.cpp),
the projection identity div(u') = b - A p, and by agreement with the 3-D
build on one plane (w = 0), which is an independent restatement.  Jet's 2-D
pressure solver is not compiled here (it pulls in the boundary-condition
solvers and colliders)."""
import numpy as np
import pytest

BIG = np.finfo(np.float64).max


def cc2(fn, nx, ny, h=1.0):
    y, x = np.meshgrid(h * (np.arange(ny) + 0.5), h * (np.arange(nx) + 0.5), indexing="ij")
    return np.ascontiguousarray(fn(x, y), dtype=np.float64)


def kat_velocity2():
    """Jet's 2-D test field on 3 x 3 cells: v = 1 on the interior y-faces."""
    u = np.zeros((3, 4))
    v = np.zeros((4, 3))
    v[1:3, :] = 1.0
    return u, v


def solve2(oracle, u, v, bsdf, fsdf, h=(1.0, 1.0)):
    m = oracle.sp_markers(bsdf, fsdf)
    A, b = oracle.sp_build_system2(m, u, v, h)
    p, it, res, ok = oracle.cg_solve(A, b, 200, 1e-12)
    uo, vo = oracle.sp_apply_gradient2(m, p[0], u, v, h)
    return m, A, b, p[0], uo, vo


def test_kat2_single_phase(oracle):
    # GridSinglePhasePressureSolver2.SolveSinglePhase (:16-58)
    u, v = kat_velocity2()
    _, _, _, p, uo, vo = solve2(oracle, u, v, cc2(lambda x, y: BIG + 0 * x, 3, 3),
                                cc2(lambda x, y: -BIG + 0 * x, 3, 3))
    assert np.abs(uo).max() < 1e-6 and np.abs(vo).max() < 1e-6
    np.testing.assert_allclose(p[1:, :] - p[:-1, :], -1.0, atol=1e-6)


def test_kat2_single_phase_with_boundary(oracle):
    # ...SolveSinglePhaseWithBoundary (:104-152): wall on the right-most column
    u, v = kat_velocity2()
    _, _, _, p, uo, vo = solve2(oracle, u, v, cc2(lambda x, y: -x + 2.0, 3, 3),
                                cc2(lambda x, y: -BIG + 0 * x, 3, 3))
    assert np.abs(uo).max() < 1e-6
    expect = np.zeros_like(vo)
    expect[1:3, 2] = 1.0
    np.testing.assert_allclose(vo, expect, atol=1e-6)
    np.testing.assert_allclose(p[1:, :2] - p[:-1, :2], -1.0, atol=1e-6)


def test_kat2_free_surface(oracle):
    # ...SolveFreeSurface (:154-199): fluidSdf = y - 2 -> p = 2 - j
    u, v = kat_velocity2()
    _, _, _, p, uo, vo = solve2(oracle, u, v, cc2(lambda x, y: BIG + 0 * x, 3, 3),
                                cc2(lambda x, y: y - 2.0, 3, 3))
    assert np.abs(uo).max() < 1e-6 and np.abs(vo).max() < 1e-6
    for j in range(2):
        np.testing.assert_allclose(p[j, :], 2.0 - j, atol=1e-6)


def random_scene2(nx, ny, rng, h):
    c = rng.uniform(0.3, 0.7, 2) * (nx * h, ny * h)
    r = 0.2 * min(nx, ny) * h
    bsdf = cc2(lambda x, y: np.hypot(x - c[0], y - c[1]) - r, nx, ny, h)
    fsdf = cc2(lambda x, y: y - 0.7 * ny * h + 0.1 * np.sin(x / h), nx, ny, h)
    u = rng.uniform(-1, 1, (ny, nx + 1))
    v = rng.uniform(-1, 1, (ny + 1, nx))
    return bsdf, fsdf, u, v


@pytest.mark.parametrize("seed,n,h", [(0, (12, 10), 1.0), (1, (17, 9), 0.5), (2, (32, 32), 0.25)])
def test_projection_identity2(oracle, seed, n, h):
    # div(u') = b - A p on every fluid cell with no Air neighbour on its -x / -y
    # side, for any p (Jet updates only +side faces of fluid cells).
    rng = np.random.default_rng(seed)
    nx, ny = n
    bsdf, fsdf, u, v = random_scene2(nx, ny, rng, h)
    m = oracle.sp_markers(bsdf, fsdf)
    assert (m == 0).sum() > 0.2 * m.size and (m == 1).any() and (m == 2).any()
    A, b = oracle.sp_build_system2(m, u, v, (h, h))
    p = rng.uniform(-1, 1, m.shape)
    p[m != 0] = 0.0
    uo, vo = oracle.sp_apply_gradient2(m, p, u, v, (h, h))
    lhs = (uo[:, 1:] - uo[:, :-1]) / h + (vo[1:, :] - vo[:-1, :]) / h
    rhs = (b - oracle.mvm(A, p[None]))[0]
    air = m == 1
    ok = m == 0
    ok[:, 1:] &= ~air[:, :-1]
    ok[1:, :] &= ~air[:-1, :]
    np.testing.assert_allclose(lhs[ok], rhs[ok], rtol=1e-12, atol=1e-10 / h)
    assert np.all(A[0][m != 0] == np.array([1.0, 0.0, 0.0, 0.0])) and np.all(b[0][m != 0] == 0.0)
    assert np.all(A[..., 3] == 0.0)


@pytest.mark.parametrize("seed,n", [(3, (12, 10)), (4, (16, 16)), (5, (9, 7))])
def test_build2_equals_3d_build_on_one_plane(oracle, seed, n):
    # The 3-D build of the same cells as one plane with w = 0 is an
    # independent restatement: equal A and b (b up to the sign of zero).
    rng = np.random.default_rng(seed)
    nx, ny = n
    bsdf, fsdf, u, v = random_scene2(nx, ny, rng, 1.0)
    m = oracle.sp_markers(bsdf, fsdf)
    A2, b2 = oracle.sp_build_system2(m, u, v, (1.0, 1.0))
    w = np.zeros((2, ny, nx))
    A3, b3 = oracle.sp_build_system(m[None].copy(), u[None].copy(), v[None].copy(), w,
                                    (1.0, 1.0, 1.0))
    assert np.array_equal(A2, A3) and np.array_equal(b2, b3)
    p = rng.uniform(-1, 1, (1, ny, nx))
    uo2, vo2 = oracle.sp_apply_gradient2(m, p[0], u, v, (1.0, 1.0))
    uo3, vo3, _ = oracle.sp_apply_gradient(m[None].copy(), p, u[None].copy(), v[None].copy(), w,
                                           (1.0, 1.0, 1.0))
    assert np.array_equal(uo2, uo3[0]) and np.array_equal(vo2, vo3[0])


def test_coarsen_markers2(oracle):
    # uniform neighbourhoods keep their code; a 2-D vote over 4 x 4 equals the
    # 3-D vote over the same cells extruded to two planes (every count x 4)
    for code in (0, 1, 2):
        assert np.all(oracle.sp_coarsen_markers2(np.full((4, 4), code, np.int8)) == code)
    rng = np.random.default_rng(6)
    for shape in ((8, 8), (6, 10), (16, 4)):
        f = rng.integers(0, 3, shape).astype(np.int8)
        want = oracle.sp_coarsen_markers(np.ascontiguousarray(np.stack([f, f])))[0]
        assert np.array_equal(oracle.sp_coarsen_markers2(f), want)
    # fluid / air tie (8 / 8) -> Air, fluid / boundary tie -> Boundary (argmax3)
    f = np.zeros((8, 8), np.int8)
    f[:, 5:] = 1
    assert oracle.sp_coarsen_markers2(f)[1, 2] == 1
    f[:, 5:] = 2
    assert oracle.sp_coarsen_markers2(f)[1, 2] == 2


def test_sp_build_mg2_hierarchy(oracle):
    rng = np.random.default_rng(7)
    bsdf, fsdf, u, v = random_scene2(64, 32, rng, 1.0 / 64)
    markers, sys = oracle.sp_build_mg2(u, v, (1.0 / 64, 1.0 / 64), bsdf, fsdf, 4)
    assert [m.shape for m in markers] == [(32, 64), (16, 32), (8, 16), (4, 8)]
    assert [a.shape for a in sys.A] == [(1, 32, 64, 4), (1, 16, 32, 4), (1, 8, 16, 4), (1, 4, 8, 4)]
    # coarse couplings scale by 1/4 per level: a coarse fluid cell with fluid
    # on all four sides has centre 4 / (2^l h)^2
    for l, (m, A) in enumerate(zip(markers, sys.A)):
        ih2 = (64.0 / 2 ** l) ** 2
        inner = np.zeros_like(m, bool)
        inner[1:-1, 1:-1] = ((m[1:-1, 1:-1] == 0) & (m[1:-1, :-2] == 0) & (m[1:-1, 2:] == 0)
                             & (m[:-2, 1:-1] == 0) & (m[2:, 1:-1] == 0))
        assert np.all(A[0][inner][:, 0] == 4 * ih2)


def decompress2(m, x):
    p = np.zeros(m.shape)
    p[m == 0] = x
    return p


@pytest.mark.parametrize("case", ["single", "free"])
def test_kat2_compressed(oracle, case):
    # ...SolveSinglePhaseCompressed (:60-102) and SolveFreeSurfaceCompressed
    # (:201-246): the compressed build, Jet's CG on it, decompressSolution
    u, v = kat_velocity2()
    bs = cc2(lambda x, y: BIG + 0 * x, 3, 3)
    fs = cc2(lambda x, y: y - 2.0, 3, 3) if case == "free" else cc2(lambda x, y: -BIG + 0 * x, 3, 3)
    m = oracle.sp_markers(bs, fs)
    val, rp, ci, b = oracle.sp_build_compressed2(m, u, v, (1.0, 1.0))
    x, it, res, ok = oracle.csr_cg_solve(val, rp, ci, b, 200, 1e-12)
    p = decompress2(m, x)
    uo, vo = oracle.sp_apply_gradient2(m, p, u, v, (1.0, 1.0))
    assert np.abs(uo).max() < 1e-6 and np.abs(vo).max() < 1e-6
    if case == "free":
        for j in range(2):
            np.testing.assert_allclose(p[j, :], 2.0 - j, atol=1e-6)
    else:
        np.testing.assert_allclose(p[1:, :] - p[:-1, :], -1.0, atol=1e-6)


@pytest.mark.parametrize("seed,n", [(8, (12, 10)), (9, (17, 9))])
def test_compressed2_equals_3d_and_flat(oracle, seed, n):
    # the 3-D compressed restatement on one plane (w = 0) gives the same
    # arrays, and the compressed rows are the flat rows on the fluid cells
    rng = np.random.default_rng(seed)
    nx, ny = n
    bsdf, fsdf, u, v = random_scene2(nx, ny, rng, 0.5)
    m = oracle.sp_markers(bsdf, fsdf)
    val, rp, ci, b = oracle.sp_build_compressed2(m, u, v, (0.5, 0.5))
    w = np.zeros((2, ny, nx))
    v3, r3, c3, b3 = oracle.sp_build_compressed(m[None].copy(), u[None].copy(), v[None].copy(), w,
                                                (0.5, 0.5, 0.5))
    assert np.array_equal(val, v3) and np.array_equal(rp, r3) and np.array_equal(ci, c3)
    assert np.array_equal(b, b3)
    A, bf = oracle.sp_build_system2(m, u, v, (0.5, 0.5))
    fl = np.flatnonzero(m.ravel() == 0)
    assert np.array_equal(b, bf.ravel()[fl])
    rows = len(fl)
    D = np.zeros((rows, rows))
    for r in range(rows):
        D[r, ci[rp[r]:rp[r + 1]].astype(int)] = val[rp[r]:rp[r + 1]]
    pos = {c: k for k, c in enumerate(fl)}
    for k, c in enumerate(fl):
        assert D[k, k] == A.reshape(-1, 4)[c, 0]
        if c + 1 in pos and (c + 1) % nx:
            assert D[k, pos[c + 1]] == A.reshape(-1, 4)[c, 1]
        if c + nx in pos:
            assert D[k, pos[c + nx]] == A.reshape(-1, 4)[c, 2]
    assert np.array_equal(D, D.T)
