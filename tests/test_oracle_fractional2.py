"""Pins the oracle's restatement of GridFractionalSinglePhasePressureSolver2
(src/jet/grid_fractional_single_phase_pressure_solver2.cpp): face weights
(:414-490), float restriction (:29-93), the variational system (:95-203) and
the pressure gradient (:565-600).

- Bitwise against Jet's own solver compiled from its sources (oracle/_ref):
  A on every level, b and the projected velocity of a given pressure, for a
  circular collider (resting and moving) under a free surface.
- Jet's three unit-test answers (src/tests/unit_tests/
  grid_fractional_single_phase_pressure_solver2_tests.cpp:17,63,109) through
  the oracle's build, solve and gradient.
- A uniform flow U with the collider moving at U is divergence free: b = 0 on
  every fluid row (the moving-boundary term cancels the weighted faces).
"""
import numpy as np
import pytest

KMAXD = np.finfo(np.float64).max


def scene2(nx, ny, seed, moving=True, h=None):
    """Circle collider (centre (0.5 nx, 0.35 ny) h, radius 0.2 ny h) below a
    liquid plane at 0.6 ny h; velocities U(-1, 1); a rigid collider motion
    (translation plus rotation about the centre) sampled at the faces."""
    rng = np.random.default_rng(seed)
    h = h or 1.0 / nx
    c = (0.5 * nx * h, 0.35 * ny * h)
    rad = 0.2 * ny * h
    X, Y = np.meshgrid((np.arange(-1, nx + 1) + 0.5) * h, (np.arange(-1, ny + 1) + 0.5) * h)
    bsdf = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2) - rad
    _, FY = np.meshgrid((np.arange(nx) + 0.5) * h, (np.arange(ny) + 0.5) * h)
    fsdf = FY - 0.6 * ny * h
    u = rng.uniform(-1, 1, (ny, nx + 1))
    v = rng.uniform(-1, 1, (ny + 1, nx))
    bu = bv = None
    if moving:
        _, UY = np.meshgrid(np.arange(nx + 1) * h, (np.arange(ny) + 0.5) * h)
        VX, _ = np.meshgrid((np.arange(nx) + 0.5) * h, np.arange(ny + 1) * h)
        bu = np.ascontiguousarray(0.3 - 0.7 * (UY - c[1]))
        bv = np.ascontiguousarray(-0.2 + 0.7 * (VX - c[0]))
    return dict(u=u, v=v, h=(h, h), bsdf=bsdf, fsdf=fsdf, bu=bu, bv=bv)


@pytest.mark.parametrize("nx,ny,levels,moving", [(16, 16, 3, True), (33, 17, 1, True),
                                                 (64, 48, 4, False), (40, 24, 3, True),
                                                 (8, 8, 4, True)])
def test_bitwise_vs_jet(oracle, ref, nx, ny, levels, moving):
    s = scene2(nx, ny, 7, moving)
    (fs, UW, VW), sys_ = oracle.fr_build_mg2(s["u"], s["v"], s["h"], s["bsdf"], s["fsdf"], levels,
                                             s["bu"], s["bv"])
    p = np.random.default_rng(3).uniform(-1, 1, (ny, nx))
    A_jet, b_jet, uo_jet, vo_jet = ref.fr_solve2(s["u"], s["v"], s["h"], (0.0, 0.0), s["bsdf"],
                                                 s["fsdf"], levels, p, s["bu"], s["bv"])
    assert len(A_jet) == len(sys_.A)
    for A, Aj in zip(sys_.A, A_jet):
        assert np.array_equal(A, Aj)
    assert np.array_equal(sys_.b[0], b_jet)
    uo, vo = oracle.fr_apply_gradient2(fs[0], UW[0], VW[0], p, s["u"], s["v"], s["h"])
    assert np.array_equal(uo, uo_jet) and np.array_equal(vo, vo_jet)


def _free_surface(n, surface):
    """Jet's SolveFreeSurface inputs: u = 0, v = 1 on the interior v faces,
    fluid below y = surface, no collider (boundary SDF = kMaxD)."""
    u = np.zeros((n, n + 1))
    v = np.ones((n + 1, n))
    v[0, :] = 0.0
    v[n, :] = 0.0
    _, FY = np.meshgrid(np.arange(n) + 0.5, np.arange(n) + 0.5)
    return u, v, np.full((n + 2, n + 2), KMAXD), FY - surface


def _solve(oracle, sys_, levels):
    from oracle.jet_oracle import MgParams

    ok, _, rel = oracle.mgpcg_clean_solve(sys_, MgParams.make(max_levels=levels, restrict_iter=2,
                                                              correct_iter=2, coarsest_iter=20,
                                                              final_iter=2), 200, 1e-12)
    assert ok and rel <= 1e-12
    return sys_.x[0][0]


@pytest.mark.parametrize("compressed", [False, True])
def test_jet_free_surface(oracle, compressed):
    # SolveFreeSurface / SolveFreeSurfaceCompressed (:17, :63): both solve the
    # same system (the compressed form keeps the fluid rows only); p = 1.5, 0.5,
    # 0 by row and the projected velocity is zero
    u, v, bsdf, fsdf = _free_surface(3, 2.0)
    (fs, UW, VW), sys_ = oracle.fr_build_mg2(u, v, (1.0, 1.0), bsdf, fsdf, 1)
    if compressed:
        fluid = fs[0].ravel() < 0
        A = np.zeros((fluid.sum(), fluid.sum()))
        rows = sys_.A[0][0].reshape(-1, 4)
        idx = -np.ones(9, int)
        idx[fluid] = np.arange(fluid.sum())
        for c in np.nonzero(fluid)[0]:
            A[idx[c], idx[c]] = rows[c, 0]
            for off, k in ((1, 1), (3, 2)):
                if c + off < 9 and fluid[c + off] and rows[c, k] != 0.0:
                    A[idx[c], idx[c + off]] = A[idx[c + off], idx[c]] = rows[c, k]
        x = np.zeros(9)
        x[fluid] = np.linalg.solve(A, sys_.b[0].ravel()[fluid])
        p = x.reshape(3, 3)
    else:
        p = _solve(oracle, sys_, 1)
    for i in range(3):
        assert abs(p[0, i] - 1.5) <= 1e-6 and abs(p[1, i] - 0.5) <= 1e-6 and abs(p[2, i]) <= 1e-6
    uo, vo = oracle.fr_apply_gradient2(fs[0], UW[0], VW[0], p, u, v, (1.0, 1.0))
    assert np.max(np.abs(uo)) <= 1e-6 and np.max(np.abs(vo)) <= 1e-6


def test_jet_free_surface_mg(oracle):
    # SolveFreeSurfaceMg (:109): 32 x 32, surface at y = 16, 5 levels; p = 15.5 - j
    # below the surface and 0 above (within 0.1), u and the fluid's v ~ 0 (0.002)
    u, v, bsdf, fsdf = _free_surface(32, 16.0)
    (fs, UW, VW), sys_ = oracle.fr_build_mg2(u, v, (1.0, 1.0), bsdf, fsdf, 5)
    p = _solve(oracle, sys_, 5)
    for j in range(32):
        want = 15.5 - j if j < 16 else 0.0
        assert abs(p[j, 16] - want) <= 0.1
    uo, vo = oracle.fr_apply_gradient2(fs[0], UW[0], VW[0], p, u, v, (1.0, 1.0))
    assert np.max(np.abs(uo)) <= 0.002 and np.max(np.abs(vo[:16])) <= 0.002


def test_uniform_flow_with_the_collider_is_divergence_free(oracle):
    s = scene2(24, 20, 9, moving=False)
    U = (0.37, -0.81)
    u = np.full_like(s["u"], U[0])
    v = np.full_like(s["v"], U[1])
    (fs, _, _), sys_ = oracle.fr_build_mg2(u, v, s["h"], s["bsdf"], s["fsdf"], 1,
                                           np.full_like(u, U[0]), np.full_like(v, U[1]))
    # fluid rows away from the domain walls (a wall face is unweighted and
    # carries the flow out of the domain)
    fluid = (fs[0] < 0) & (sys_.A[0][0][..., 0] != 1.0)
    fluid[:, 0] = fluid[:, -1] = False
    fluid[0, :] = fluid[-1, :] = False
    b = sys_.b[0][0]
    assert fluid.sum() > 50
    assert np.max(np.abs(b[fluid])) <= 1e-12 * 24 / s["h"][0]
