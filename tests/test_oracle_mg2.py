"""The oracle's 2-D multigrid (FdmMgUtils2 transfers, FdmMgpcgSolver2)
pinned against Jet's compiled solver, Jet's unit-test answers and closed
forms.  A hierarchy of one-plane levels with more than one level is a 2-D
hierarchy; its relax, residual and products are the 3-D ones with the z
terms absent."""
import numpy as np
import pytest

from systems import dense, mg_poisson_hierarchy2, poisson_helper


@pytest.mark.parametrize("shape", [(2, 2), (4, 6), (32, 16), (6, 4), (2, 8)])
def test_restrict2_correct2_bitwise_vs_jet(oracle, ref, shape):
    rng = np.random.default_rng(11)
    nx, ny = shape
    f = rng.uniform(-1, 1, (ny, nx))
    c = rng.uniform(-1, 1, (ny // 2, nx // 2))
    assert np.array_equal(oracle.restrict2(f), ref.restrict2(f))
    assert np.array_equal(oracle.correct2(c, f), ref.correct2(c, f))


def test_restrict2_correct2_closed_forms(oracle):
    # Weights sum to one: constants map to themselves (clamped edges too);
    # linear fields are exact away from the clamped edges.
    n = 16
    np.testing.assert_allclose(oracle.restrict2(np.full((n, n), -2.75)), -2.75, rtol=1e-15)
    i = np.arange(n, dtype=float)
    lin = 3.0 * i[None, :] - 0.5 * i[:, None]
    ci = 2 * np.arange(n // 2) + 0.5
    expect = 3.0 * ci[None, :] - 0.5 * ci[:, None]
    np.testing.assert_allclose(oracle.restrict2(np.ascontiguousarray(lin))[1:-1, 1:-1],
                               expect[1:-1, 1:-1], rtol=1e-14, atol=1e-13)
    m = 8
    f = np.zeros((2 * m, 2 * m))
    np.testing.assert_allclose(oracle.correct2(np.full((m, m), 1.25), f), 1.25, rtol=1e-15)
    ic = np.arange(m, dtype=float)
    clin = ic[None, :] - 2 * ic[:, None]
    fi = (np.arange(2 * m) - 0.5) / 2.0
    expect = fi[None, :] - 2 * fi[:, None]
    np.testing.assert_allclose(oracle.correct2(np.ascontiguousarray(clin), f)[1:-1, 1:-1],
                               expect[1:-1, 1:-1], rtol=1e-14, atol=1e-13)


def test_restrict2_is_the_3d_restriction_of_a_z_constant_field(oracle):
    # A field that does not vary over two planes restricts in 3-D (z taps
    # {0, 0, 1, 1}, weights summing to one) to the 2-D restriction.
    rng = np.random.default_rng(12)
    f = rng.uniform(-1, 1, (12, 10))
    f3 = np.ascontiguousarray(np.stack([f, f]))
    np.testing.assert_allclose(oracle.restrict(f3)[0], oracle.restrict2(f), rtol=1e-14, atol=1e-15)
    c = rng.uniform(-1, 1, (6, 5))
    c3 = np.ascontiguousarray(c[None])
    out3 = oracle.correct(c3, np.zeros((2, 12, 10)))
    np.testing.assert_allclose(out3[0], oracle.correct2(c, np.zeros((12, 10))), rtol=1e-14,
                               atol=1e-15)


def test_level_sizing2_jet_kat(oracle, ref):
    # FdmMgUtils2.ResizeArrayWithFinest (src/tests/unit_tests/
    # fdm_mg_linear_system2_tests.cpp.cpp:13-32)
    assert oracle.levels_with_finest2((100, 200), 4) == [(100, 200), (50, 100), (25, 50)]
    assert oracle.levels_with_finest2((32, 16), 6) == [(32, 16), (16, 8), (8, 4), (4, 2), (2, 1)]
    assert len(oracle.levels_with_finest2((16, 16), 6)) == 5
    for shape, maxl in (((100, 200), 4), ((32, 16), 6), ((16, 16), 6), ((7, 8), 3), ((64, 64), 0)):
        assert oracle.levels_with_finest2(shape, maxl) == ref.levels_with_finest2(shape, maxl)


@pytest.mark.parametrize("red_black", [True, False])
def test_mgpcg2_jet_kat_and_bitwise_vs_jet(oracle, ref, red_black):
    # FdmMgpcgSolver2.Solve (src/tests/unit_tests/fdm_mgpcg_solver2_tests.cpp:
    # 13-49): MGPCG(200, 6, 5, 5, 10, 10, 1e-4) on 6 levels from 4 x 4 must
    # converge; the oracle follows Jet's solver iteration for iteration.
    from oracle.jet_oracle import MgParams

    p = MgParams.make(max_levels=6, restrict_iter=5, correct_iter=5, coarsest_iter=10,
                      final_iter=10, max_tol=1e-4, red_black=red_black)
    so = mg_poisson_hierarchy2(4, 6)
    sr = mg_poisson_hierarchy2(4, 6)
    oko, io, ro = oracle.mgpcg_solve(so, p, 200)
    okr, ir, rr = ref.mgpcg_solve2(sr, p, 200)
    assert oko and okr
    assert io == ir
    np.testing.assert_allclose(ro, rr, rtol=1e-12)
    np.testing.assert_allclose(so.x[0], sr.x[0], rtol=0, atol=1e-12 * np.abs(sr.x[0]).max())


def test_mgpcg2_clean_solve_matches_dense_solve(oracle):
    # The clean-mode FP64 solve on a 2-D hierarchy reaches the dense
    # solution of a non-singular system (Dirichlet shift on the top row of
    # every level).
    from oracle.jet_oracle import MgParams

    s = mg_poisson_hierarchy2(2, 4)
    for l, A in enumerate(s.A):
        A[:, -1, :, 0] += 0.25 ** l
    p = MgParams.make(max_levels=4, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                      final_iter=2, max_tol=1e-12)
    ok, it, rel = oracle.mgpcg_clean_solve(s, p, 100, 1e-12)
    assert ok and rel <= 1e-11 and it < 40
    x = np.linalg.solve(dense(s.A[0]), s.b[0].ravel())
    np.testing.assert_allclose(s.x[0].ravel(), x, rtol=0, atol=1e-9 * np.abs(x).max())
