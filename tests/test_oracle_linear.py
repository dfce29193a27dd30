"""Pins the CPU oracle's linear-algebra restatement.

Each group is checked against something other than the oracle's own formula:
Jet's compiled FdmBlas3/CG/MG/MGPCG (oracle/_ref, bitwise or
iteration-for-iteration), Jet's unit-test known answers, closed forms and
dense brute force.
"""
import numpy as np
import pytest

from systems import dense, mg_poisson_hierarchy, poisson_helper, random_stencil

SHAPES = [(1, 1, 1), (2, 1, 1), (3, 4, 5), (7, 1, 3), (16, 16, 16), (33, 17, 9)]


@pytest.mark.parametrize("shape", SHAPES)
def test_blas_bitwise_vs_jet(oracle, ref, shape):
    rng = np.random.default_rng(1)
    nx, ny, nz = shape
    A = random_stencil(nx, ny, nz, rng, spd=False)
    x = rng.uniform(-1, 1, (nz, ny, nx))
    b = rng.uniform(-1, 1, (nz, ny, nx))
    assert np.array_equal(oracle.mvm(A, x), ref.mvm(A, x))
    assert np.array_equal(oracle.residual(A, x, b), ref.residual(A, x, b))
    assert oracle.dot(x, b) == ref.dot(x, b)
    assert oracle.linfnorm(x) == ref.linfnorm(x)
    assert np.array_equal(oracle.axpy(0.37, x, b), ref.axpy(0.37, x, b))


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 3, 1), (4, 3, 2), (5, 5, 5)])
def test_mvm_dense_brute_force(oracle, shape):
    rng = np.random.default_rng(2)
    nx, ny, nz = shape
    A = random_stencil(nx, ny, nz, rng, spd=False)
    x = rng.uniform(-1, 1, (nz, ny, nx))
    b = rng.uniform(-1, 1, (nz, ny, nx))
    M = dense(A)
    np.testing.assert_allclose(oracle.mvm(A, x).ravel(), M @ x.ravel(), rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(oracle.residual(A, x, b).ravel(), b.ravel() - M @ x.ravel(),
                               rtol=1e-13, atol=1e-14)


def test_mvm_integer_stencil_exact(oracle):
    # h = 1 Poisson rows and integer x: every product and sum is exact.
    A, _ = poisson_helper(6, 5, 4)
    rng = np.random.default_rng(3)
    x = rng.integers(-1000, 1000, (4, 5, 6)).astype(np.float64)
    M = dense(A).astype(np.int64)
    assert np.array_equal(oracle.mvm(A, x).ravel(), (M @ x.ravel().astype(np.int64)).astype(np.float64))


def test_cg_2x2_known_answer(oracle):
    # Jet Cg.Solve (src/tests/unit_tests/cg_tests.cpp:16-55): [[4,1],[1,3]] x = (1,2)
    # encoded as a 2x1x1 stencil (center, right).
    A = np.zeros((1, 1, 2, 4))
    A[0, 0, 0] = (4.0, 1.0, 0.0, 0.0)
    A[0, 0, 1] = (3.0, 0.0, 0.0, 0.0)
    b = np.array([[[1.0, 2.0]]])
    x, it, res, ok = oracle.cg_solve(A, b, 0, 0.0)
    assert np.all(x == 0.0) and it == 0 and res == np.sqrt(5.0) and not ok is None
    x, it, res, ok = oracle.cg_solve(A, b, 10, 0.0)
    assert x[0, 0, 0] == pytest.approx(1.0 / 11.0, rel=4 * np.finfo(float).eps)
    assert x[0, 0, 1] == pytest.approx(7.0 / 11.0, rel=4 * np.finfo(float).eps)
    assert res <= np.finfo(float).eps and it <= 2


@pytest.mark.parametrize("n", [3, 8, 13])
def test_cg_closed_form_helper(oracle, n):
    # Helper system solution x = c - j (all-Neumann, compare modulo constant).
    A, b = poisson_helper(n, n, n)
    x, it, res, ok = oracle.cg_solve(A, b, 1000, 1e-12)
    assert ok
    j = np.arange(n)[None, :, None] * np.ones((n, n, n))
    d = x + j
    assert np.ptp(d) < 1e-9


@pytest.mark.parametrize("shape,tol", [((3, 3, 3), 1e-9), ((16, 12, 10), 1e-10), ((24, 24, 24), 1e-8)])
def test_cg_matches_jet_iteration_for_iteration(oracle, ref, shape, tol):
    rng = np.random.default_rng(4)
    nx, ny, nz = shape
    A = random_stencil(nx, ny, nz, rng)
    b = rng.uniform(-1, 1, (nz, ny, nx))
    xo, io, ro, oko = oracle.cg_solve(A, b, 500, tol)
    xr, ir, rr, okr = ref.cg_solve(A, b, 500, tol)
    assert (io, oko) == (ir, okr)
    assert ro == rr
    assert np.array_equal(xo, xr)


def test_cg_max_iterations_returns_false(oracle, ref):
    A, b = poisson_helper(12, 12, 12)
    xo, io, ro, oko = oracle.cg_solve(A, b, 3, 1e-14)
    xr, ir, rr, okr = ref.cg_solve(A, b, 3, 1e-14)
    assert io == ir == 3 and not oko and not okr and ro == rr


@pytest.mark.parametrize("red_black", [False, True])
@pytest.mark.parametrize("sor", [1.0, 1.5])
def test_gauss_seidel_bitwise_vs_jet(oracle, ref, red_black, sor):
    rng = np.random.default_rng(5)
    A = random_stencil(9, 8, 7, rng)
    b = rng.uniform(-1, 1, (7, 8, 9))
    x = rng.uniform(-1, 1, (7, 8, 9))
    assert np.array_equal(oracle.gs_relax(A, b, sor, x, red_black), ref.gs_relax(A, b, sor, x, red_black))
    assert np.array_equal(oracle.jacobi_relax(A, b, x), ref.jacobi_relax(A, b, x))


@pytest.mark.parametrize("red_black", [False, True])
def test_relax_monotone_residual(oracle, red_black):
    # Jet FdmGaussSeidelSolver3.Relax / RelaxRedBlack
    # (src/tests/unit_tests/fdm_gauss_seidel_solver3_tests.cpp:43-84).
    A, b = poisson_helper(32, 32, 32)
    x = np.zeros_like(b)
    norm0 = oracle.l2norm(oracle.residual(A, x, b))
    for it in range(20):
        x = oracle.gs_relax(A, b, 1.0, x, red_black)
        norm = oracle.l2norm(oracle.residual(A, x, b))
        if it > 0 or not red_black:
            assert norm < norm0
        norm0 = norm


def test_gauss_seidel_is_lexicographic(oracle):
    # One lexicographic sweep on a 1-D chain equals forward substitution of
    # (D + L) x = b: a closed form independent of the relax loop.
    rng = np.random.default_rng(6)
    A = random_stencil(6, 1, 1, rng)
    b = rng.uniform(-1, 1, (1, 1, 6))
    M = dense(A)
    x_expected = np.linalg.solve(np.tril(M), b.ravel())
    np.testing.assert_allclose(oracle.gs_relax(A, b, 1.0, np.zeros_like(b)).ravel(), x_expected,
                               rtol=1e-13)


@pytest.mark.parametrize("shape", [(2, 2, 2), (4, 6, 8), (16, 16, 16)])
def test_restrict_correct_bitwise_vs_jet(oracle, ref, shape):
    rng = np.random.default_rng(7)
    nx, ny, nz = shape
    f = rng.uniform(-1, 1, (nz, ny, nx))
    c = rng.uniform(-1, 1, (nz // 2, ny // 2, nx // 2))
    assert np.array_equal(oracle.restrict(f), ref.restrict(f))
    assert np.array_equal(oracle.correct(c, f), ref.correct(c, f))


def test_restrict_closed_forms(oracle):
    # Weights sum to one: constants restrict to themselves everywhere
    # (including clamped edges); linear fields are exact away from the
    # clamped faces (the stencil is centred on fine index 2i+0.5).
    n = 16
    f = np.full((n, n, n), 3.25)
    np.testing.assert_allclose(oracle.restrict(f), 3.25, rtol=1e-15)
    i = np.arange(n, dtype=float)
    lin = (2.0 * i[None, None, :] - 0.5 * i[None, :, None] + 0.25 * i[:, None, None])
    c = oracle.restrict(np.ascontiguousarray(lin))
    ci = 2 * np.arange(n // 2) + 0.5
    expect = 2.0 * ci[None, None, :] - 0.5 * ci[None, :, None] + 0.25 * ci[:, None, None]
    np.testing.assert_allclose(c[1:-1, 1:-1, 1:-1], expect[1:-1, 1:-1, 1:-1], rtol=1e-14, atol=1e-13)


def test_correct_closed_forms(oracle):
    # Prolongation weights sum to one: a constant coarse correction adds the
    # constant; a linear coarse field prolongs exactly in the interior.
    n = 8
    f = np.zeros((2 * n, 2 * n, 2 * n))
    out = oracle.correct(np.full((n, n, n), -1.5), f)
    np.testing.assert_allclose(out, -1.5, rtol=1e-15)
    ic = np.arange(n, dtype=float)
    lin = ic[None, None, :] + 2 * ic[None, :, None] - ic[:, None, None]
    out = oracle.correct(np.ascontiguousarray(lin), f)
    fi = (np.arange(2 * n) - 0.5) / 2.0  # fine centre in coarse index units
    expect = fi[None, None, :] + 2 * fi[None, :, None] - fi[:, None, None]
    np.testing.assert_allclose(out[1:-1, 1:-1, 1:-1], expect[1:-1, 1:-1, 1:-1], rtol=1e-14, atol=1e-13)


def test_correct_k_clamp_uses_ny_like_jet(oracle, ref):
    # Non-cubic grid where n.y != n.z exposes Jet's k-clamp quirk
    # (src/jet/fdm_mg_linear_system3.cpp:159): the oracle must reproduce it.
    rng = np.random.default_rng(8)
    f = rng.uniform(-1, 1, (8, 4, 6))
    c = rng.uniform(-1, 1, (4, 2, 3))
    assert np.array_equal(oracle.correct(c, f), ref.correct(c, f))


@pytest.mark.parametrize("shape,maxl", [((128, 128, 128), 5), ((48, 48, 48), 9), ((96, 96, 96), 9),
                                        ((64, 32, 16), 9), ((5, 6, 7), 3), ((8, 8, 8), 0)])
def test_level_sizing_vs_jet(oracle, ref, shape, maxl):
    assert oracle.levels_with_finest(shape, maxl) == ref.levels_with_finest(shape, maxl)


def test_mg_solve_matches_jet_and_reduces_residual(oracle, ref):
    # Jet FdmMgSolver3.Solve at 4 levels from 4^3 (32^3 finest) to keep it
    # quick (the reference test uses 6 levels).
    from oracle.jet_oracle import MgParams

    p = MgParams.make(max_levels=4, restrict_iter=5, correct_iter=5, coarsest_iter=20,
                      final_iter=20, max_tol=1e-9)
    so = mg_poisson_hierarchy(4, 4)
    sr = mg_poisson_hierarchy(4, 4)
    norm0 = oracle.l2norm(oracle.residual(so.A[0], so.x[0], so.b[0]))
    oko, normo = oracle.mg_solve(so, p)
    okr = ref.mg_solve(sr, p)
    assert oko == okr
    for l in range(4):
        assert np.array_equal(so.x[l], sr.x[l])
    assert oracle.l2norm(oracle.residual(so.A[0], so.x[0], so.b[0])) < norm0


@pytest.mark.parametrize("red_black", [False, True])
def test_mgpcg_matches_jet(oracle, ref, red_black):
    # Jet FdmMgpcgSolver3.Solve (src/tests/unit_tests/fdm_mgpcg_solver3_tests.cpp:13-60):
    # MGPCG(50, 4, 5, 5, 10, 10, 1e-4, 1.5) on 4 levels from 4^3 must converge.
    from oracle.jet_oracle import MgParams

    p = MgParams.make(max_levels=4, restrict_iter=5, correct_iter=5, coarsest_iter=10,
                      final_iter=10, max_tol=1e-4, sor=1.5, red_black=red_black)
    so = mg_poisson_hierarchy(4, 4)
    sr = mg_poisson_hierarchy(4, 4)
    oko, io, ro = oracle.mgpcg_solve(so, p, 50)
    okr, ir, rr = ref.mgpcg_solve(sr, p, 50)
    assert oko and okr
    assert io == ir and ro == rr
    assert np.array_equal(so.x[0], sr.x[0])


def test_mgpcg_solution_matches_cg(oracle):
    # Different algorithms, same fine system: Jet-faithful MGPCG, the
    # clean-mode reference solve and CG agree.  A Dirichlet row shift on every
    # level makes the system non-singular (on the all-Neumann helper Jet's
    # warm-started SOR preconditioner stalls near 1e-4, which oracle/_ref
    # reproduces).
    from oracle.jet_oracle import MgParams

    def system():
        s = mg_poisson_hierarchy(4, 3)
        for l, A in enumerate(s.A):
            A[:, -1, :, 0] += 0.25 ** l
        return s

    p = MgParams.make(max_levels=3, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                      final_iter=2, max_tol=1e-10)
    s = system()
    xc, it, res, ok = oracle.cg_solve(s.A[0], s.b[0], 2000, 1e-12)
    assert ok and res <= 1e-12
    oracle.mgpcg_solve(s, p, 200)
    assert np.abs(s.x[0] - xc).max() < 1e-6 * np.abs(xc).max()
    s = system()
    ok, it, rel = oracle.mgpcg_clean_solve(s, p, 100, 1e-12)
    assert ok and rel <= 1e-12 and it < 30
    assert np.abs(s.x[0] - xc).max() < 1e-11 * np.abs(xc).max()
