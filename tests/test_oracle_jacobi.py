"""Pins the oracle's weighted-Jacobi V-cycle (or_mg_clean_vcycle_jacobi), the
reference of the GPU's JG_SMOOTH_JACOBI smoother.

- A one-level cycle with one sweep and omega = 1 is one Jet Jacobi relax from
  x = 0, bit for bit (Jet's compiled FdmJacobiSolver3::relax through oracle/_ref
  where it is built, the oracle's bitwise-pinned restatement otherwise).
- Two sweeps from zero with weight omega have a closed form cell by cell:
  x1 = omega b / c, x2 = x1 + omega ((b - s(x1)) / c - x1) with s the
  off-diagonal sum, checked against a dense restatement of A.
- With as many sweeps after the coarse correction as before, the cycle is a
  symmetric positive definite linear map (what PCG needs).
- The preconditioned solve reaches the solution of a dense direct solve.
- A planted mistake (the weight dropped from the sweep) is caught by the
  closed form.
"""
import numpy as np
import pytest

from systems import dense, random_stencil


def hierarchy(shape, levels, seed):
    from oracle.jet_oracle import MgSystem

    rng = np.random.default_rng(seed)
    s = MgSystem()
    nx, ny, nz = shape
    for l in range(levels):
        A = random_stencil(nx >> l, ny >> l, nz >> l, rng)
        s.A.append(A)
        s.b.append(np.zeros(A.shape[:3]))
        s.x.append(np.zeros(A.shape[:3]))
    return s


def params(**kw):
    from oracle.jet_oracle import MgParams

    return MgParams.make(**kw)


def test_one_sweep_is_jets_jacobi_relax(oracle, request):
    s = hierarchy((9, 7, 5), 1, seed=91)
    b = np.random.default_rng(92).uniform(-1, 1, s.x[0].shape)
    z = oracle.mg_clean_vcycle_jacobi(s, params(coarsest_iter=1), 1.0, b)
    assert np.array_equal(z, oracle.jacobi_relax(s.A[0], b, np.zeros_like(b)))
    try:
        ref = request.getfixturevalue("ref")
    except pytest.skip.Exception:
        return
    assert np.array_equal(z, ref.jacobi_relax(s.A[0], b, np.zeros_like(b)))


def _two_sweeps_closed_form(A, b, omega):
    c = A[..., 0]
    M = dense(A)
    off = M - np.diag(np.diag(M))
    x1 = omega * (b / c)
    s = (off @ x1.ravel()).reshape(b.shape)
    return x1 + omega * ((b - s) / c - x1)


@pytest.mark.parametrize("omega", [6.0 / 7.0, 0.5])
def test_two_sweeps_closed_form(oracle, omega):
    s = hierarchy((5, 4, 3), 1, seed=93)
    b = np.random.default_rng(94).uniform(-1, 1, s.x[0].shape)
    z = oracle.mg_clean_vcycle_jacobi(s, params(coarsest_iter=2), omega, b)
    np.testing.assert_allclose(z, _two_sweeps_closed_form(s.A[0], b, omega), rtol=1e-13, atol=1e-15)


def test_planted_unweighted_sweep_is_caught(oracle):
    s = hierarchy((5, 4, 3), 1, seed=93)
    b = np.random.default_rng(94).uniform(-1, 1, s.x[0].shape)
    planted = oracle.mg_clean_vcycle_jacobi(s, params(coarsest_iter=2), 1.0, b)
    want = _two_sweeps_closed_form(s.A[0], b, 6.0 / 7.0)
    assert not np.allclose(planted, want, rtol=1e-6)


@pytest.mark.parametrize("shape,levels", [((16, 16, 16), 3), ((24, 16, 8), 2)])
def test_symmetric_cycle_is_spd_and_linear(oracle, shape, levels):
    s = hierarchy(shape, levels, seed=95)
    p = params(restrict_iter=2, correct_iter=2, coarsest_iter=6, final_iter=2)
    rng = np.random.default_rng(96)
    r1 = rng.uniform(-1, 1, s.x[0].shape)
    r2 = rng.uniform(-1, 1, s.x[0].shape)
    om = 6.0 / 7.0
    z1 = oracle.mg_clean_vcycle_jacobi(s, p, om, r1)
    z2 = oracle.mg_clean_vcycle_jacobi(s, p, om, r2)
    a, b = float(np.sum(r1 * z2)), float(np.sum(r2 * z1))
    assert abs(a - b) <= 1e-12 * max(abs(a), 1.0)
    assert float(np.sum(r1 * z1)) > 0 and float(np.sum(r2 * z2)) > 0
    z12 = oracle.mg_clean_vcycle_jacobi(s, p, om, 2.0 * r1 - 3.0 * r2)
    np.testing.assert_allclose(z12, 2.0 * z1 - 3.0 * z2, rtol=1e-11, atol=1e-12)


def test_preconditioned_solve_matches_direct(oracle):
    s = hierarchy((8, 8, 8), 3, seed=97)
    rng = np.random.default_rng(98)
    s.b[0] = rng.uniform(-1, 1, s.x[0].shape)
    ok, it, rel = oracle.mgpcg_clean_solve_jacobi(
        s, params(max_levels=3, restrict_iter=2, correct_iter=2, coarsest_iter=10, final_iter=2),
        6.0 / 7.0, 100, 1e-12)
    assert ok and rel <= 1e-12 and it < 30
    want = np.linalg.solve(dense(s.A[0]), s.b[0].ravel()).reshape(s.b[0].shape)
    np.testing.assert_allclose(s.x[0], want, rtol=1e-9, atol=1e-11)
