"""Pins the oracle's restatement of GridBackwardEulerDiffusionSolver3's
system build bitwise against Jet's own compiled solver (oracle/_ref: the
system it hands to its linear solver is captured), and the diffusion result
against Jet's unit-test answer (src/tests/unit_tests/
grid_backward_diffusion_solver3_tests.cpp:13-40)."""
import numpy as np
import pytest

BIG = np.finfo(np.float64).max

# Jet's GridBackwardEulerDiffusionSolver3.Solve answer (slice k = 1; the
# k = 0 and k = 2 slices repeat the j = 0 / j = 2 rows of it by symmetry)
KAT_MID = np.array([[0.005291, 0.041270, 0.005291],
                    [0.041270, 0.680423, 0.041270],
                    [0.005291, 0.041270, 0.005291]])
KAT_EDGE = np.array([[0.001058, 0.005291, 0.001058],
                     [0.005291, 0.041270, 0.005291],
                     [0.001058, 0.005291, 0.001058]])


def kat_input():
    f = np.zeros((3, 3, 3))
    f[1, 1, 1] = 1.0
    return f, np.full((3, 3, 3), BIG), np.full((3, 3, 3), -BIG)


def random_case(n, seed):
    nx, ny, nz = n
    rng = np.random.default_rng(seed)
    f = rng.uniform(-1, 1, (nz, ny, nx))
    bs = np.where(rng.random((nz, ny, nx)) < 0.15, -1.0, 1.0)
    fs = np.where(rng.random((nz, ny, nx)) < 0.75, -1.0, 1.0)
    return f, bs, fs


@pytest.mark.parametrize("n,dirichlet,seed", [((9, 7, 6), False, 1), ((9, 7, 6), True, 2),
                                              ((12, 5, 8), True, 3), ((1, 1, 1), True, 4),
                                              ((4, 9, 3), False, 5)])
def test_diffusion_system_matches_jet(oracle, ref, n, dirichlet, seed):
    f, bs, fs = random_case(n, seed)
    h = (0.5, 0.25, 0.4)
    A, b = oracle.diffusion_build(f, bs, fs, 0.3, 0.7, h, dirichlet)
    Aj, bj = ref.diffusion_system(f, bs, fs, 0.3, 0.7, h, dirichlet)
    assert np.array_equal(A, Aj) and np.array_equal(b, bj)


def test_diffusion_kat(oracle):
    f, bs, fs = kat_input()
    A, b = oracle.diffusion_build(f, bs, fs, 1.0, 1.0 / 12.0, (1.0, 1.0, 1.0))
    x, it, res, ok = oracle.cg_solve(A, b, 200, 1e-12)
    np.testing.assert_allclose(x[1], KAT_MID, atol=1e-6)
    np.testing.assert_allclose(x[0], KAT_EDGE, atol=1e-6)
    np.testing.assert_allclose(x[2], KAT_EDGE, atol=1e-6)


def test_jet_diffusion_kat(ref):
    # Jet's own solve (its default ICCG) of the test, through the shim
    f, bs, fs = kat_input()
    x = ref.diffusion_solve(f, bs, fs, 1.0, 1.0 / 12.0, (1.0, 1.0, 1.0))
    np.testing.assert_allclose(x[1], KAT_MID, atol=1e-6)


def test_diffusion_conserves_with_neumann(oracle):
    # All-fluid Neumann diffusion conserves the total: the row sums of
    # I + dt mu L are 1 (closed form of the operator)
    f, bs, fs = random_case((8, 6, 5), 9)
    fs[:] = -1.0
    bs[:] = 1.0
    A, b = oracle.diffusion_build(f, bs, fs, 0.5, 0.2, (1.0, 1.0, 1.0))
    x, it, res, ok = oracle.cg_solve(A, b, 500, 1e-13)
    assert abs(x.sum() - f.sum()) <= 1e-10 * np.abs(f).sum()
