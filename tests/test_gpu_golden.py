"""GPU kernels through the C ABI against the committed vectors of Jet's own
compiled code (tests/golden/jet_linear.npz, see tests/golden/make_golden.py).

FdmBlas3::mvm / residual, one relaxRedBlack sweep, one Jacobi step and the
multigrid restriction / correction evaluate Jet's expressions in Jet's
operand order with contraction off: the FP64 results must be bit-identical
to Jet's.  FdmCgSolver3 differs from Jet only in the summation order of its
dot products: iterations within 2 of Jet's, x within 1e-9 relative.
"""
import os

import numpy as np
import pytest

from jgtest import dev, empty, host, jetgpu, rel_err, torch_cuda

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "jet_linear.npz")


@pytest.fixture(scope="module")
def solver():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = jetgpu.Solver(0)
    yield s
    s.close()


@pytest.fixture(scope="module")
def g():
    with np.load(GOLDEN, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def xyz(a):
    nz, ny, nx = a.shape[:3]
    return (nx, ny, nz)


def test_blas_bitwise_vs_jet(solver, g):
    A, x, b = g["blas_A"], g["blas_x"], g["blas_b"]
    out = empty(x.shape)
    solver.mvm(xyz(x), dev(A), dev(x), out)
    assert np.array_equal(host(out), g["blas_mvm"])
    solver.residual(xyz(x), dev(A), dev(x), dev(b), out)
    assert np.array_equal(host(out), g["blas_residual"])


def test_relax_bitwise_vs_jet(solver, g):
    A, b, x0 = g["relax_A"], g["relax_b"], g["relax_x0"]
    dx = dev(x0)
    solver.relax_red_black(xyz(x0), dev(A), dev(b), 1.3, dx)
    assert np.array_equal(host(dx), g["relax_rb"])
    out = empty(x0.shape)
    solver.relax_jacobi(xyz(x0), dev(A), dev(b), dev(x0), out)
    assert np.array_equal(host(out), g["relax_jacobi"])


def test_transfer_bitwise_vs_jet(solver, g):
    fine, coarse = g["mg_fine"], g["mg_coarse"]
    out = empty(coarse.shape)
    solver.restrict(xyz(coarse), dev(fine), out)
    assert np.array_equal(host(out), g["mg_restrict"])
    df = dev(fine)
    solver.correct(xyz(fine), dev(coarse), df)
    assert np.array_equal(host(df), g["mg_correct"])


def test_cg_vs_jet(solver, g):
    A, b = g["cg_A"], g["cg_b"]
    solver.set_levels([xyz(b)])
    solver.upload(0, A, b)
    info = solver.cg(300, 1e-10)
    x = solver.download_x()
    assert info.converged
    assert abs(int(info.iterations) - int(g["cg_iterations"])) <= 2
    assert info.last_residual <= 1e-10
    assert rel_err(x, g["cg_x"]) <= 1e-9
