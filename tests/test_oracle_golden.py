"""The oracle against the committed vectors of Jet's own compiled code
(tests/golden/jet_linear.npz, written by tests/golden/make_golden.py from
oracle/_ref): FdmBlas3, red-black Gauss-Seidel, Jacobi and the multigrid
transfer operators bit for bit, FdmCgSolver3 iteration for iteration.
Runs wherever the oracle builds, with or without the reference tree."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "jet_linear.npz")


@pytest.fixture(scope="module")
def g():
    with np.load(GOLDEN, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def test_blas_bitwise_vs_jet_vectors(oracle, g):
    assert np.array_equal(oracle.mvm(g["blas_A"], g["blas_x"]), g["blas_mvm"])
    assert np.array_equal(oracle.residual(g["blas_A"], g["blas_x"], g["blas_b"]), g["blas_residual"])


def test_relax_bitwise_vs_jet_vectors(oracle, g):
    A, b, x0 = g["relax_A"], g["relax_b"], g["relax_x0"]
    assert np.array_equal(oracle.gs_relax(A, b, 1.3, x0, red_black=True), g["relax_rb"])
    assert np.array_equal(oracle.jacobi_relax(A, b, x0), g["relax_jacobi"])


def test_transfer_bitwise_vs_jet_vectors(oracle, g):
    assert np.array_equal(oracle.restrict(g["mg_fine"]), g["mg_restrict"])
    assert np.array_equal(oracle.correct(g["mg_coarse"], g["mg_fine"]), g["mg_correct"])


def test_cg_vs_jet_vectors(oracle, g):
    x, it, res, ok = oracle.cg_solve(g["cg_A"], g["cg_b"], 300, 1e-10)
    assert ok and it == int(g["cg_iterations"])
    assert np.array_equal(x, g["cg_x"])
    assert res == float(g["cg_last_residual"])
