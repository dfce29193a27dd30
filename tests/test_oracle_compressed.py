"""Pins the oracle's restatement of Jet's compressed (CSR) path:
FdmCompressedBlas3 mvm / residual and FdmCgSolver3::solveCompressed bitwise
and iteration for iteration against Jet's own compiled code (oracle/_ref),
and the compressed single-phase build against the flat build (itself pinned
by Jet's pressure known answers) and Jet's SolveFreeSurfaceCompressed answer
(src/tests/unit_tests/grid_single_phase_pressure_solver3_tests.cpp:191-247).
"""
import numpy as np
import pytest

from systems import csr_from_stencil, poisson_csr_kat

BIG = np.finfo(np.float64).max


def random_markers(n, seed, fluid=0.7):
    nx, ny, nz = n
    rng = np.random.default_rng(seed)
    m = np.where(rng.random((nz, ny, nx)) < fluid, 0, np.where(rng.random((nz, ny, nx)) < 0.5, 1, 2))
    return m.astype(np.int8), rng


def random_faces(rng, n):
    nx, ny, nz = n
    return (rng.uniform(-1, 1, (nz, ny, nx + 1)), rng.uniform(-1, 1, (nz, ny + 1, nx)),
            rng.uniform(-1, 1, (nz + 1, ny, nx)))


@pytest.mark.parametrize("n,h,seed", [((14, 11, 9), 0.5, 1), ((8, 8, 8), 1.0, 2), ((5, 17, 3), 0.25, 3),
                                      ((1, 1, 1), 1.0, 4)])
def test_compressed_build_equals_flat_build_on_fluid_cells(oracle, n, h, seed):
    m, rng = random_markers(n, seed)
    if not (m == 0).any():
        m.flat[0] = 0
    u, v, w = random_faces(rng, n)
    val, rp, ci, b = oracle.sp_build_compressed(m, u, v, w, (h, h, h))
    A, bf = oracle.sp_build_system(m, u, v, w, (h, h, h))
    fluid = m == 0
    wv, wrp, wci = csr_from_stencil(A, fluid)
    assert np.array_equal(rp, wrp) and np.array_equal(ci, wci) and np.array_equal(val, wv)
    assert np.array_equal(b, bf[fluid])


def test_compressed_free_surface_kat(oracle):
    # Jet's SolveFreeSurfaceCompressed: 3^3 cells, v = 1 on interior y faces,
    # fluid below y = 2: p = 2 - j on the fluid rows (j = 0, 1).
    n = (3, 3, 3)
    u, v, w = np.zeros((3, 3, 4)), np.zeros((3, 4, 3)), np.zeros((4, 3, 3))
    v[:, 1:3, :] = 1.0
    z, y, x = np.meshgrid(np.arange(3) + 0.5, np.arange(3) + 0.5, np.arange(3) + 0.5, indexing="ij")
    m = oracle.sp_markers(np.full(n, BIG), np.ascontiguousarray(y - 2.0))
    val, rp, ci, b = oracle.sp_build_compressed(m, u, v, w, (1.0, 1.0, 1.0))
    p, ok, it, res = oracle.csr_cg_solve(val, rp, ci, b, 100, 1e-12)
    assert ok and b.size == 18
    jj = np.argwhere(m == 0)[:, 1]  # fluid rows in x-fastest order: (k, j, i)
    np.testing.assert_allclose(p, 2.0 - jj, atol=1e-6)


def test_cg_compressed_kat(oracle):
    # Jet's FdmCgSolver3.SolveCompressed: residual below the 1e-9 tolerance
    val, rp, ci, b = poisson_csr_kat((3, 3, 3))
    x, ok, it, res = oracle.csr_cg_solve(val, rp, ci, b, 100, 1e-9)
    assert ok and res < 1e-9


@pytest.mark.parametrize("n,seed", [((14, 11, 9), 1), ((9, 9, 9), 2), ((2, 3, 4), 3)])
def test_csr_blas_and_cg_match_jet(oracle, ref, n, seed):
    m, rng = random_markers(n, seed)
    u, v, w = random_faces(rng, n)
    val, rp, ci, b = oracle.sp_build_compressed(m, u, v, w, (0.5, 0.5, 0.5))
    if b.size == 0:
        pytest.skip("no fluid rows")
    x = rng.uniform(-1, 1, b.size)
    assert np.array_equal(oracle.csr_mvm(val, rp, ci, x), ref.csr_mvm(val, rp, ci, x))
    assert np.array_equal(oracle.csr_residual(val, rp, ci, x, b), ref.csr_residual(val, rp, ci, x, b))
    xo, oko, ito, reso = oracle.csr_cg_solve(val, rp, ci, b, 500, 1e-10)
    xr, okr, itr, resr = ref.csr_cg_solve(val, rp, ci, b, 500, 1e-10)
    # Jet's VectorN::dot is a parallel reduction, so its summation order (and
    # the last bits) depend on the thread count; iterations and solution agree
    assert oko == okr and abs(ito - itr) <= 2
    assert np.abs(xo - xr).max() <= 1e-8 * max(1.0, np.abs(xr).max())


def test_csr_kat_matches_jet(oracle, ref):
    val, rp, ci, b = poisson_csr_kat((3, 3, 3))
    xo, oko, ito, reso = oracle.csr_cg_solve(val, rp, ci, b, 100, 1e-9)
    xr, okr, itr, resr = ref.csr_cg_solve(val, rp, ci, b, 100, 1e-9)
    assert oko == okr and ito == itr and abs(reso - resr) <= 1e-12
