"""GPU 2-D single-phase pressure build (GridSinglePhasePressureSolver2 on a
one-plane hierarchy) against the oracle: markers and rows on every level
and the right-hand side bit for bit, the gradient bit for bit, and Jet's 2-D
known answers through the GPU build, the GPU 2-D MGPCG and the GPU gradient."""
import numpy as np
import pytest

from jgtest import dev, empty, host, jetgpu, torch_cuda
from test_gpu_pressure import copy_to_host
from test_oracle_pressure2 import BIG, cc2, kat_velocity2, random_scene2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = jetgpu.Solver(0)
    yield s
    s.close()


@pytest.mark.parametrize("n,levels,seed", [((64, 32), 4, 1), ((96, 96), 5, 2), ((30, 18), 2, 3),
                                           ((1024, 512), 6, 4), ((5, 7), 1, 5)])
def test_sp_build2_matches_oracle(solver, oracle, n, levels, seed):
    print(f"case sp_build2 n={n} levels={levels}", flush=True)
    nx, ny = n
    h = 1.0 / max(n)
    bsdf, fsdf, u, v = random_scene2(nx, ny, np.random.default_rng(seed), h)
    markers, sys_ = oracle.sp_build_mg2(u, v, (h, h), bsdf, fsdf, levels)
    solver.set_levels([(m.shape[1], m.shape[0], 1) for m in markers])
    solver.sp_build2(dev(u), dev(v), dev(bsdf), dev(fsdf), (h, h))
    for l, m in enumerate(markers):
        gm = np.empty(m.shape, np.int8)
        A = np.empty(sys_.A[l].shape)
        copy_to_host(jetgpu.lib().jg_solver_markers(solver.h, l), gm)
        copy_to_host(solver.matrix_ptr(l), A)
        assert np.array_equal(gm, m), l
        assert np.array_equal(A, sys_.A[l]), l
    b = np.empty(sys_.b[0].shape)
    copy_to_host(solver.rhs_ptr(), b)
    assert np.array_equal(b, sys_.b[0])
    # the gradient of a random p placed in the resident solution
    p = np.random.default_rng(seed + 10).uniform(-1, 1, (ny, nx))
    x = torch_cuda().as_tensor(jetgpu._DeviceView(solver.solution_ptr(), nx * ny, "<f8"),
                               device="cuda")
    x.copy_(dev(p).reshape(-1))
    torch_cuda().cuda.synchronize()
    uo, vo = dev(u), dev(v)
    solver.sp_apply_gradient2(dev(u), dev(v), uo, vo, (h, h))
    wu, wv = oracle.sp_apply_gradient2(markers[0], p, u, v, (h, h))
    assert np.array_equal(host(uo), wu) and np.array_equal(host(vo), wv)


@pytest.mark.parametrize("case", ["single", "boundary", "free"])
@pytest.mark.parametrize("precision", [jetgpu.JG_FP64, jetgpu.JG_MIXED])
def test_kat2_through_gpu(solver, case, precision):
    # Jet's 2-D single-phase KATs (src/tests/unit_tests/
    # grid_single_phase_pressure_solver2_tests.cpp:16-199) with the build,
    # the solve (2-D MGPCG on 3 x 3, one level) and the gradient on the GPU
    u, v = kat_velocity2()
    bs = cc2(lambda x, y: -x + 2.0, 3, 3) if case == "boundary" else cc2(lambda x, y: BIG + 0 * x, 3, 3)
    fs = cc2(lambda x, y: y - 2.0, 3, 3) if case == "free" else cc2(lambda x, y: -BIG + 0 * x, 3, 3)
    solver.set_levels([(3, 3, 1)])
    solver.sp_build2(dev(u), dev(v), dev(bs), dev(fs), (1.0, 1.0))
    gp = jetgpu.mg_params(max_levels=1, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                          final_iter=2)
    info = solver.mgpcg(100, gp, jetgpu.gpu_options(precision=precision,
                                                    stop_rule=jetgpu.JG_STOP_RELATIVE,
                                                    rel_tol=1e-12))
    assert info.rel_residual <= 1e-10
    p = solver.download_x()[0]
    uo, vo = dev(u), dev(v)
    solver.sp_apply_gradient2(dev(u), dev(v), uo, vo, (1.0, 1.0))
    uo, vo = host(uo), host(vo)
    assert np.abs(uo).max() < 1e-6
    if case == "boundary":
        expect = np.zeros_like(vo)
        expect[1:3, 2] = 1.0
        np.testing.assert_allclose(vo, expect, atol=1e-6)
        np.testing.assert_allclose(p[1:, :2] - p[:-1, :2], -1.0, atol=1e-6)
    elif case == "free":
        assert np.abs(vo).max() < 1e-6
        for j in range(2):
            np.testing.assert_allclose(p[j, :], 2.0 - j, atol=1e-6)
    else:
        assert np.abs(vo).max() < 1e-6
        np.testing.assert_allclose(p[1:, :] - p[:-1, :], -1.0, atol=1e-6)


def test_sp_build2_rejects_3d_hierarchy(solver):
    solver.set_levels([(8, 8, 8), (4, 4, 4)])
    z = empty((8 * 8 * 9,))
    with pytest.raises(RuntimeError):
        solver.sp_build2(z, z, z, z, (1.0, 1.0))


def test_3d_builds_reject_2d_hierarchy(solver):
    # the 3-D coarsening reads fine planes 2k and 2k+1; a one-plane hierarchy must be refused
    # before any launch
    solver.set_levels([(8, 8, 1), (4, 4, 1)])
    z = empty((9 * 9 * 2,))
    with pytest.raises(RuntimeError):
        solver.sp_build(z, z, z, z, z, 1.0)
    with pytest.raises(RuntimeError):
        solver.fr_build(z, z, z, z, z, 1.0)


@pytest.mark.parametrize("n,seed", [((64, 32), 11), ((96, 96), 12), ((7, 5), 13), ((1024, 512), 14)])
def test_sp_build_compressed2_matches_oracle(solver, oracle, n, seed):
    print(f"case sp_build_compressed2 n={n}", flush=True)
    nx, ny = n
    h = 1.0 / max(n)
    bsdf, fsdf, u, v = random_scene2(nx, ny, np.random.default_rng(seed), h)
    m = oracle.sp_markers(bsdf, fsdf)
    val, rp, ci, b = oracle.sp_build_compressed2(m, u, v, (h, h))
    solver.set_levels([(nx, ny, 1)])
    rows, nnz = solver.sp_build_compressed2(dev(u), dev(v), dev(bsdf), dev(fsdf), (h, h))
    assert (rows, nnz) == (b.size, val.size)
    gval, grp, gci, gb, _ = solver.csr_download(rows, nnz)
    assert np.array_equal(gval, val) and np.array_equal(gb, b)
    assert np.array_equal(grp.astype(np.uint64), rp) and np.array_equal(gci.astype(np.uint64), ci)


@pytest.mark.parametrize("case", ["single", "free"])
def test_kat2_compressed_through_gpu(solver, case):
    # SolveSinglePhaseCompressed / SolveFreeSurfaceCompressed (src/tests/
    # unit_tests/grid_single_phase_pressure_solver2_tests.cpp:60-102, 201-246)
    # with the compressed build, Jet's CG, decompressSolution and the
    # gradient on the device
    u, v = kat_velocity2()
    bs = cc2(lambda x, y: BIG + 0 * x, 3, 3)
    fs = cc2(lambda x, y: y - 2.0, 3, 3) if case == "free" else cc2(lambda x, y: -BIG + 0 * x, 3, 3)
    solver.set_levels([(3, 3, 1)])
    solver.sp_build_compressed2(dev(u), dev(v), dev(bs), dev(fs), (1.0, 1.0))
    info = solver.csr_solve(200, 1e-12)
    assert info.converged
    solver.csr_decompress()
    p = solver.download_x()[0]
    uo, vo = dev(u), dev(v)
    solver.sp_apply_gradient2(dev(u), dev(v), uo, vo, (1.0, 1.0))
    uo, vo = host(uo), host(vo)
    assert np.abs(uo).max() < 1e-6 and np.abs(vo).max() < 1e-6
    if case == "free":
        for j in range(2):
            np.testing.assert_allclose(p[j, :], 2.0 - j, atol=1e-6)
    else:
        np.testing.assert_allclose(p[1:, :] - p[:-1, :], -1.0, atol=1e-6)
