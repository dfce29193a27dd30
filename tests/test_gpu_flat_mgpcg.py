"""The flat-entry MGPCG (jg_solver_coarsen, jg_mgpcg_flat_solve_host) through
the C ABI, against the oracle.

- Coarse matrices built on the device from level 0 are bitwise equal to the
  oracle's or_mg_coarsen on every level (FP64, fixed summation order,
  contraction off): random stencils, diffusion systems with colliders and
  Dirichlet faces, and one-plane hierarchies.
- GridBackwardEulerDiffusionSolver3's system (src/jet/
  grid_backward_euler_diffusion_solver3.cpp:194-263) of the sphere-pool scene,
  solved by the flat entry in FP64 and mixed precision, reaches the oracle's
  1e-12 CG solution to 1e-8, in about the iterations the oracle's MGPCG on
  the same hierarchy takes and far fewer than Jet's CG.
- The device-resident path (jg_diffusion_build on a multi-level hierarchy,
  jg_solver_coarsen, jg_solver_mgpcg) equals the host-array entry.
"""
import numpy as np
import pytest

from jgtest import dev, jetgpu, rel_err, torch_cuda
from systems import random_stencil
from test_gpu_pressure import copy_to_host

pytestmark = pytest.mark.gpu

DT, MU = 1.0 / 60.0, 0.05


@pytest.fixture(scope="module")
def solver():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = jetgpu.Solver(0)
    yield s
    s.close()


def device_hierarchy(solver, oracle, A0, levels):
    want = oracle.mg_flat_hierarchy(A0, np.zeros(A0.shape[:3]), levels)
    solver.set_levels([(a.shape[2], a.shape[1], a.shape[0]) for a in want.A])
    solver.upload(0, A0, None)
    solver.coarsen()
    got = []
    for l, ref in enumerate(want.A):
        A = np.empty(ref.shape)
        copy_to_host(solver.matrix_ptr(l), A)
        got.append(A)
    return got, want


def diffusion_case(shape, dirichlet, seed):
    rng = np.random.default_rng(seed)
    f = rng.uniform(-1, 1, shape)
    bs = np.where(rng.random(shape) < 0.15, -1.0, 1.0)
    fs = np.where(rng.random(shape) < 0.8, -1.0, 1.0)
    return f, bs, fs


@pytest.mark.parametrize("shape,levels", [((16, 24, 32), 3), ((4, 4, 4), 2), ((1, 64, 96), 5),
                                          ((32, 32, 32), 6), ((2, 48, 8), 2)])
def test_coarsen_random_stencil_matches_oracle(solver, oracle, shape, levels):
    print(f"case coarsen random {shape} levels={levels}", flush=True)
    nz, ny, nx = shape
    A0 = random_stencil(nx, ny, nz, np.random.default_rng(31))
    got, want = device_hierarchy(solver, oracle, A0, levels)
    assert len(got) == levels
    for l in range(levels):
        assert np.array_equal(got[l], want.A[l]), l


@pytest.mark.parametrize("shape,levels,dirichlet", [((24, 40, 32), 3, False),
                                                    ((24, 40, 32), 3, True),
                                                    ((1, 80, 48), 4, True)])
def test_coarsen_diffusion_matches_oracle(solver, oracle, shape, levels, dirichlet):
    print(f"case coarsen diffusion {shape} levels={levels} dirichlet={dirichlet}", flush=True)
    f, bs, fs = diffusion_case(shape, dirichlet, 32)
    A0, _ = oracle.diffusion_build(f, bs, fs, 0.3, 0.7, (0.5, 0.25, 0.5), dirichlet)
    got, want = device_hierarchy(solver, oracle, A0, levels)
    for l in range(levels):
        assert np.array_equal(got[l], want.A[l]), l


def sphere_pool_diffusion(oracle, n, seed=3):
    sc = jetgpu.generate_scene_host(jetgpu.scene_desc(jetgpu.JG_SCENE_SPHERE_POOL, n, seed=seed))
    g = np.arange(n) / n
    f = np.ascontiguousarray(np.sin(6.0 * g)[:, None, None] * np.cos(4.0 * g)[None, :, None]
                             + g[None, None, :])
    A, b = oracle.diffusion_build(f, sc["bsdf"], sc["fsdf"], DT, MU, (1.0 / n,) * 3)
    return f, sc, A, b


@pytest.mark.parametrize("precision", [jetgpu.JG_FP64, jetgpu.JG_MIXED])
def test_flat_mgpcg_diffusion_matches_oracle(solver, oracle, precision):
    from oracle.jet_oracle import MgParams

    n, levels = 64, 5
    print(f"case flat mgpcg diffusion n={n} precision={precision}", flush=True)
    _, _, A, b = sphere_pool_diffusion(oracle, n)
    x_ref, cg_it, _, ok = oracle.cg_solve(A, b, 2000, 1e-12 * np.linalg.norm(b))
    assert ok
    kw = dict(max_levels=levels, restrict_iter=2, correct_iter=2, coarsest_iter=20, final_iter=2)
    s = oracle.mg_flat_hierarchy(A, b, levels)
    ok, it_ref, rel = oracle.mgpcg_clean_solve(s, MgParams.make(**kw), 200, 1e-10)
    assert ok and rel <= 1e-10
    x, info = solver.mgpcg_flat_host(A, b, 200, jetgpu.mg_params(**kw),
                                     jetgpu.gpu_options(precision=precision,
                                                        stop_rule=jetgpu.JG_STOP_RELATIVE,
                                                        rel_tol=1e-10))
    print(f"flat mgpcg: gpu {info.iterations} it, oracle {it_ref} it, CG {cg_it} it")
    assert info.converged and info.rel_residual <= 1e-10
    assert rel_err(x, x_ref) <= 1e-8
    assert abs(int(info.iterations) - int(it_ref)) <= (2 if precision == jetgpu.JG_FP64 else 6)
    assert info.iterations * 3 < cg_it


def test_flat_mgpcg_device_path_equals_host_entry(solver, oracle):
    import torch

    n, levels = 64, 5
    f, sc, A, b = sphere_pool_diffusion(oracle, n)
    p = jetgpu.mg_params(max_levels=levels, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                         final_iter=2)
    o = jetgpu.gpu_options(stop_rule=jetgpu.JG_STOP_RELATIVE, rel_tol=1e-10)
    x_host, info_host = solver.mgpcg_flat_host(A, b, 200, p, o)
    solver.set_levels([(n >> l,) * 3 for l in range(levels)])
    solver.diffusion_build(dev(f), dev(sc["bsdf"]), dev(sc["fsdf"]), DT, MU, 1.0 / n)
    solver.coarsen()
    info = solver.mgpcg(200, p, o)
    torch.cuda.synchronize()
    assert info.iterations == info_host.iterations
    assert np.array_equal(solver.download_x(), x_host)


def test_flat_mgpcg_one_plane(solver, oracle):
    # a one-plane system takes x/y halving and the 2-D transfers
    rng = np.random.default_rng(33)
    A = random_stencil(128, 96, 1, rng)
    b = rng.uniform(-1, 1, (1, 96, 128))
    x_ref, _, _, ok = oracle.cg_solve(A, b, 5000, 1e-13 * np.linalg.norm(b))
    assert ok
    x, info = solver.mgpcg_flat_host(A, b, 200, jetgpu.mg_params(max_levels=4, restrict_iter=2,
                                                                correct_iter=2, coarsest_iter=20,
                                                                final_iter=2),
                                     jetgpu.gpu_options(stop_rule=jetgpu.JG_STOP_RELATIVE,
                                                        rel_tol=1e-11))
    assert info.converged and rel_err(x, x_ref) <= 1e-8
