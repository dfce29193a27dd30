"""GPU assembly and projection of the 2-D variational (fractional) pressure
system (jg_fr_build2 / jg_fr_apply_gradient2) against the oracle restatement of
GridFractionalSinglePhasePressureSolver2, which is itself bitwise equal to
Jet's compiled solver (tests/test_oracle_fractional2.py).

Weights and restricted fields are float computations, the system and the
gradient FP64 ones, all in the oracle's operand order with contraction off:
they must be bit-identical on every level.  Jet's three 2-D unit-test answers
(src/tests/unit_tests/grid_fractional_single_phase_pressure_solver2_tests.cpp:
17,63,109) pass through the GPU build, the GPU solves and the GPU gradient.
"""
import numpy as np
import pytest

from jgtest import dev, host, jetgpu, rel_err, torch_cuda
from test_gpu_pressure import copy_to_host
from test_oracle_fractional2 import KMAXD, scene2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = jetgpu.Solver(0)
    yield s
    s.close()


def build_both(solver, oracle, s, levels):
    (fs, UW, VW), sys_ = oracle.fr_build_mg2(s["u"], s["v"], s["h"], s["bsdf"], s["fsdf"], levels,
                                             s["bu"], s["bv"])
    solver.set_levels([(a.shape[1], a.shape[0], 1) for a in fs])
    g = {k: dev(s[k]) if s[k] is not None else None for k in ("bu", "bv")}
    solver.fr_build2(dev(s["u"]), dev(s["v"]), dev(s["bsdf"]), dev(s["fsdf"]), s["h"], **g)
    return (fs, UW, VW), sys_


@pytest.mark.parametrize("nx,ny,levels,moving", [(16, 16, 3, True), (33, 17, 1, True),
                                                 (64, 48, 4, False), (40, 24, 3, True),
                                                 (256, 128, 6, True)])
def test_fr_build2_matches_oracle(solver, oracle, nx, ny, levels, moving):
    print(f"case fr_build2 {nx}x{ny} levels={levels} moving={moving}", flush=True)
    s = scene2(nx, ny, 7, moving)
    (fs, UW, VW), sys_ = build_both(solver, oracle, s, levels)
    for l in range(len(fs)):
        for which, ref in ((0, UW[l]), (1, VW[l]), (3, fs[l])):
            got = np.empty(ref.shape, np.float32)
            copy_to_host(solver.fr_array(l, which), got)
            assert np.array_equal(got, ref), (l, which)
        A = np.empty(sys_.A[l].shape)
        copy_to_host(solver.matrix_ptr(l), A)
        assert np.array_equal(A, sys_.A[l]), l
    b = np.empty(sys_.b[0].shape)
    copy_to_host(solver.rhs_ptr(), b)
    assert np.array_equal(b, sys_.b[0])
    # partial weights, ghost-fluid rows and solid rows are all present
    assert 0.0 < (UW[0] < 1).mean() < 0.5 and (UW[0] == 0).any()


def test_fr_gradient2_matches_oracle(solver, oracle):
    import torch

    s = scene2(48, 40, 11, True)
    (fs, UW, VW), _ = build_both(solver, oracle, s, 1)
    p = np.random.default_rng(12).uniform(-1, 1, (40, 48))
    x = torch.as_tensor(jetgpu._DeviceView(solver.solution_ptr(), p.size, "<f8"), device="cuda")
    x.copy_(torch.from_numpy(p.ravel()).cuda())
    torch.cuda.synchronize()
    uo, vo = dev(s["u"]), dev(s["v"])
    solver.fr_apply_gradient2(dev(s["u"]), dev(s["v"]), uo, vo, s["h"])
    want_u, want_v = oracle.fr_apply_gradient2(fs[0], UW[0], VW[0], p, s["u"], s["v"], s["h"])
    assert np.array_equal(host(uo), want_u) and np.array_equal(host(vo), want_v)
    # in place (output aliases the input)
    u, v = dev(s["u"]), dev(s["v"])
    solver.fr_apply_gradient2(u, v, u, v, s["h"])
    assert np.array_equal(host(u), want_u) and np.array_equal(host(v), want_v)


def free_surface(n, surface):
    u = np.zeros((n, n + 1))
    v = np.ones((n + 1, n))
    v[0, :] = 0.0
    v[n, :] = 0.0
    _, FY = np.meshgrid(np.arange(n) + 0.5, np.arange(n) + 0.5)
    return dict(u=u, v=v, h=(1.0, 1.0), bsdf=np.full((n + 2, n + 2), KMAXD), fsdf=FY - surface,
                bu=None, bv=None)


def test_jet_free_surface_3x3_on_gpu(solver, oracle):
    # SolveFreeSurface (:17): p = 1.5, 0.5, 0 by row; the projected velocity is zero
    s = free_surface(3, 2.0)
    build_both(solver, oracle, s, 1)
    info = solver.cg(100, 1e-12)
    p = solver.download_x()[0]
    for i in range(3):
        assert abs(p[0, i] - 1.5) <= 1e-6 and abs(p[1, i] - 0.5) <= 1e-6 and abs(p[2, i]) <= 1e-6
    u, v = dev(s["u"]), dev(s["v"])
    solver.fr_apply_gradient2(u, v, u, v, s["h"])
    assert np.max(np.abs(host(u))) <= 1e-6 and np.max(np.abs(host(v))) <= 1e-6
    assert info.converged


@pytest.mark.parametrize("precision", [jetgpu.JG_FP64, jetgpu.JG_MIXED])
def test_jet_free_surface_mg_on_gpu(solver, oracle, precision):
    # SolveFreeSurfaceMg (:109): 32 x 32, surface at y = 16, 5 levels, MGPCG on the GPU
    s = free_surface(32, 16.0)
    _, sys_ = build_both(solver, oracle, s, 5)
    p = jetgpu.mg_params(max_levels=5, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                         final_iter=2)
    info = solver.mgpcg(200, p, jetgpu.gpu_options(precision=precision,
                                                   stop_rule=jetgpu.JG_STOP_RELATIVE, rel_tol=1e-10))
    assert info.converged and info.rel_residual <= 1e-8
    x = solver.download_x()[0]
    for j in range(32):
        want = 15.5 - j if j < 16 else 0.0
        assert abs(x[j, 16] - want) <= 0.1
    u, v = dev(s["u"]), dev(s["v"])
    solver.fr_apply_gradient2(u, v, u, v, s["h"])
    assert np.max(np.abs(host(u))) <= 0.002 and np.max(np.abs(host(v)[:16])) <= 0.002


@pytest.mark.parametrize("precision", [jetgpu.JG_FP64, jetgpu.JG_MIXED])
def test_fr2_mgpcg_matches_oracle(solver, oracle, precision):
    # the collider scene: GPU build + GPU MGPCG against the oracle's 1e-12 solve
    from oracle.jet_oracle import MgParams

    s = scene2(128, 96, 13, True)
    (fs, _, _), sys_ = build_both(solver, oracle, s, 5)
    ok, _, rel = oracle.mgpcg_clean_solve(sys_, MgParams.make(max_levels=5, restrict_iter=2,
                                                              correct_iter=2, coarsest_iter=20,
                                                              final_iter=2), 200, 1e-12)
    assert ok and rel <= 1e-12
    p = jetgpu.mg_params(max_levels=5, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                         final_iter=2)
    info = solver.mgpcg(300, p, jetgpu.gpu_options(precision=precision,
                                                   stop_rule=jetgpu.JG_STOP_RELATIVE, rel_tol=1e-10))
    assert info.rel_residual <= 1e-8
    assert rel_err(solver.download_x(), sys_.x[0]) <= 1e-6


def test_fr_build2_rejects_3d_hierarchy(solver):
    solver.set_levels([(8, 8, 8), (4, 4, 4)])
    z = dev(np.zeros(10))
    with pytest.raises(jetgpu.JgError, match="2-D"):
        solver.fr_build2(z, z, z, z, (1.0, 1.0))
