"""GPU assembly, projection and solves of the variational (fractional)
pressure system against the oracle restatement of
GridFractionalSinglePhasePressureSolver3.

Face weights and restricted fields are float computations and the matrix and
right-hand side FP64 computations done with the same IEEE operations in the
same order as the oracle, so they must be bit-identical on every level.
Solutions follow the solver parity bar: relative residual <= 1e-8 and
||p - p_ref|| / ||p_ref|| <= 1e-6, also at the 512^3 size of the torus
workload (the oracle's solve there takes about a minute of host time).
"""
import numpy as np
import pytest

from jgtest import dev, empty, host, jetgpu, rel_err, torch_cuda
from test_gpu_pressure import copy_to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = jetgpu.Solver(0)
    yield s
    s.close()


def torus(n, seed):
    return jetgpu.generate_scene_host(jetgpu.scene_desc(jetgpu.JG_SCENE_TORUS_POOL, n, seed=seed),
                                      vertex=True)


def build_both(solver, oracle, s, levels, collider=None):
    kw = {}
    if collider is not None:
        kw = dict(bu=collider[0], bv=collider[1], bw=collider[2])
    (fs, UW, VW, WW), sys_ = oracle.fr_build_mg(s["u"], s["v"], s["w"], (1.0,) * 3, (0.0,) * 3,
                                                s["vertex"], s["fsdf"], levels, **kw)
    solver.set_levels([(a.shape[2], a.shape[1], a.shape[0]) for a in fs])
    g = {k: dev(v) for k, v in kw.items()}
    solver.fr_build(dev(s["u"]), dev(s["v"]), dev(s["w"]), dev(s["vertex"]), dev(s["fsdf"]), 1.0,
                    **g)
    return (fs, UW, VW, WW), sys_


@pytest.mark.parametrize("n,levels", [(32, 3), (48, 4)])
@pytest.mark.parametrize("moving", [False, True])
def test_fr_build_matches_oracle(solver, oracle, n, levels, moving):
    print(f"case fr_build n={n} levels={levels} moving={moving}", flush=True)
    s = torus(n, seed=3)
    collider = None
    if moving:
        rng = np.random.default_rng(4)
        collider = tuple(rng.uniform(-1, 1, s[k].shape) for k in ("u", "v", "w"))
    (fs, UW, VW, WW), sys_ = build_both(solver, oracle, s, levels, collider)
    assert len(fs) == levels
    for l in range(levels):
        for which, ref in enumerate((UW[l], VW[l], WW[l], fs[l])):
            got = np.empty(ref.shape, np.float32)
            copy_to_host(solver.fr_array(l, which), got)
            assert np.array_equal(got, ref), (l, which)
        A = np.empty(sys_.A[l].shape)
        copy_to_host(solver.matrix_ptr(l), A)
        assert np.array_equal(A, sys_.A[l]), l
    b = np.empty(sys_.b[0].shape)
    copy_to_host(solver.rhs_ptr(), b)
    assert np.array_equal(b, sys_.b[0])
    # the scene exercises partial weights, ghost-fluid rows and solid rows
    assert 0.0 < (UW[0] < 1).mean() < 0.5 and (UW[0] == 0).any()


def test_fr_gradient_matches_oracle(solver, oracle):
    s = torus(40, seed=5)
    (fs, UW, VW, WW), sys_ = build_both(solver, oracle, s, 1)
    info = solver.cg(3000, 1e-9)
    p = solver.download_x()
    uo, vo, wo = dev(s["u"]), dev(s["v"]), dev(s["w"])
    solver.fr_apply_gradient(dev(s["u"]), dev(s["v"]), dev(s["w"]), uo, vo, wo, 1.0)
    ou, ov, ow = oracle.fr_apply_gradient(fs[0], UW[0], VW[0], WW[0], p, s["u"], s["v"], s["w"],
                                          (1.0,) * 3)
    for g, o_ in ((uo, ou), (vo, ov), (wo, ow)):
        assert np.array_equal(host(g), o_)


def test_fr_free_surface_known_answer(solver):
    # GridFractionalSinglePhasePressureSolver3.SolveFreeSurface
    # (src/tests/unit_tests/grid_fractional_single_phase_pressure_solver3_tests.cpp:17-74):
    # ghost-fluid theta = 0.5 at the surface gives p = 1.5 - j, and the
    # projected velocity vanishes.
    n = 3
    u, v, w = np.zeros((n, n, n + 1)), np.zeros((n, n + 1, n)), np.zeros((n + 1, n, n))
    v[:, 1:3, :] = 1.0
    z, y, x = np.meshgrid(np.arange(n) + 0.5, np.arange(n) + 0.5, np.arange(n) + 0.5, indexing="ij")
    fsdf = np.ascontiguousarray(y - 2.0)
    vertex = np.full((n + 1, n + 1, n + 1), np.finfo(np.float64).max)
    solver.set_levels([(n, n, n)])
    du, dv, dw = dev(u), dev(v), dev(w)
    solver.fr_build(du, dv, dw, dev(vertex), dev(fsdf), 1.0)
    solver.cg(200, 1e-12)
    solver.fr_apply_gradient(du, dv, dw, du, dv, dw, 1.0)
    p = solver.download_x()
    for a in (du, dv, dw):
        assert np.abs(host(a)).max() < 1e-6
    for j in range(2):
        np.testing.assert_allclose(p[:, j, :], 1.5 - j, atol=1e-6)


@pytest.mark.parametrize("precision", [jetgpu.JG_FP64, jetgpu.JG_MIXED])
def test_fr_mgpcg_torus_matches_oracle(solver, oracle, precision):
    from oracle.jet_oracle import MgParams

    n, levels = 64, 4
    print(f"case fr_mgpcg n={n} precision={precision}", flush=True)
    s = torus(n, seed=7)
    _, sys_ = build_both(solver, oracle, s, levels)
    ok, it, rel = oracle.mgpcg_clean_solve(sys_, MgParams.make(max_levels=levels, restrict_iter=2,
                                                                correct_iter=2, coarsest_iter=20,
                                                                final_iter=2, sor=1.3), 200, 1e-12)
    assert ok and rel <= 1e-12
    p = jetgpu.mg_params(max_levels=levels, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                         final_iter=2, sor=1.3)
    o = jetgpu.gpu_options(precision=precision, stop_rule=jetgpu.JG_STOP_RELATIVE, rel_tol=1e-10)
    info = solver.mgpcg(300, p, o)
    x = solver.download_x()
    assert info.rel_residual <= 1e-8
    assert rel_err(x, sys_.x[0]) <= 1e-6


def test_fr_torus_512_matches_oracle(solver):
    # The variational 512^3 workload (torus R = 150, r = 50 cells around the
    # y axis, liquid below 0.35 n), 7 levels, generated and assembled on the
    # device.  Both precisions reach 1e-8 relative residual, agree to 1e-6,
    # and are within 1e-6 of the oracle's FP64 solution.
    import torch

    n, levels = 512, 7
    print(f"case fr_torus n={n} levels={levels}", flush=True)
    d = jetgpu.scene_desc(jetgpu.JG_SCENE_TORUS_POOL, n, seed=11)
    f64 = torch.float64
    u = torch.empty((n, n, n + 1), dtype=f64, device="cuda")
    v = torch.empty((n, n + 1, n), dtype=f64, device="cuda")
    w = torch.empty((n + 1, n, n), dtype=f64, device="cuda")
    bs = torch.empty((n, n, n), dtype=f64, device="cuda")
    fs = torch.empty((n, n, n), dtype=f64, device="cuda")
    vs = torch.empty((n + 1, n + 1, n + 1), dtype=f64, device="cuda")
    torch.cuda.synchronize()
    solver.set_levels([(n >> l,) * 3 for l in range(levels)])
    solver.generate_scene(d, u, v, w, bs, fs, vs)
    solver.fr_build(u, v, w, vs, fs, 1.0)
    fluid = float(((fs < 0) & (bs >= 0)).double().mean().item())
    del u, v, w, bs, fs, vs
    torch.cuda.empty_cache()
    assert 0.2 < fluid < 0.4
    p = jetgpu.mg_params(max_levels=levels, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                         final_iter=2, sor=1.3)
    xs = {}
    for prec in (jetgpu.JG_FP64, jetgpu.JG_MIXED):
        info = solver.mgpcg(300, p, jetgpu.gpu_options(precision=prec,
                                                        stop_rule=jetgpu.JG_STOP_RELATIVE,
                                                        rel_tol=1e-9))
        print(f"  precision={prec} iterations={info.iterations} rel={info.rel_residual:.2e} "
              f"ms={info.solve_ms:.1f}", flush=True)
        assert info.rel_residual <= 1e-8
        xs[prec] = solver.download_x()
    assert rel_err(xs[jetgpu.JG_MIXED], xs[jetgpu.JG_FP64]) <= 1e-6
    # the FP64-oracle error of both: the oracle's weights, system and
    # clean-mode MGPCG (1e-11 relative residual) from the host copy of the
    # same generated scene
    import time

    from oracle.jet_oracle import MgParams, Oracle

    t0 = time.perf_counter()
    oracle = Oracle()
    sc = jetgpu.generate_scene_host(d, vertex=True)
    _, sys_ = oracle.fr_build_mg(sc["u"], sc["v"], sc["w"], (1.0,) * 3, (0.0,) * 3, sc["vertex"],
                                 sc["fsdf"], levels)
    del sc
    print(f"  oracle build {time.perf_counter() - t0:.1f} s", flush=True)
    ok, it, rel = oracle.mgpcg_clean_solve(sys_, MgParams.make(max_levels=levels, restrict_iter=2,
                                                                correct_iter=2, coarsest_iter=20,
                                                                final_iter=2, sor=1.3), 300, 1e-11)
    print(f"  oracle solve {it} iterations rel={rel:.1e} at {time.perf_counter() - t0:.1f} s",
          flush=True)
    assert ok and rel <= 1e-11
    for prec, x in xs.items():
        err = rel_err(x, sys_.x[0])
        print(f"  precision={prec} err_vs_oracle={err:.2e}", flush=True)
        assert err <= 1e-6
