"""The headline workload at its full size: the 256^3 sphere pool (8 solid
spheres, ~60% fluid below a free surface, 6 levels, V(2,2), 20 coarsest
sweeps, SOR 1.3), generated and assembled on the device, solved from x = 0 to
||b - A x|| <= 1e-8 ||b|| in FP64 and in mixed FP32/FP64 -- the solves the
benchmark times.  Both must be within 1e-6 of the oracle's FP64 solution of
the same system (its clean-mode MGPCG to 1e-12 relative residual, built from
the host copy of the same generated scene), the parity bar of the solver.
"""
import numpy as np
import pytest

from jgtest import empty, jetgpu, rel_err, torch_cuda

pytestmark = pytest.mark.gpu

N, LEVELS, OMEGA = 256, 6, 1.3


@pytest.fixture(scope="module")
def headline(oracle):
    if torch_cuda() is None:
        pytest.skip("no GPU")
    from oracle.jet_oracle import MgParams

    print(f"case headline n={N}: oracle build and 1e-12 solve", flush=True)
    desc = jetgpu.scene_desc(jetgpu.JG_SCENE_SPHERE_POOL, N, seed=1)
    scene = jetgpu.generate_scene_host(desc)
    markers, sys_ = oracle.sp_build_mg(scene["u"], scene["v"], scene["w"], (1.0,) * 3, (0.0,) * 3,
                                       scene["bsdf"], scene["fsdf"], LEVELS)
    del scene
    ok, it, rel = oracle.mgpcg_clean_solve(
        sys_, MgParams.make(max_levels=LEVELS, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                            final_iter=2, sor=OMEGA), 200, 1e-12)
    assert ok and rel <= 1e-12
    p_ref = sys_.x[0].copy()
    del sys_
    s = jetgpu.Solver(0)
    s.set_levels([(N >> l,) * 3 for l in range(LEVELS)])
    u, v, w = empty((N, N, N + 1)), empty((N, N + 1, N)), empty((N + 1, N, N))
    bs, fs = empty((N, N, N)), empty((N, N, N))
    s.generate_scene(desc, u, v, w, bs, fs)
    s.sp_build(u, v, w, bs, fs, 1.0)
    fluid = float(((bs >= 0) & (fs < 0)).double().mean().item())
    del u, v, w, bs, fs
    yield s, p_ref, markers[0], fluid
    s.close()


@pytest.mark.parametrize("precision", [jetgpu.JG_FP64, jetgpu.JG_MIXED])
def test_headline_256_matches_oracle(headline, precision):
    s, p_ref, m0, fluid = headline
    print(f"case headline n={N} precision={precision}", flush=True)
    assert 0.59 <= fluid <= 0.61
    p = jetgpu.mg_params(max_levels=LEVELS, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                         final_iter=2, sor=OMEGA)
    o = jetgpu.gpu_options(precision=precision, stop_rule=jetgpu.JG_STOP_RELATIVE, rel_tol=1e-8)
    info = s.mgpcg(500, p, o)
    x = s.download_x()
    err = rel_err(x, p_ref)
    print(f"  iterations={info.iterations} refinements={info.refinements} "
          f"rel={info.rel_residual:.2e} err_vs_oracle={err:.2e}", flush=True)
    assert info.converged and info.rel_residual <= 1e-8
    assert err <= 1e-6
    # solid and air cells are identity rows with b = 0
    assert np.max(np.abs(x[m0 != 0])) <= 1e-6 * np.max(np.abs(p_ref))
    assert info.iterations <= 20
