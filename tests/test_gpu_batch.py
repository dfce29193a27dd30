"""Batched MGPCG: independent pressure systems of uneven sizes solved by one
captured hipGraph with per-system convergence flags.

Per system the batch runs the same kernels in the same order as a single
jg_solver_mgpcg solve; the only differences are the sweep tiling the
autotuner picks (tuning is shared between identical hierarchies) and the
replayed no-op iterations after convergence.  So each batched solution must
match the single-system solve to rounding (1e-10 relative, iterations within
one), and meet the parity bar: relative residual <= 1e-8, FP64 and mixed
solutions within 1e-6 of each other.
"""
import numpy as np
import pytest

from jgtest import empty, jetgpu, rel_err, torch_cuda

pytestmark = pytest.mark.gpu

PARAMS = dict(restrict_iter=2, correct_iter=2, coarsest_iter=20, final_iter=2, sor=1.3)


def levels_for(n, max_levels=5):
    lv = 1
    while lv < max_levels and n % (1 << lv) == 0 and (n >> lv) >= 2:
        lv += 1
    return lv


def fill_system(solver, n, seed):
    """Batch-member scene (1-4 spheres, air above a random plane) built on the device."""
    d = jetgpu.scene_desc(jetgpu.JG_SCENE_BATCH_MEMBER, n, seed=seed)
    u, v, w = empty((n, n, n + 1)), empty((n, n + 1, n)), empty((n + 1, n, n))
    bs, fs = empty((n, n, n)), empty((n, n, n))
    solver.generate_scene(d, u, v, w, bs, fs)
    solver.sp_build(u, v, w, bs, fs, 1.0)


def solve_single(n, seed, levels, precision):
    s = jetgpu.Solver(0)
    try:
        s.set_levels([(n >> l,) * 3 for l in range(levels)])
        fill_system(s, n, seed)
        info = s.mgpcg(500, jetgpu.mg_params(max_levels=levels, **PARAMS),
                       jetgpu.gpu_options(precision=precision, stop_rule=jetgpu.JG_STOP_RELATIVE,
                                          rel_tol=1e-9))
        return s.download_x(), info
    finally:
        s.close()


def make_batch(sizes, seeds):
    levels = [levels_for(n) for n in sizes]
    b = jetgpu.Batch([(n, n, n) for n in sizes], levels)
    for sys_, n, seed in zip(b.systems, sizes, seeds):
        fill_system(sys_, n, seed)
    return b, levels


@pytest.fixture(scope="module")
def gpu():
    if torch_cuda() is None:
        pytest.skip("no GPU")


@pytest.mark.parametrize("precision", [jetgpu.JG_FP64, jetgpu.JG_MIXED])
def test_batch_matches_single_solves(gpu, precision):
    sizes = [32, 48, 64, 48, 32, 16]
    seeds = [101, 102, 103, 104, 105, 106]
    print(f"case batch small precision={precision}", flush=True)
    b, levels = make_batch(sizes, seeds)
    try:
        infos, ms = b.mgpcg(500, jetgpu.mg_params(max_levels=5, **PARAMS),
                            jetgpu.gpu_options(precision=precision, stop_rule=jetgpu.JG_STOP_RELATIVE,
                                               rel_tol=1e-9))
        xs = [s.download_x() for s in b.systems]
    finally:
        b.close()
    for i, (n, seed, lv) in enumerate(zip(sizes, seeds, levels)):
        x1, info1 = solve_single(n, seed, lv, precision)
        assert infos[i].rel_residual <= 1e-8, i
        assert abs(int(infos[i].iterations) - int(info1.iterations)) <= 1, i
        assert rel_err(xs[i], x1) <= 1e-10, i


def test_batch_64_uneven_systems(gpu):
    # 64 systems, edge lengths drawn from {32, 48, 64, 96, 128}; FP64 and
    # mixed precision, each through one replayed graph.
    rng = np.random.default_rng(2024)
    sizes = [int(n) for n in rng.choice([32, 48, 64, 96, 128], 64)]
    seeds = [1000 + i for i in range(64)]
    print(f"case batch 64 cells={sum(n ** 3 for n in sizes)}", flush=True)
    b, levels = make_batch(sizes, seeds)
    try:
        res = {}
        for precision in (jetgpu.JG_FP64, jetgpu.JG_MIXED):
            p = jetgpu.mg_params(max_levels=5, **PARAMS)
            o = jetgpu.gpu_options(precision=precision, stop_rule=jetgpu.JG_STOP_RELATIVE,
                                   rel_tol=1e-9)
            b.mgpcg(500, p, o)  # tuning and graph capture
            infos, ms = b.mgpcg(500, p, o)
            print(f"  precision={precision} batch ms={ms:.2f} "
                  f"max iterations={max(i.iterations for i in infos)}", flush=True)
            assert all(i.rel_residual <= 1e-8 for i in infos)
            res[precision] = [s.download_x() for s in b.systems]
        for a, c in zip(res[jetgpu.JG_FP64], res[jetgpu.JG_MIXED]):
            assert rel_err(c, a) <= 1e-6
    finally:
        b.close()
    # one member of each size against the oracle's FP64 solution of the same
    # system (host copy of the same generated scene, clean-mode MGPCG to 1e-12)
    from oracle.jet_oracle import MgParams, Oracle

    oracle = Oracle()
    first = {}
    for i, n in enumerate(sizes):
        first.setdefault(n, i)
    for n, i in sorted(first.items()):
        scene = jetgpu.generate_scene_host(jetgpu.scene_desc(jetgpu.JG_SCENE_BATCH_MEMBER, n,
                                                             seed=seeds[i]))
        _, sys_ = oracle.sp_build_mg(scene["u"], scene["v"], scene["w"], (1.0,) * 3, (0.0,) * 3,
                                     scene["bsdf"], scene["fsdf"], levels[i])
        ok, _, rel = oracle.mgpcg_clean_solve(sys_, MgParams.make(max_levels=levels[i], **PARAMS),
                                              300, 1e-12)
        assert ok and rel <= 1e-12
        for precision in (jetgpu.JG_FP64, jetgpu.JG_MIXED):
            err = rel_err(res[precision][i], sys_.x[0])
            print(f"  member {i} n={n} precision={precision} err_vs_oracle={err:.2e}", flush=True)
            assert err <= 1e-6, (i, n, precision)
