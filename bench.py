#!/usr/bin/env python3
"""Headline benchmark: 256^3 MGPCG pressure solve, mixed FP32/FP64 vs FP64.

Workload (the metric's configuration): a 256^3 single-phase pressure system
(GridSinglePhasePressureSolver3 layout) with 8 seeded solid spheres of radius
U(8,32) cells and ~60% fluid below a free surface, velocities U(-1,1) with the
blocked boundary condition; 6 multigrid levels (256 -> 8), 2 pre / 2 post
red-black sweeps, 20 coarsest sweeps.  Inputs are generated and the system
is assembled on the GPU, so everything is resident in HBM before timing.

A step is one full solve from x = 0 to ||b - A x|| <= 1e-8 ||b||:
  - mixed: FP64 residual/refinement, FP32 stencil + V-cycle, FP64 dots;
  - fp64 : plain FP64 MGPCG.
Timed with a device synchronisation on both sides of K solves.  The CPU leg
(rank 0, N=1) solves the same system with the oracle's FP64 reference solve
on the host cores, which also gives p_ref for the error report.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "fluid-engine-dev_amd")
sys.path.insert(0, PKG)
sys.path.insert(0, ROOT)

METRIC = ("256³ MGPCG solve ms + HBM GB/s: mixed FP32/FP64 vs FP64, rel err ≤1e-6")
HBM_PEAK_GBS = 8000.0
# Counter-measured HBM bytes (rocprofv3 FETCH_SIZE / WRITE_SIZE passes over
# `bench.py --only traffic-probe`, scaled by a copy kernel of known size):
# written by `bench.py --collect-traffic`, kept as a summary under profiles/.
TRAFFIC_FILE_REL = "profiles/traffic_mgpcg_sphere_pool_256.json"
TRAFFIC_FILE = os.path.join(ROOT, TRAFFIC_FILE_REL)
# Algorithmic bytes per cell of one MGPCG iteration (4 matrix coefficients
# per row, compulsory vector traffic, every level of a V(2,2) cycle, the
# outer iteration's SpMV, updates and dots): FP64 467 B, FP32 inner
# iteration of the mixed solve 234 B, plus 80 B per FP64 refinement step.
ALG_BYTES_PER_CELL = {"fp64": 467.0, "mixed": 234.0, "refine": 80.0}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--levels", type=int, default=6)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--rel-tol", type=float, default=1e-8)
    ap.add_argument("--omega", type=float, default=1.3,
                    help="SOR factor of the red-black smoother (1.3 needs the fewest "
                         "iterations on this workload: 13 against 15 at 1.0 and 1.5)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline / oracle leg")
    ap.add_argument("--only", choices=["compressed", "diffusion", "mgpcg_2d", "traffic-probe"],
                    help="run one secondary leg alone, GPU only (for a per-leg kernel trace), "
                         "and print its JSON; traffic-probe: the program the counter passes of "
                         "--collect-traffic run")
    ap.add_argument("--collect-traffic", action="store_true",
                    help="measure HBM traffic with two rocprofv3 counter passes and write the "
                         "summary to gpurun_out/traffic/")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is a design note, not a measured path)"}))
        sys.exit(2)
    if args.collect_traffic:
        collect_traffic(args)
        return
    if args.only == "traffic-probe":
        traffic_probe(args)
        return
    if args.only:
        import jetgpu as jg

        args.no_cpu = True
        legs = {"compressed": lambda: csr_leg(jg, args.n, args),
                "diffusion": lambda: diffusion_leg(jg, args.n, args),
                "mgpcg_2d": lambda: mg2_leg(jg, args)}
        print(json.dumps({"leg": args.only, **legs[args.only]()["gpu"]}))
        return

    import numpy as np
    import torch

    import jetgpu as jg

    n, L = args.n, args.levels
    solver = jg.Solver(0)
    solver.set_levels([(n >> l,) * 3 for l in range(L)])
    f64 = torch.float64
    u = torch.empty((n, n, n + 1), dtype=f64, device="cuda")
    v = torch.empty((n, n + 1, n), dtype=f64, device="cuda")
    w = torch.empty((n + 1, n, n), dtype=f64, device="cuda")
    bs = torch.empty((n, n, n), dtype=f64, device="cuda")
    fs = torch.empty((n, n, n), dtype=f64, device="cuda")
    torch.cuda.synchronize()
    desc = jg.scene_desc(jg.JG_SCENE_SPHERE_POOL, n, seed=args.seed)
    solver.generate_scene(desc, u, v, w, bs, fs)
    solver.sp_build(u, v, w, bs, fs, 1.0)
    # the flat multi-level build (markers and A on every level, b): timed on a second call
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.sp_build(u, v, w, bs, fs, 1.0)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    fluid_fraction = float(((bs >= 0) & (fs < 0)).double().mean().item())
    del u, v, w, bs, fs
    torch.cuda.empty_cache()
    log(f"bench: system built n={n} levels={L} fluid={fluid_fraction:.3f}")

    params = jg.mg_params(max_levels=L, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                          final_iter=2, sor=args.omega)
    opts = {
        "fp64": jg.gpu_options(precision=jg.JG_FP64, stop_rule=jg.JG_STOP_RELATIVE,
                               rel_tol=args.rel_tol),
        "mixed": jg.gpu_options(precision=jg.JG_MIXED, stop_rule=jg.JG_STOP_RELATIVE,
                                rel_tol=args.rel_tol),
    }
    for _ in range(args.warmup):
        for name in ("fp64", "mixed"):
            solver.mgpcg(500, params, opts[name])
    log("bench: warmup done")

    res = {}
    for name in ("fp64", "mixed"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        infos = [solver.mgpcg(500, params, opts[name]) for _ in range(args.steps)]
        torch.cuda.synchronize()
        wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        x = solver.download_x()
        res[name] = {"ms": wall_ms, "event_ms": float(np.mean([i.solve_ms for i in infos])),
                     "iterations": infos[-1].iterations, "refinements": infos[-1].refinements,
                     "rel_residual": infos[-1].rel_residual, "x": x}
        log(f"bench: {name} {wall_ms:.3f} ms/solve, {infos[-1].iterations} it, "
            f"rel {infos[-1].rel_residual:.2e}")

    # Dominant kernel of the mixed solve: the finest level's smoothing.  A
    # V(2,2) cycle on level 0 is two launches of it (the tuner pairs the two
    # sweeps of each side into one launch wherever that is faster): the
    # pre-smoothing launch (x from zero, the residual written for the
    # restriction) and the post-smoothing launch (the coarse correction added
    # as x is loaded).  jg_solver_time_kernel replays exactly those launches,
    # with the solve's sweep plan and tuning; HIP events on the solver's
    # stream around 50 back-to-back launches.
    kern = {}
    for kname, kid in (("pre0", jg.JG_KERNEL_PRE0), ("post0", jg.JG_KERNEL_POST0),
                       ("sweep", jg.JG_KERNEL_SWEEP), ("dir_spmv", jg.JG_KERNEL_DIR_SPMV),
                       ("residual", jg.JG_KERNEL_RESIDUAL), ("update_xr", jg.JG_KERNEL_UPDATE_XR),
                       ("copy", jg.JG_KERNEL_COPY)):
        ms, nbytes = solver.time_kernel(kid, jg.JG_MIXED, 50)
        kern[kname] = {"avg_ms": ms, "bytes": nbytes, "gbs": nbytes / ms / 1e6}
    traffic = load_traffic(n)
    roof = roofline(kern, traffic)
    n_cells = n ** 3
    out = {
        "metric": METRIC,
        "value": round(res["mixed"]["ms"], 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["mixed"]["ms"], 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"sphere pool {n}^3: 8 spheres r~U(8,32)*n/256, ~60% fluid, "
                               f"{L} levels, V(2,2), 20 coarsest sweeps, SOR {args.omega:g}, "
                               f"rel tol {args.rel_tol:g}",
                   "n": n, "levels": L, "fluid_fraction": round(fluid_fraction, 4)},
        "fp64_ms": round(res["fp64"]["ms"], 4),
        "speedup_mixed_vs_fp64": round(res["fp64"]["ms"] / res["mixed"]["ms"], 3),
        "iterations": {k: res[k]["iterations"] for k in res},
        "refinements_mixed": res["mixed"]["refinements"],
        "rel_residual": {k: res[k]["rel_residual"] for k in res},
        "cells_per_s": {k: n_cells / (res[k]["ms"] * 1e-3) for k in res},
        "kernels_mixed": {k: {"avg_us": round(vv["avg_ms"] * 1e3, 2), "alg_gbs": round(vv["gbs"], 1)}
                          for k, vv in kern.items()},
        "roofline": roof,
        "solve_alg_gbs": {
            "fp64": round(ALG_BYTES_PER_CELL["fp64"] * n_cells * res["fp64"]["iterations"]
                          / (res["fp64"]["ms"] * 1e-3) / 1e9, 1),
            "mixed": round((ALG_BYTES_PER_CELL["mixed"] * res["mixed"]["iterations"]
                            + ALG_BYTES_PER_CELL["refine"] * res["mixed"]["refinements"])
                           * n_cells / (res["mixed"]["ms"] * 1e-3) / 1e9, 1)},
    }
    if traffic and traffic.get("solve"):
        # the metric's HBM GB/s: counter-measured bytes of one whole solve
        # (every kernel of it, from the FETCH_SIZE / WRITE_SIZE passes) over
        # the solve time measured here
        out["solve_hbm"] = {
            k: {"bytes": traffic["solve"][k]["hbm_bytes"],
                "gbs": round(traffic["solve"][k]["hbm_bytes"] / (res[k]["ms"] * 1e-3) / 1e9, 1),
                "frac": round(traffic["solve"][k]["hbm_bytes"] / (res[k]["ms"] * 1e-3) / 1e9
                              / HBM_PEAK_GBS, 4),
                "dispatches": traffic["solve"][k]["dispatches"]}
            for k in ("fp64", "mixed") if k in traffic["solve"]}
        out["solve_hbm"]["source"] = TRAFFIC_FILE_REL

    out["system_build"] = {"gpu_ms": round(build_ms, 3), "levels": L,
                           "what": "GridSinglePhasePressureSolver3 markers and A on every level, "
                                   "b on level 0 (jg_sp_build)"}
    out["projection_step"] = step_leg(jg, solver, n, params, opts, res, args)
    bc = bc_leg(jg, solver, n, args)
    out["constrain_velocity"] = bc["gpu"]
    csr = csr_leg(jg, n, args)
    out["compressed_cg"] = csr["gpu"]
    dif = diffusion_leg(jg, n, args)
    out["diffusion"] = dif["gpu"]
    mg2 = mg2_leg(jg, args)
    out["mgpcg_2d"] = mg2["gpu"]
    small = small_configs_leg(jg, args)
    out["closed_sphere_cg_64"] = small["cg64"]["gpu"]
    out["dam_break_mgpcg_128"] = small["dam128"]["gpu"]
    batch = batch_leg(jg, args)
    out["batch_64_uneven"] = batch["gpu"]
    torus = torus_leg(jg, args)
    out["torus_fractional_512"] = torus["gpu"]
    if not args.no_cpu:
        out.update(cpu_leg(jg, n, L, args, res))
        out["constrain_velocity"].update(bc_cpu_leg(bc, n))
        out["compressed_cg"].update(csr_cpu_leg(csr))
        out["diffusion"].update(diffusion_cpu_leg(dif))
        out["mgpcg_2d"].update(mg2_cpu_leg(mg2))
        c1, c2 = small_configs_cpu_leg(small)
        out["closed_sphere_cg_64"].update(c1)
        out["dam_break_mgpcg_128"].update(c2)
        out["batch_64_uneven"].update(batch_cpu_leg(batch))
        out["torus_fractional_512"].update(torus_cpu_leg(torus))
    print(json.dumps(out))
    solver.close()


def load_traffic(n):
    """The counter-measured traffic summary of this workload, if it is there
    and was measured at this size."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t if t.get("n") == n else None


def roofline(kern, traffic):
    """Roofline of the dominant kernel: the level-0 smoothing launches of one
    V-cycle (pre and post), per launch.  achieved = algorithmic bytes per
    launch / average launch duration (both measured here); traffic = the
    counter-measured HBM bytes per launch of the same two launches."""
    pre, post = kern["pre0"], kern["post0"]
    t_ms = (pre["avg_ms"] + post["avg_ms"]) / 2.0
    alg = (pre["bytes"] + post["bytes"]) / 2.0
    achieved = alg / t_ms / 1e6
    names = None
    hbm = None
    if traffic and traffic.get("launches"):
        tl = traffic["launches"]
        names = " + ".join(tl[k]["kernel"] for k in ("pre0", "post0"))
        hbm = (tl["pre0"]["hbm_bytes"] + tl["post0"]["hbm_bytes"]) / 2.0
    roof = {
        "bound": "hbm",
        "kernel": names or "level-0 smoothing launches of one V(2,2) cycle (pre, post)",
        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": round(hbm) if hbm else None,
        "per": "launch, averaged over the pre- and the post-smoothing launch of level 0",
        "algorithmic_bytes": alg, "avg_us": round(t_ms * 1e3, 2),
        "launches": {k: {"avg_us": round(kern[k]["avg_ms"] * 1e3, 2),
                         "algorithmic_bytes": kern[k]["bytes"],
                         "alg_bytes_per_cell": round(kern[k]["bytes"] / traffic_cells(traffic), 3)
                         if traffic else None}
                     for k in ("pre0", "post0")},
    }
    if hbm:
        roof["measured_gbs"] = round(hbm / t_ms / 1e6, 1)
        roof["measured_frac"] = round(hbm / t_ms / 1e6 / HBM_PEAK_GBS, 4)
        roof["traffic_source"] = TRAFFIC_FILE_REL
    return roof


def traffic_cells(traffic):
    return float(traffic["n"]) ** 3


def traffic_probe(args):
    """The program the counter passes run: the workload's system, warm-up
    solves (tuning), then, each between two marker kernels (a torch fill that
    the library never launches): one mixed solve, one FP64 solve, the pre- and
    the post-smoothing launch of level 0 and a copy kernel (the counters'
    scale), each 1 + 20 times."""
    import torch

    import jetgpu as jg

    n, L = args.n, args.levels
    solver = jg.Solver(0)
    solver.set_levels([(n >> l,) * 3 for l in range(L)])
    f64 = torch.float64
    u = torch.empty((n, n, n + 1), dtype=f64, device="cuda")
    v = torch.empty((n, n + 1, n), dtype=f64, device="cuda")
    w = torch.empty((n + 1, n, n), dtype=f64, device="cuda")
    bs = torch.empty((n, n, n), dtype=f64, device="cuda")
    fs = torch.empty((n, n, n), dtype=f64, device="cuda")
    mark = torch.empty(64, dtype=torch.float32, device="cuda")  # no fill kernel of its own
    torch.cuda.synchronize()
    solver.generate_scene(jg.scene_desc(jg.JG_SCENE_SPHERE_POOL, n, seed=args.seed), u, v, w, bs, fs)
    solver.sp_build(u, v, w, bs, fs, 1.0)
    del u, v, w, bs, fs
    params = jg.mg_params(max_levels=L, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                          final_iter=2, sor=args.omega)
    opts = {k: jg.gpu_options(precision=p, stop_rule=jg.JG_STOP_RELATIVE, rel_tol=args.rel_tol)
            for k, p in (("fp64", jg.JG_FP64), ("mixed", jg.JG_MIXED))}
    for k in ("fp64", "mixed"):
        solver.mgpcg(500, params, opts[k])
    reps = TRAFFIC_REPS

    def marker(i):
        torch.cuda.synchronize()
        mark.fill_(float(i))
        torch.cuda.synchronize()
        log(f"probe: marker {i}")

    marker(1)
    solver.mgpcg(500, params, opts["mixed"])
    marker(2)
    solver.mgpcg(500, params, opts["fp64"])
    marker(3)
    solver.time_kernel(jg.JG_KERNEL_PRE0, jg.JG_MIXED, reps)
    marker(4)
    solver.time_kernel(jg.JG_KERNEL_POST0, jg.JG_MIXED, reps)
    marker(5)
    solver.time_kernel(jg.JG_KERNEL_COPY, jg.JG_MIXED, reps)
    marker(6)
    solver.close()


TRAFFIC_REPS = 20
TRAFFIC_SEGMENTS = ("mixed", "fp64", "pre0", "post0", "copy")


def read_counters(outdir, counter):
    """Per-dispatch (kernel name, value in KiB) of one rocprofv3 --pmc pass,
    in dispatch order (rows of one dispatch summed)."""
    import csv
    import glob

    files = glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError(f"no counter_collection.csv under {outdir}")
    disp = {}
    for fn in files:
        with open(fn, newline="") as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                key = int(row.get("Dispatch_Id") or row.get("Correlation_Id"))
                name, val = disp.get(key, (row["Kernel_Name"], 0.0))
                disp[key] = (name, val + float(row["Counter_Value"]))
    return [disp[k] for k in sorted(disp)]


def segments(disp):
    """Split dispatches at the marker kernels (torch fills)."""
    segs, cur, started = [], [], False
    for name, val in disp:
        if "FillFunctor" in name:
            if started:
                segs.append(cur)
            cur, started = [], True
            continue
        if started:
            cur.append((name, val))
    return segs


def collect_traffic(args):
    """Two rocprofv3 counter passes (FETCH_SIZE, then WRITE_SIZE) over the
    probe program, each a child process; writes the summary JSON (per-launch
    bytes of the level-0 smoothing launches and whole-solve bytes) to
    gpurun_out/traffic/."""
    import subprocess
    from collections import Counter

    out_root = os.path.join(ROOT, "gpurun_out", "traffic")
    os.makedirs(out_root, exist_ok=True)
    env = dict(os.environ, TMPDIR="/tmp")
    raw = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = os.path.join(out_root, counter.lower())
        cmd = ["rocprofv3", "--pmc", counter, "-f", "csv", "-d", d, "-o", "probe", "--",
               sys.executable, os.path.join(ROOT, "bench.py"), "--only", "traffic-probe",
               "--n", str(args.n), "--levels", str(args.levels)]
        log("bench: " + " ".join(cmd))
        subprocess.run(cmd, cwd="/tmp", env=env, check=True, timeout=600)
        segs = segments(read_counters(d, counter))
        if len(segs) != len(TRAFFIC_SEGMENTS):
            raise RuntimeError(f"{counter}: {len(segs)} segments, expected {len(TRAFFIC_SEGMENTS)}")
        raw[counter] = dict(zip(TRAFFIC_SEGMENTS, segs))

    def repeated(seg):
        """The kernel a timing segment launches 1 + TRAFFIC_REPS times (one
        warm launch, then the timed ones), and its average over the timed
        launches."""
        names = Counter(nm for nm, _ in seg)
        name = [nm for nm, c in names.items() if c == TRAFFIC_REPS + 1][0]
        vals = [v for nm, v in seg if nm == name][1:]
        return name, sum(vals) / len(vals)

    n = args.n
    copy_bytes = 4.0 * n ** 3  # k_convert<float, float>: n^3 floats read and written
    cname, cf = repeated(raw["FETCH_SIZE"]["copy"])
    _, cw = repeated(raw["WRITE_SIZE"]["copy"])
    fscale, wscale = copy_bytes / (cf * 1024.0), copy_bytes / (cw * 1024.0)
    res = {"n": n, "levels": args.levels,
           "command": "python3 bench.py --collect-traffic (rocprofv3 --pmc FETCH_SIZE, then "
                      "--pmc WRITE_SIZE, -- python3 bench.py --only traffic-probe)",
           "copy": {"kernel": cname, "bytes_each_way": copy_bytes, "fetch_kib": cf, "write_kib": cw,
                    "fetch_scale": round(fscale, 4), "write_scale": round(wscale, 4)},
           "launches": {}, "solve": {}}
    for k in ("pre0", "post0"):
        name, f = repeated(raw["FETCH_SIZE"][k])
        _, w = repeated(raw["WRITE_SIZE"][k])
        res["launches"][k] = {"kernel": name, "fetch_kib": f, "write_kib": w,
                              "hbm_bytes": round(f * 1024.0 * fscale + w * 1024.0 * wscale)}
    for k in ("mixed", "fp64"):
        f = raw["FETCH_SIZE"][k]
        w = raw["WRITE_SIZE"][k]
        per = {}
        for (nm, fv), (_, wv) in zip(f, w):
            b = fv * 1024.0 * fscale + wv * 1024.0 * wscale
            c, tot = per.get(nm, (0, 0.0))
            per[nm] = (c + 1, tot + b)
        total = sum(t for _, t in per.values())
        top = sorted(per.items(), key=lambda kv: -kv[1][1])[:12]
        res["solve"][k] = {"dispatches": len(f), "hbm_bytes": round(total),
                           "top": [{"kernel": nm, "dispatches": c, "hbm_bytes": round(t)}
                                   for nm, (c, t) in top]}
    path = os.path.join(out_root, os.path.basename(TRAFFIC_FILE))
    with open(path, "w") as f:
        json.dump(res, f, indent=1)
    log(f"bench: traffic summary written to {path}")
    print(json.dumps(res))


BC_DEPTH = 5  # ceil(max CFL) with GridFluidSolver3's default max CFL of 5


def step_leg(jg, solver, n, params, opts, res, args):
    """One pressure-projection step as a simulation issues it, every step from
    the same velocities: GridSinglePhasePressureSolver3's build (markers and A
    on every level, b), the MGPCG solve and the pressure gradient into a second
    set of face arrays.  A rebuilt A invalidates the FP32 copy of the
    matrices, so each mixed step pays its refresh (one pass over the FP64
    rows of every level) inside the timed region."""
    import torch

    f64 = torch.float64
    shapes = [(n, n, n + 1), (n, n + 1, n), (n + 1, n, n)]
    u, v, w = (torch.empty(sh, dtype=f64, device="cuda") for sh in shapes)
    uo, vo, wo = (torch.empty(sh, dtype=f64, device="cuda") for sh in shapes)
    bs, fs = (torch.empty((n, n, n), dtype=f64, device="cuda") for _ in range(2))
    solver.generate_scene(jg.scene_desc(jg.JG_SCENE_SPHERE_POOL, n, seed=args.seed), u, v, w, bs, fs)

    def step(o):
        solver.sp_build(u, v, w, bs, fs, 1.0)
        info = solver.mgpcg(500, params, o)
        solver.sp_apply_gradient(u, v, w, uo, vo, wo, 1.0)
        return info

    out = {"what": "jg_sp_build + jg_solver_mgpcg + jg_sp_apply_gradient per step, the FP32 "
                   "matrix copy refreshed every mixed step"}
    for name in ("fp64", "mixed"):
        step(opts[name])
        reps = max(args.steps, 3)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        infos = [step(opts[name]) for _ in range(reps)]
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        out[name] = {"ms": round(ms, 4), "iterations": infos[-1].iterations,
                     "rel_residual": infos[-1].rel_residual,
                     "solve_only_ms": round(res[name]["ms"], 4),
                     "rest_ms": round(ms - res[name]["ms"], 4)}
        log(f"bench: projection step {name} {ms:.3f} ms")
    del u, v, w, uo, vo, wo, bs, fs
    torch.cuda.empty_cache()
    return out


def bc_leg(jg, solver, n, args):
    """The step after the projection: GridBlockedBoundaryConditionSolver3::
    constrainVelocity on the same scene's faces, its spheres as a resting
    collider (cell-centre SDF), extrapolation depth 5, all walls closed."""
    import torch

    f64 = torch.float64
    shapes = [(n, n, n + 1), (n, n + 1, n), (n + 1, n, n)]
    u, v, w = (torch.empty(sh, dtype=f64, device="cuda") for sh in shapes)
    bs = torch.empty((n, n, n), dtype=f64, device="cuda")
    fs = torch.empty((n, n, n), dtype=f64, device="cuda")
    desc = jg.scene_desc(jg.JG_SCENE_SPHERE_POOL, n, seed=args.seed)
    solver.generate_scene(desc, u, v, w, bs, fs)
    torch.cuda.synchronize()
    vel0 = [a.cpu().numpy() for a in (u, v, w)] if not args.no_cpu else None
    h, o = (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    call = lambda: solver.constrain_velocity(u, v, w, bs, h, o, None, BC_DEPTH, jg.JG_DIR_ALL, True)
    call()  # first call: scratch allocation, and the result checked against the oracle
    first = [a.cpu().numpy() for a in (u, v, w)] if vel0 is not None else None
    reps = 10
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    faces = sum(a.numel() for a in (u, v, w))
    log(f"bench: constrain_velocity {ms:.3f} ms ({faces} faces)")
    bsh = bs.cpu().numpy() if vel0 is not None else None
    del u, v, w, bs, fs
    torch.cuda.empty_cache()
    return {"gpu": {"gpu_ms": round(ms, 4), "faces": faces, "depth": BC_DEPTH, "blocked": True,
                    "collider": "the scene's spheres, resting (cell-centre SDF)"},
            "vel0": vel0, "first": first, "sdf": bsh}


def csr_leg(jg, n, args):
    """The compressed path on the same scene: GridSinglePhasePressureSolver3's
    compressed build on the device (fluid rows only) and FdmCgSolver3::
    solveCompressed (Jet's CG, no preconditioner) to ||r|| <= 1e-8 ||b||."""
    import torch

    f64 = torch.float64
    shapes = [(n, n, n + 1), (n, n + 1, n), (n + 1, n, n), (n, n, n), (n, n, n)]
    u, v, w, bs, fs = (torch.empty(sh, dtype=f64, device="cuda") for sh in shapes)
    s = jg.Solver(0)
    s.set_levels([(n, n, n)])
    s.generate_scene(jg.scene_desc(jg.JG_SCENE_SPHERE_POOL, n, seed=args.seed), u, v, w, bs, fs)
    torch.cuda.synchronize()
    s.sp_build_compressed(u, v, w, bs, fs, 1.0)  # first build: allocation
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rows, nnz = s.sp_build_compressed(u, v, w, bs, fs, 1.0)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    b = torch.as_tensor(jg._DeviceView(jg.lib().jg_csr_array(s.h, 1), rows, "<f8"), device="cuda")
    tol = 1e-8 * float(torch.linalg.norm(b).item())
    t0 = time.perf_counter()
    info = s.csr_solve(20000, tol)
    solve_ms = (time.perf_counter() - t0) * 1e3
    it = max(1, info.iterations)
    # algorithmic bytes of one CG iteration: the fused direction + product
    # (values and 32-bit columns per entry, row pointer, z / d in, d / q out),
    # x and r update (4 reads, 2 writes), r.r (r = z without preconditioner)
    per_it = 12 * nnz + (4 + 32) * rows + 48 * rows + 8 * rows
    log(f"bench: compressed rows={rows} nnz={nnz} build {build_ms:.2f} ms, CG {it} it "
        f"{solve_ms:.1f} ms, rel {info.rel_residual:.1e}")
    sample = None
    if not args.no_cpu:  # a bounded sample of the same system for the CPU leg
        val, rp, ci, bh, _ = s.csr_download(rows, nnz)
        sample = (val, rp, ci, bh)
    s.close()
    del u, v, w, bs, fs, b
    torch.cuda.empty_cache()
    return {"gpu": {"rows": rows, "nnz": nnz, "build_ms": round(build_ms, 3),
                    "solve_ms": round(solve_ms, 2), "iterations": info.iterations,
                    "rel_residual": info.rel_residual,
                    "ms_per_iteration": round(solve_ms / it, 4),
                    "alg_bytes_per_iteration": per_it,
                    "alg_gbs": round(per_it * it / (solve_ms * 1e-3) / 1e9, 1)},
            "sample": sample}


DIF_DT, DIF_MU = 1.0 / 60.0, 0.05  # c = dt mu / h^2 = 54.6 at h = 1/256
DIF_LEVELS = 6
DIF_MG = dict(restrict_iter=2, correct_iter=2, coarsest_iter=20, final_iter=2)


def diffusion_leg(jg, n, args):
    """GridBackwardEulerDiffusionSolver3 on the scene's n^3 cell-centred grid
    (boundary and fluid SDFs of the sphere pool, Neumann at the colliders):
    the system built on the device and solved to ||r|| <= 1e-8 ||b|| by
    Jet's CG and by the flat-entry MGPCG (Jet's level sizing, 6 levels, the
    coarse matrices built on the device from A), FP64 and mixed; the field a
    smooth pattern."""
    import torch

    f64 = torch.float64
    shapes = [(n, n, n + 1), (n, n + 1, n), (n + 1, n, n), (n, n, n), (n, n, n)]
    u, v, w, bs, fs = (torch.empty(sh, dtype=f64, device="cuda") for sh in shapes)
    s = jg.Solver(0)
    s.set_levels([(n, n, n)])
    s.generate_scene(jg.scene_desc(jg.JG_SCENE_SPHERE_POOL, n, seed=args.seed), u, v, w, bs, fs)
    g = torch.arange(n, dtype=f64, device="cuda") / n
    f = (torch.sin(6.0 * g)[:, None, None] * torch.cos(4.0 * g)[None, :, None]
         + g[None, None, :]).contiguous()
    h = 1.0 / n
    torch.cuda.synchronize()
    s.diffusion_build(f, bs, fs, DIF_DT, DIF_MU, h)  # first build: allocation
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.diffusion_build(f, bs, fs, DIF_DT, DIF_MU, h)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    b = torch.as_tensor(jg._DeviceView(s.rhs_ptr(), n ** 3, "<f8"), device="cuda")
    tol = 1e-8 * float(torch.linalg.norm(b).item())
    del b
    t0 = time.perf_counter()
    info = s.cg(2000, tol)
    solve_ms = (time.perf_counter() - t0) * 1e3
    it = max(1, info.iterations)
    log(f"bench: diffusion {n}^3 build {build_ms:.2f} ms, CG {it} it {solve_ms:.2f} ms")
    xs = {"cg": s.download_x()} if not args.no_cpu else {}
    gpu = {"cells": n ** 3, "dt": DIF_DT, "mu": DIF_MU, "h": h,
           "build_ms": round(build_ms, 3), "solve_ms": round(solve_ms, 3),
           "iterations": info.iterations, "rel_residual": info.rel_residual,
           "ms_per_iteration": round(solve_ms / it, 4)}

    # the flat-entry MGPCG on the same system
    L = DIF_LEVELS
    s.set_levels([(n >> l,) * 3 for l in range(L)])
    s.diffusion_build(f, bs, fs, DIF_DT, DIF_MU, h)
    s.coarsen()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.diffusion_build(f, bs, fs, DIF_DT, DIF_MU, h)
    s.coarsen()
    torch.cuda.synchronize()
    coarsen_ms = (time.perf_counter() - t0) * 1e3 - build_ms
    p = jg.mg_params(max_levels=L, sor=args.omega, **DIF_MG)
    flat = {"levels": L, "schedule": "V(2,2), 20 coarsest sweeps, SOR %g" % args.omega,
            "build_and_coarsen_ms": round(build_ms + coarsen_ms, 3)}
    for name, prec in (("fp64", jg.JG_FP64), ("mixed", jg.JG_MIXED)):
        o = jg.gpu_options(precision=prec, stop_rule=jg.JG_STOP_RELATIVE, rel_tol=1e-8)
        s.mgpcg(500, p, o)  # tuning
        reps = 3
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fi = s.mgpcg(500, p, o)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        flat[name] = {"solve_ms": round(ms, 3), "iterations": fi.iterations,
                      "rel_residual": fi.rel_residual}
        if not args.no_cpu:
            xs[f"flat_{name}"] = s.download_x()
        log(f"bench: diffusion flat MGPCG {name} {ms:.2f} ms, {fi.iterations} it")
    flat["speedup_fp64_vs_cg"] = round(solve_ms / flat["fp64"]["solve_ms"], 2)
    flat["speedup_mixed_vs_cg"] = round(solve_ms / flat["mixed"]["solve_ms"], 2)
    gpu["flat_mgpcg"] = flat
    sample = None
    if not args.no_cpu:
        sample = tuple(a.cpu().numpy() for a in (f, bs, fs))
    s.close()
    del u, v, w, bs, fs, f
    torch.cuda.empty_cache()
    return {"gpu": gpu, "sample": sample, "h": h, "x": xs, "omega": args.omega}


def diffusion_cpu_leg(dif):
    """The oracle's diffusion build and 5 CG iterations on the same input
    (one core), its flat-entry MGPCG solve (the same aggregated hierarchy,
    OpenMP) to 1e-12, and the error of every GPU solution against that."""
    import numpy as np

    from oracle.jet_oracle import MgParams, Oracle

    f, bs, fs = dif["sample"]
    o = Oracle()
    t0 = time.perf_counter()
    A, b = o.diffusion_build(f, bs, fs, DIF_DT, DIF_MU, (dif["h"],) * 3)
    build_ms = (time.perf_counter() - t0) * 1e3
    iters = 5
    t0 = time.perf_counter()
    o.cg_solve(A, b, iters, 0.0)
    ms = (time.perf_counter() - t0) * 1e3
    log(f"bench: diffusion oracle build {build_ms:.0f} ms, {ms / iters:.1f} ms per CG iteration")
    out = {"cpu_build_ms": round(build_ms, 1), "cpu_ms_per_iteration": round(ms / iters, 2),
           "cpu_cores": 1, "cpu_kind": "port",
           "cpu_sample": f"the oracle's build and {iters} CG iterations on the same grid"}
    t0 = time.perf_counter()
    hier = o.mg_flat_hierarchy(A, b, DIF_LEVELS)
    ok, it, rel = o.mgpcg_clean_solve(hier, MgParams.make(max_levels=DIF_LEVELS, sor=dif["omega"],
                                                          **DIF_MG), 200, 1e-12)
    mg_ms = (time.perf_counter() - t0) * 1e3
    ref = hier.x[0]
    nr = np.linalg.norm(ref)
    out["flat_mgpcg_cpu"] = {"ms": round(mg_ms, 1), "iterations": it, "rel_residual": rel,
                             "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                             "kind": "port"}
    out["err_vs_oracle"] = {k: float(np.linalg.norm(x - ref) / nr) for k, x in dif["x"].items()}
    log(f"bench: diffusion oracle flat MGPCG {mg_ms:.0f} ms ({it} it), errors {out['err_vs_oracle']}")
    return out


# 4096^2: as many cells as the 256^3 workload, in one plane (the stencil
# kernels' per-tile reduction partials bound a plane at about 2^24 cells)
MG2_N, MG2_LEVELS = 4096, 8


def poisson2_hierarchy(n, levels, seed):
    """A 2-D pressure-like hierarchy (FdmMgLinearSystem2 as one-plane levels):
    5-point Laplacian scaled by (1/h_l)^2 with h_l = 2^l, closed walls and an
    open (Dirichlet) top edge, b random on the finest level."""
    import numpy as np

    A, sizes = [], []
    for l in range(levels):
        m = n >> l
        s = 0.25 ** l
        j = np.arange(m)[:, None]
        i = np.arange(m)[None, :]
        cnt = ((i > 0).astype(float) + (i < m - 1) + (j > 0) + (j < m - 1) + (j == m - 1))
        a = np.zeros((1, m, m, 4))
        a[0, :, :, 0] = cnt * s
        a[0, :, :, 1] = np.where(i < m - 1, -s, 0.0)
        a[0, :, :, 2] = np.where(j < m - 1, -s, 0.0)
        A.append(a)
        sizes.append((m, m, 1))
    b = np.random.default_rng(seed).uniform(-1.0, 1.0, (1, n, n))
    return A, b, sizes


def mg2_leg(jg, args):
    """FdmMgpcgSolver2 on a 4096^2 pressure-like system (8 levels, V(2,2), 20
    coarsest sweeps), FP64 and mixed, to ||r|| <= 1e-8 ||b||."""
    import torch

    A, b, sizes = poisson2_hierarchy(MG2_N, MG2_LEVELS, args.seed)
    s = jg.Solver(0)
    s.set_levels(sizes)
    for l, a in enumerate(A):
        s.upload(l, a, b if l == 0 else None)
    p = jg.mg_params(max_levels=MG2_LEVELS, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                     final_iter=2, sor=args.omega)
    out = {}
    for name, prec in (("fp64", jg.JG_FP64), ("mixed", jg.JG_MIXED)):
        o = jg.gpu_options(precision=prec, stop_rule=jg.JG_STOP_RELATIVE, rel_tol=1e-8)
        s.mgpcg(200, p, o)  # warm-up: tuning and FP32 copies
        reps = 3
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            info = s.mgpcg(200, p, o)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        out[name] = {"solve_ms": round(ms, 3), "iterations": info.iterations,
                     "rel_residual": info.rel_residual}
        log(f"bench: 2-D MGPCG {MG2_N}^2 {name}: {ms:.2f} ms, {info.iterations} it")
    s.close()
    torch.cuda.empty_cache()
    gpu = {"grid": [MG2_N, MG2_N], "levels": MG2_LEVELS, **out,
           "speedup_mixed_vs_fp64": round(out["fp64"]["solve_ms"] / out["mixed"]["solve_ms"], 3),
           "ms_per_iteration_fp64": round(out["fp64"]["solve_ms"]
                                          / max(1, out["fp64"]["iterations"]), 4)}
    return {"gpu": gpu, "system": (A, b), "omega": args.omega}


def mg2_cpu_leg(mg2):
    """The oracle's FP64 clean-mode MGPCG on the same 2-D hierarchy, 5
    iterations: the CPU time per iteration beside the GPU's."""
    import numpy as np

    from oracle.jet_oracle import MgParams, MgSystem, Oracle


    A, b = mg2["system"]
    sys_ = MgSystem(A=list(A), x=[np.zeros(a.shape[:3]) for a in A],
                    b=[b] + [np.zeros(a.shape[:3]) for a in A[1:]])
    p = MgParams.make(max_levels=len(A), restrict_iter=2, correct_iter=2, coarsest_iter=20,
                      final_iter=2, sor=mg2["omega"])
    iters = 5
    t0 = time.perf_counter()
    Oracle().mgpcg_clean_solve(sys_, p, iters, 0.0)
    ms = (time.perf_counter() - t0) * 1e3
    log(f"bench: 2-D MGPCG oracle {ms / iters:.0f} ms per iteration")
    return {"cpu_ms_per_iteration": round(ms / iters, 1), "cpu_cores": 1, "cpu_kind": "port",
            "cpu_sample": f"{iters} iterations of the oracle's FP64 MGPCG on the same hierarchy"}


def csr_cpu_leg(csr):
    """The oracle's compressed CG (serial C) on the same rows, 20 iterations:
    the CPU time per iteration beside the GPU's."""
    from oracle.jet_oracle import Oracle

    val, rp, ci, b = csr["sample"]
    iters = 20
    t0 = time.perf_counter()
    Oracle().csr_cg_solve(val, rp.astype("uint64"), ci.astype("uint64"), b, iters, 0.0)
    ms = (time.perf_counter() - t0) * 1e3
    log(f"bench: compressed CG oracle {ms / iters:.1f} ms per iteration")
    return {"cpu_ms_per_iteration": round(ms / iters, 2), "cpu_cores": 1, "cpu_kind": "port",
            "cpu_sample": f"{iters} iterations of the oracle's compressed CG on the same system"}


def bc_cpu_leg(bc, n):
    """The oracle's constrainVelocity on the host (one core) on the same
    input: the CPU time beside the GPU's, and bitwise equality of the first
    GPU call's result."""
    import numpy as np

    from oracle.jet_oracle import DIR_ALL, Oracle

    t0 = time.perf_counter()
    got = Oracle().constrain_velocity(*bc["vel0"], bc["sdf"], (1.0,) * 3, (0.0,) * 3, None,
                                      BC_DEPTH, DIR_ALL, True)
    cpu_ms = (time.perf_counter() - t0) * 1e3
    same = all(np.array_equal(a, b) for a, b in zip(got, bc["first"]))
    log(f"bench: constrain_velocity oracle {cpu_ms:.0f} ms, GPU result bitwise equal: {same}")
    return {"cpu_ms": round(cpu_ms, 1), "cpu_cores": 1, "cpu_kind": "port",
            "gpu_bitwise_equal_to_oracle": bool(same)}


def _device_scene(jg, s, kind, n, seed):
    """Generate a scene on the device and build its system on every level of
    the solver's hierarchy; returns the host copies of its inputs."""
    import torch

    f64 = torch.float64
    shapes = [(n, n, n + 1), (n, n + 1, n), (n + 1, n, n), (n, n, n), (n, n, n)]
    u, v, w, bs, fs = (torch.empty(sh, dtype=f64, device="cuda") for sh in shapes)
    s.generate_scene(jg.scene_desc(kind, n, seed=seed), u, v, w, bs, fs)
    s.sp_build(u, v, w, bs, fs, 1.0)
    torch.cuda.synchronize()
    host = {k: a.cpu().numpy() for k, a in zip(("u", "v", "w", "bsdf", "fsdf"), (u, v, w, bs, fs))}
    del u, v, w, bs, fs
    torch.cuda.empty_cache()
    return host


def _timed(fn, reps):
    import torch

    fn()  # warm-up (allocation, tuning, FP32 copies)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        info = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps, info


def small_configs_leg(jg, args):
    """The two smaller configurations of the workload set, on the GPU:
    - 64^3 closed sphere (all fluid, a solid sphere of radius 16 at the
      centre, solid walls: all-Neumann) solved by Jet's FdmCgSolver3(2000,
      1e-8), i.e. CG to the absolute ||r|| <= 1e-8 of Jet's rule;
    - 128^3 dam break (fluid in x < 64, y < 51.2, air above, gravity on v)
      by FP64 MGPCG with 5 levels and V(2,2), 20 coarsest sweeps, to
      ||b - A x|| <= 1e-8 ||b|| (and mixed precision beside it)."""
    out = {}
    s = jg.Solver(0)
    s.set_levels([(64, 64, 64)])
    host = _device_scene(jg, s, jg.JG_SCENE_CLOSED_SPHERE, 64, args.seed)
    ms, info = _timed(lambda: s.cg(2000, 1e-8), 3)
    out["cg64"] = {"gpu": {"workload": "64^3 closed sphere r=16, all-Neumann, FdmCgSolver3(2000, 1e-8)",
                           "gpu_ms": round(ms, 3), "iterations": info.iterations,
                           "last_residual": info.last_residual, "rel_residual": info.rel_residual},
                   "scene": host, "x": s.download_x()}
    log(f"bench: 64^3 closed sphere CG {ms:.2f} ms, {info.iterations} it")
    s.close()
    L = 5
    s = jg.Solver(0)
    s.set_levels([(128 >> l,) * 3 for l in range(L)])
    host = _device_scene(jg, s, jg.JG_SCENE_DAM_BREAK, 128, args.seed)
    p = jg.mg_params(max_levels=L, restrict_iter=2, correct_iter=2, coarsest_iter=20, final_iter=2,
                     sor=args.omega)
    gpu = {"workload": "128^3 dam break, 5 levels, V(2,2), 20 coarsest sweeps, rel tol 1e-8"}
    xs = {}
    for name, prec in (("fp64", jg.JG_FP64), ("mixed", jg.JG_MIXED)):
        o = jg.gpu_options(precision=prec, stop_rule=jg.JG_STOP_RELATIVE, rel_tol=1e-8)
        ms, info = _timed(lambda: s.mgpcg(500, p, o), 5)
        gpu[name] = {"gpu_ms": round(ms, 3), "iterations": info.iterations,
                     "rel_residual": info.rel_residual}
        xs[name] = s.download_x()
        log(f"bench: 128^3 dam break {name} {ms:.2f} ms, {info.iterations} it")
    gpu["speedup_mixed_vs_fp64"] = round(gpu["fp64"]["gpu_ms"] / gpu["mixed"]["gpu_ms"], 3)
    out["dam128"] = {"gpu": gpu, "scene": host, "x": xs, "levels": L, "omega": args.omega}
    s.close()
    return out


BATCH_PARAMS = dict(restrict_iter=2, correct_iter=2, coarsest_iter=20, final_iter=2)


def _batch_levels(n, max_levels=5):
    lv = 1
    while lv < max_levels and n % (1 << lv) == 0 and (n >> lv) >= 2:
        lv += 1
    return lv


def batch_leg(jg, args):
    """64 independent pressure systems of uneven sizes (edges drawn from
    {32, 48, 64, 96, 128}, 1-4 spheres and a free surface each, built on the
    device), solved together by one replayed graph with per-system stop
    flags, FP64 and mixed, to ||b - A x|| <= 1e-8 ||b|| per system."""
    import numpy as np
    import torch

    rng = np.random.default_rng(2024)
    sizes = [int(n) for n in rng.choice([32, 48, 64, 96, 128], 64)]
    seeds = [1000 + i for i in range(64)]
    levels = [_batch_levels(n) for n in sizes]
    b = jg.Batch([(n, n, n) for n in sizes], levels)
    f64 = torch.float64
    for sys_, n, seed in zip(b.systems, sizes, seeds):
        u, v, w = (torch.empty(sh, dtype=f64, device="cuda")
                   for sh in ((n, n, n + 1), (n, n + 1, n), (n + 1, n, n)))
        bs, fs = (torch.empty((n, n, n), dtype=f64, device="cuda") for _ in range(2))
        sys_.generate_scene(jg.scene_desc(jg.JG_SCENE_BATCH_MEMBER, n, seed=seed), u, v, w, bs, fs)
        sys_.sp_build(u, v, w, bs, fs, 1.0)
    torch.cuda.synchronize()
    p = jg.mg_params(max_levels=5, sor=args.omega, **BATCH_PARAMS)
    gpu = {"systems": len(sizes), "cells": int(sum(n ** 3 for n in sizes)),
           "edges": {str(n): sizes.count(n) for n in sorted(set(sizes))}}
    xs = {}
    for name, prec in (("fp64", jg.JG_FP64), ("mixed", jg.JG_MIXED)):
        o = jg.gpu_options(precision=prec, stop_rule=jg.JG_STOP_RELATIVE, rel_tol=1e-8)
        b.mgpcg(500, p, o)  # tuning and graph capture
        reps = 3
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            infos, _ = b.mgpcg(500, p, o)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        gpu[name] = {"ms": round(ms, 3), "max_iterations": max(i.iterations for i in infos),
                     "max_rel_residual": max(i.rel_residual for i in infos),
                     "cells_per_s": gpu["cells"] / (ms * 1e-3)}
        xs[name] = [s.download_x() for s in b.systems]
        log(f"bench: batch of 64 {name} {ms:.2f} ms, max {gpu[name]['max_iterations']} it")
    gpu["speedup_mixed_vs_fp64"] = round(gpu["fp64"]["ms"] / gpu["mixed"]["ms"], 3)
    b.close()
    torch.cuda.empty_cache()
    # one member of each size is checked against the oracle by the CPU leg
    first = {}
    for i, n in enumerate(sizes):
        first.setdefault(n, i)
    sample = {i: (sizes[i], seeds[i], levels[i], xs["fp64"][i], xs["mixed"][i])
              for i in first.values()}
    return {"gpu": gpu, "sample": sample, "omega": args.omega}


TORUS_N, TORUS_LEVELS, TORUS_SEED = 512, 7, 11
TORUS_MG = dict(restrict_iter=2, correct_iter=2, coarsest_iter=20, final_iter=2, sor=1.3)


def torus_leg(jg, args):
    """The variational workload at 512^3: GridFractionalSinglePhasePressureSolver3's
    weights and system on all 7 levels built on the device from the torus
    pool (torus R = 150, r = 50 cells, liquid below 0.35 n), solved by MGPCG
    in FP64 and mixed precision to ||b - A x|| <= 1e-8 ||b||."""
    import torch

    n, L = TORUS_N, TORUS_LEVELS
    d = jg.scene_desc(jg.JG_SCENE_TORUS_POOL, n, seed=TORUS_SEED)
    f64 = torch.float64
    s = jg.Solver(0)
    s.set_levels([(n >> l,) * 3 for l in range(L)])
    u = torch.empty((n, n, n + 1), dtype=f64, device="cuda")
    v = torch.empty((n, n + 1, n), dtype=f64, device="cuda")
    w = torch.empty((n + 1, n, n), dtype=f64, device="cuda")
    bs, fs = (torch.empty((n, n, n), dtype=f64, device="cuda") for _ in range(2))
    vs = torch.empty((n + 1, n + 1, n + 1), dtype=f64, device="cuda")
    s.generate_scene(d, u, v, w, bs, fs, vs)
    s.fr_build(u, v, w, vs, fs, 1.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.fr_build(u, v, w, vs, fs, 1.0)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    fluid = float(((fs < 0) & (bs >= 0)).double().mean().item())
    del u, v, w, bs, fs, vs
    torch.cuda.empty_cache()
    p = jg.mg_params(max_levels=L, **TORUS_MG)
    gpu = {"cells": n ** 3, "levels": L, "fluid_fraction": round(fluid, 4),
           "build_ms": round(build_ms, 3)}
    xs = {}
    for name, prec in (("fp64", jg.JG_FP64), ("mixed", jg.JG_MIXED)):
        o = jg.gpu_options(precision=prec, stop_rule=jg.JG_STOP_RELATIVE, rel_tol=1e-8)
        s.mgpcg(300, p, o)  # tuning
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = s.mgpcg(300, p, o)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        gpu[name] = {"ms": round(ms, 3), "iterations": info.iterations,
                     "rel_residual": info.rel_residual}
        if not args.no_cpu:
            xs[name] = s.download_x()
        log(f"bench: torus 512^3 {name} {ms:.2f} ms, {info.iterations} it")
    gpu["speedup_mixed_vs_fp64"] = round(gpu["fp64"]["ms"] / gpu["mixed"]["ms"], 3)
    s.close()
    torch.cuda.empty_cache()
    return {"gpu": gpu, "x": xs}


def torus_cpu_leg(torus):
    """The oracle's weights, system and FP64 MGPCG (the GPU's algorithm, to
    1e-11) of the same generated scene on the host cores, and the error of
    both GPU solutions against it."""
    import numpy as np

    import jetgpu as jg
    from oracle.jet_oracle import MgParams, Oracle

    o = Oracle()
    L = TORUS_LEVELS
    t0 = time.perf_counter()
    sc = jg.generate_scene_host(jg.scene_desc(jg.JG_SCENE_TORUS_POOL, TORUS_N, seed=TORUS_SEED),
                                vertex=True)
    _, sys_ = o.fr_build_mg(sc["u"], sc["v"], sc["w"], (1.0,) * 3, (0.0,) * 3, sc["vertex"],
                            sc["fsdf"], L)
    del sc
    build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    ok, it, rel = o.mgpcg_clean_solve(sys_, MgParams.make(max_levels=L, **TORUS_MG), 300, 1e-11)
    solve_ms = (time.perf_counter() - t0) * 1e3
    ref = sys_.x[0]
    nr = np.linalg.norm(ref)
    errs = {k: float(np.linalg.norm(x - ref) / nr) for k, x in torus["x"].items()}
    log(f"bench: torus oracle build {build_s:.1f} s, solve {solve_ms:.0f} ms ({it} it), "
        f"errors {errs}")
    return {"err_vs_oracle": errs, "oracle_rel_residual": rel,
            "cpu_oracle": {"ms": round(solve_ms, 1), "iterations": it, "rel_tol": 1e-11,
                           "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                           "kind": "port",
                           "sample": "the oracle's FP64 MGPCG (the GPU's algorithm) on the same "
                                     "system to 1e-11 relative residual; its build "
                                     f"({build_s:.1f} s: numpy scene, C weights and system) is "
                                     "not timed"}}


def batch_cpu_leg(batch):
    """The oracle's FP64 solve (1e-12) of one member of each size, and the
    error of the batched FP64 and mixed solutions against it."""
    import numpy as np

    import jetgpu as jg
    from oracle.jet_oracle import MgParams, Oracle

    o = Oracle()
    errs = {}
    for i, (n, seed, lv, x64, x32) in sorted(batch["sample"].items()):
        sc = jg.generate_scene_host(jg.scene_desc(jg.JG_SCENE_BATCH_MEMBER, n, seed=seed))
        _, sys_ = o.sp_build_mg(sc["u"], sc["v"], sc["w"], (1.0,) * 3, (0.0,) * 3, sc["bsdf"],
                                sc["fsdf"], lv)
        o.mgpcg_clean_solve(sys_, MgParams.make(max_levels=lv, sor=batch["omega"], **BATCH_PARAMS),
                            300, 1e-12)
        ref = sys_.x[0]
        nr = np.linalg.norm(ref)
        errs[f"member {i} ({n}^3)"] = {"fp64": float(np.linalg.norm(x64 - ref) / nr),
                                       "mixed": float(np.linalg.norm(x32 - ref) / nr)}
    log(f"bench: batch members vs oracle {errs}")
    return {"err_vs_oracle": errs}


def small_configs_cpu_leg(small):
    """CPU baselines of the two smaller configurations, on the host cores with
    Jet's own solvers compiled from its sources (oracle/_ref/libjetref.so,
    OpenMP), and the FP64-oracle error of each GPU solution against the
    oracle's 1e-12 solve of the same system."""
    import numpy as np

    from oracle.jet_oracle import MgParams, Oracle, Ref

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = Oracle()
    try:
        ref = Ref()
    except OSError:
        ref = None
    c1, c2 = {}, {}

    # 64^3 closed sphere: Jet's FdmCgSolver3(2000, 1e-8), the whole solve
    sc = small["cg64"]["scene"]
    markers, sys_ = o.sp_build_mg(sc["u"], sc["v"], sc["w"], (1.0,) * 3, (0.0,) * 3, sc["bsdf"],
                                  sc["fsdf"], 4)
    A, b = sys_.A[0], sys_.b[0]
    solve = (lambda: ref.cg_solve(A, b, 2000, 1e-8)) if ref else (lambda: o.cg_solve(A, b, 2000, 1e-8))
    t0 = time.perf_counter()
    _, it, res, _ = solve()
    cpu_ms = (time.perf_counter() - t0) * 1e3
    o.mgpcg_clean_solve(sys_, MgParams.make(max_levels=4, restrict_iter=2, correct_iter=2,
                                            coarsest_iter=20, final_iter=2), 200, 1e-12)
    # all-Neumann: the solution is defined up to a constant on the fluid
    # cells (the solid sphere's cells are identity rows, p = 0 there)
    fluid = markers[0] == 0
    p_ref = sys_.x[0][fluid] - sys_.x[0][fluid].mean()
    x = small["cg64"]["x"][fluid]
    c1.update({"cpu_baseline": {"value": round(cpu_ms, 2), "unit": "ms", "cores": threads,
                                "kind": "reference" if ref else "port",
                                "sample": f"the whole solve by Jet's FdmCgSolver3(2000, 1e-8) "
                                          f"compiled from its sources, {it} iterations, "
                                          f"lastResidual {res:.2e}"},
               "err_vs_oracle_mean_free": float(np.linalg.norm((x - x.mean()) - p_ref)
                                                / np.linalg.norm(p_ref))})
    log(f"bench: 64^3 CG on the CPU {cpu_ms:.0f} ms, {it} it")

    # 128^3 dam break: the oracle's FP64 MGPCG to 1e-8 (the GPU's algorithm on
    # the host) and a bounded sample of Jet's own FdmMgpcgSolver3 (red-black
    # smoothing, the same V(2,2)): its time per iteration
    sc = small["dam128"]["scene"]
    L = small["dam128"]["levels"]
    _, sys_ = o.sp_build_mg(sc["u"], sc["v"], sc["w"], (1.0,) * 3, (0.0,) * 3, sc["bsdf"],
                            sc["fsdf"], L)
    p = MgParams.make(max_levels=L, restrict_iter=2, correct_iter=2, coarsest_iter=20, final_iter=2,
                      sor=small["dam128"]["omega"])
    t0 = time.perf_counter()
    _, it_cpu, rel_cpu = o.mgpcg_clean_solve(sys_, p, 200, 1e-8)
    cpu_ms = (time.perf_counter() - t0) * 1e3
    o.mgpcg_clean_solve(sys_, p, 200, 1e-12)
    p_ref = sys_.x[0].copy()
    c2["cpu_baseline"] = {"value": round(cpu_ms, 1), "unit": "ms", "cores": threads, "kind": "port",
                          "sample": f"the whole solve by the oracle's FP64 MGPCG (the GPU's "
                                    f"algorithm), {it_cpu} iterations to rel residual {rel_cpu:.1e}"}
    if ref is not None:
        iters = 5
        pj = MgParams.make(max_levels=L, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                           final_iter=2, max_tol=0.0, red_black=1)
        t0 = time.perf_counter()
        ref.mgpcg_solve(sys_, pj, iters)
        jet_ms = (time.perf_counter() - t0) * 1e3
        c2["cpu_reference_ms_per_iteration"] = round(jet_ms / iters, 1)
        c2["cpu_reference_sample"] = (f"{iters} iterations of Jet's own FdmMgpcgSolver3 (red-black, "
                                      f"V(2,2), compiled from its sources, OpenMP, {threads} threads)")
    c2["err_vs_oracle"] = {k: float(np.linalg.norm(x - p_ref) / np.linalg.norm(p_ref))
                           for k, x in small["dam128"]["x"].items()}
    log(f"bench: 128^3 dam break oracle {cpu_ms:.0f} ms, errors {c2['err_vs_oracle']}")
    return c1, c2


def cpu_leg(jg, n, L, args, res):
    """Oracle FP64 solve of the same system on the host cores: the CPU
    baseline (timed to the same 1e-8 relative residual) and p_ref (a second
    solve to 1e-12) for the error of both GPU solutions."""
    import numpy as np

    from oracle.jet_oracle import MgParams, Oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    oracle = Oracle()
    scene = jg.generate_scene_host(jg.scene_desc(jg.JG_SCENE_SPHERE_POOL, n, seed=args.seed))
    t0 = time.perf_counter()
    markers, sys_ = oracle.sp_build_mg(scene["u"], scene["v"], scene["w"], (1.0,) * 3, (0.0,) * 3,
                                       scene["bsdf"], scene["fsdf"], L)
    build_cpu_ms = (time.perf_counter() - t0) * 1e3
    log(f"bench: oracle system build {build_cpu_ms:.0f} ms")
    del scene
    p = MgParams.make(max_levels=L, restrict_iter=2, correct_iter=2, coarsest_iter=20, final_iter=2,
                      sor=args.omega)
    t0 = time.perf_counter()
    ok, it_cpu, rel_cpu = oracle.mgpcg_clean_solve(sys_, p, 200, args.rel_tol)
    cpu_ms = (time.perf_counter() - t0) * 1e3
    log(f"bench: cpu oracle solve {cpu_ms:.0f} ms, {it_cpu} it, rel {rel_cpu:.1e}")
    ok_ref, it_ref, rel_ref = oracle.mgpcg_clean_solve(sys_, p, 200, 1e-12)
    p_ref = sys_.x[0]
    nref = np.linalg.norm(p_ref)
    err = {k: float(np.linalg.norm(res[k]["x"] - p_ref) / nref) for k in res}
    log(f"bench: error vs oracle {err} (p_ref rel residual {rel_ref:.1e})")
    return {
        "system_build_cpu": {"cpu_ms": round(build_cpu_ms, 1), "cores": threads, "kind": "port",
                             "sample": "the oracle's build of the same levels (it also builds the "
                                       "coarse right-hand sides from down-sampled faces)"},
        "err_vs_oracle": err,
        "oracle_p_ref_rel_residual": rel_ref,
        "cpu_baseline": {"value": round(cpu_ms, 1), "unit": "ms", "cores": threads, "kind": "port",
                         "sample": f"full {n}^3 solve to rel residual {args.rel_tol:g} by the oracle's "
                                   f"FP64 MGPCG (restated FdmBlas3 ops, red-black V(2,2)), "
                                   f"{it_cpu} iterations, OpenMP over k"},
    }


if __name__ == "__main__":
    main()
