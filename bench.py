#!/usr/bin/env python3
"""MGCG state-solve benchmark on one MI355X.

Metric (the reference's headline path): seconds of MGCG state solve per
SIMP design iteration on the 256x128x128-element cantilever (coarse 16x8x8,
5 multigrid levels, rtol 1e-5, V(4,4) damped Jacobi, omega 0.6), mixed
precision (FP32 V-cycle, FP64 outer CG and coarse Cholesky), next to the
all-FP64 form on the same run; plus the achieved HBM bandwidth of the
dominant stencil kernel.

A step is one design iteration of top3dmgcg (stencil_optimization.c:150-189)
driven through the C ABI: solve, volume, sensitivity, filter gradient, OC
update, change, filter.  W untimed warm-up steps, then exactly K timed steps
between device synchronisations.  Data are the synthetic cantilever (load and
supports of the reference), uniform initial density 0.2.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "topopt-in-openmp-gpu_amd"))

HBM_PEAK_GBS = 8000.0
SWEEPS_PER_VCYCLE = 7  # 4 pre (the first from zero needs no operator) + 4 post - 1


def run_form(top3d_amd, torch, nelx, nely, nelz, nl, precision, warmup, steps):
    g = top3d_amd.Grid(nelx, nely, nelz, nl, precision=precision)
    st = top3d_amd.DesignState(g, volfrac=0.2, rmin=1.5)
    recs = []

    def step():
        t0 = time.perf_counter()
        rec = top3d_amd.design_step(st, cgtol=1e-5)
        stt = g.stats()
        rec["solve_s"] = stt["solve_ms"] / 1e3 + stt["setup_ms"] / 1e3
        rec["setup_s"] = stt["setup_ms"] / 1e3
        rec["step_s"] = time.perf_counter() - t0
        return rec

    for _ in range(warmup):
        r = step()
        print(f"[bench] {precision} warm-up step: solve {r['solve_s']:.4f}s cg {r['cgiter'] + 1}",
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    g.sweep_timing(True)  # HIP events around the runs of fine sweeps of the timed solves
    t0 = time.perf_counter()
    for _ in range(steps):
        recs.append(step())
        print(f"[bench] {precision} step {len(recs)}: solve {recs[-1]['solve_s']:.4f}s "
              f"cg {recs[-1]['cgiter'] + 1} c {recs[-1]['c']:.6e}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sweep_ms, sweeps, sweep_bytes = g.sweep_timing(False)
    recs[0]["sweep_live"] = {"avg_ms": sweep_ms, "launches": sweeps, "bytes": sweep_bytes}
    return g, st, recs, wall


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE scaled by a reference kernel), or None."""
    import re
    path = os.path.join(ROOT, "profiles", f"fine_sweep_f32_{workload}_pmc.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(rf"traffic_bytes_per_launch {kernel}: (\d+)", text)
    return float(m.group(1)) if m else None


def cpu_baseline(nelx, nely, nelz, nl, gpu_iters):
    """FP64 oracle on the host: per-solve set-up plus one CG iteration of the
    first design iteration, extrapolated to the GPU's CG iteration count."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    g = pyoracle.Grid(nelx, nely, nelz, nl)
    F = g.load()
    _, wx, wy, wz, ndof = g.padding(0)
    x = np.zeros(g.nelem())
    i, k, j = np.meshgrid(np.arange(1, nelx + 1), np.arange(1, nelz + 1),
                          np.arange(1, nely + 1), indexing="ij")
    x[(i * (wy - 1) * (wz - 1) + k * (wy - 1) + j).ravel()] = float(np.float32(0.2))
    U = np.zeros(ndof)
    g.solve(x, F, U, maxiter=1)
    setup_s, loop_s = g.solve_times()
    g.close()
    value = setup_s + loop_s * gpu_iters
    return {"value": value, "unit": "s/design-iter", "cores": threads, "kind": "port",
            "sample": (f"FP64 oracle (oracle/top3d_oracle.c, OpenMP {threads} threads) on the same "
                       f"{nelx}x{nely}x{nelz} nl={nl} cantilever from uniform density: per-solve "
                       f"set-up {setup_s:.2f}s + 1 CG iteration {loop_s:.2f}s measured, "
                       f"extrapolated to {gpu_iters:.1f} CG iterations per solve (the GPU FP64 "
                       f"run's mean)")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--coarse", default="16,8,8")
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--compare-fp64", type=int, default=1)
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--kernel-only", default="",
                    help="profiling aid: one mixed design iteration, then only REPS launches of "
                         "this kernel (and a 256 MiB device copy as a byte-count reference)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; --gpus N > 1 is not supported "
              "(the x-slab multi-GPU decomposition is design-only)", file=sys.stderr)
        sys.exit(2)
    import torch
    import top3d_amd

    cx, cy, cz = (int(v) for v in args.coarse.split(","))
    f = 1 << (args.levels - 1)
    nelx, nely, nelz, nl = cx * f, cy * f, cz * f, args.levels
    dofs = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)

    if args.kernel_only:
        g, _, _, _ = run_form(top3d_amd, torch, nelx, nely, nelz, nl, "mixed", 0, 1)
        ms, by = g.time_kernel(args.kernel_only, args.kernel_reps)
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):  # one elementwise kernel: 256 MiB read, 256 MiB written
            torch.add(src, 1.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"kernel": args.kernel_only, "avg_ms": ms, "bytes": by,
                          "copy_bytes_each_way": src.numel() * 4}), flush=True)
        g.close()
        return

    g, st, recs, wall = run_form(top3d_amd, torch, nelx, nely, nelz, nl, "mixed",
                                 args.warmup, args.steps)
    solve_s = sum(r["solve_s"] for r in recs) / len(recs)
    cg_mixed = [r["cgiter"] + 1 for r in recs]

    # kernel timings on the mixed run's buffers (HIP events, library stream)
    kernels = {}
    for name in ("fine_sweep_f32", "level1_sweep_f32", "fine_apply_f64", "fine_sweep_f64"):
        ms, by = g.time_kernel(name, args.kernel_reps)
        kernels[name] = {"avg_ms": ms, "bytes": by, "GB_s": by / (ms * 1e-3) / 1e9,
                         "hbm_frac": by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    # dominant kernel of the mixed solve: the FP32 fine sweep (7 per V-cycle),
    # timed live over the timed design iterations
    live = recs[0]["sweep_live"]
    dom = "fine_sweep_f32"
    traffic = measured_traffic(dom, f"{nelx}x{nely}x{nelz}")
    live_gbs = live["bytes"] / (live["avg_ms"] * 1e-3) / 1e9 if live["avg_ms"] > 0 else 0.0
    g.close()

    fp64 = None
    if args.compare_fp64:
        g64, _, recs64, _ = run_form(top3d_amd, torch, nelx, nely, nelz, nl, "fp64",
                                     args.warmup, args.steps)
        s64 = sum(r["solve_s"] for r in recs64) / len(recs64)
        cdiff = max(abs(a["c"] - b["c"]) / abs(b["c"]) for a, b in zip(recs, recs64))
        fp64 = {"solve_s_per_design_iter": s64,
                "setup_s_per_design_iter": sum(r["setup_s"] for r in recs64) / len(recs64),
                "cg_iters": [r["cgiter"] + 1 for r in recs64],
                "compliance": [r["c"] for r in recs64],
                "mixed_vs_fp64_max_rel_compliance_diff": cdiff,
                "fine_sweep_f64_live": recs64[0]["sweep_live"],
                "speedup_mixed_over_fp64": s64 / solve_s}
        g64.close()

    cpu = None
    if args.cpu_baseline:
        iters = (sum(fp64["cg_iters"]) / len(fp64["cg_iters"]) if fp64
                 else sum(cg_mixed) / len(cg_mixed))
        print("[bench] CPU baseline (oracle) ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(nelx, nely, nelz, nl, iters)

    out = {
        "metric": "MGCG solve s/design-iter at 256x128x128, mixed vs FP64 @rtol 1e-5; matvec GB/s",
        "value": solve_s,
        "unit": "s/design-iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle / f64 CG",
        "data": "synthetic cantilever (reference load and supports), uniform initial density 0.2",
        "config": {"workload": f"SIMP design iterations, MGCG state solve, cantilever "
                               f"{nelx}x{nely}x{nelz} elements, coarse {cx}x{cy}x{cz}, "
                               f"{nl} levels, V(4,4), rtol 1e-5",
                   "elements": nelx * nely * nelz, "dofs": dofs, "levels": nl,
                   "precision": "mixed"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": live_gbs,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": live_gbs / HBM_PEAK_GBS,
                     "traffic": traffic, "algorithmic_bytes": live["bytes"],
                     "avg_launch_ms": live["avg_ms"], "launches_timed": live["launches"]},
        "cpu_baseline": cpu,
        "cg_iters": cg_mixed,
        "compliance": [r["c"] for r in recs],
        "design_iter_s": wall / args.steps,
        "setup_s_per_design_iter": sum(r["setup_s"] for r in recs) / len(recs),
        "dofs_per_s": dofs / solve_s,
        "kernels": kernels,
        "fp64": fp64,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
