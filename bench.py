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

The line also carries the 512x256x256 run (coarse 16x8x8, 6 levels, mixed,
OC fed by the autograd compliance; 5 design iterations after 1 warm-up):
s/design-iter, DOFs/s, the live fine-sweep roofline and peak device memory.

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


def host_description():
    """nproc, usable CPUs and the CPU model of this host (for the baseline)."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "cpu_model": model}


def cpu_baseline(nelx, nely, nelz, nl, gpu_same, small_gpu):
    """FP64 oracle on the host: the first design iteration of the cantilever
    (uniform density 0.2, cold start; state solve with its set-up and CG,
    volume, sensitivity, filter gradient, OC, change, filter), run whole and
    timed, not extrapolated; then the reference's small cantilevers beside
    the GPU's times for them."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    host = host_description()
    pyoracle.Grid(16, 8, 8, 3).close()  # load the library before the clock starts
    try:
        import ctypes
        threads = ctypes.CDLL("libgomp.so.1").omp_get_max_threads()
    except OSError:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or host["usable_cpus"]
    t0 = time.perf_counter()
    recs, _, _ = pyoracle.top3d(nelx, nely, nelz, nl, 1)
    sec = time.perf_counter() - t0
    small = small_config_cpu(pyoracle, small_gpu, threads)
    return {"value": sec, "unit": "s/design-iter", "cores": threads, "kind": "port",
            "sample": (f"FP64 oracle (oracle/top3d_oracle.c, OpenMP {threads} threads) running "
                       f"the first design iteration of the same {nelx}x{nely}x{nelz} nl={nl} "
                       f"cantilever whole (uniform density 0.2, cold start, "
                       f"{recs[0]['cgiter'] + 1} CG iterations, then sensitivity, filters and "
                       f"OC), timed end to end; the GPU takes "
                       f"{gpu_same['mixed_s']:.4f} s (mixed) and {gpu_same['fp64_s']:.4f} s "
                       f"(FP64) for that same design iteration"),
            "host": host, "omp_num_threads": threads,
            "compliance": recs[0]["c"], "cg_iters": recs[0]["cgiter"] + 1,
            "gpu_same_iteration": gpu_same, "small_configs": small}


SMALL_CONFIGS = (("48x24x24, 3 levels", (48, 24, 24, 3), 5),
                 ("128x64x64, 4 levels", (128, 64, 64, 4), 2))


def small_config_gpu(top3d_amd, torch):
    """The reference's own small cantilevers (-x 12 -y 6 -z 6 -l 3 and
    -x 16 -y 8 -z 8 -l 4): the first design iterations run whole on the GPU
    in both forms, on a grid whose kernels are loaded (one design iteration
    first, untimed)."""
    out = {}
    for name, shape, iters in SMALL_CONFIGS:
        for form in ("fp64", "mixed"):
            g = top3d_amd.Grid(*shape, precision=form)
            top3d_amd.design_loop(g, 1)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            recs, _ = top3d_amd.design_loop(g, iters)
            torch.cuda.synchronize()
            out[(name, form)] = ((time.perf_counter() - t0) / len(recs),
                                 [r["cgiter"] + 1 for r in recs])
            g.close()
    return out


def small_config_cpu(pyoracle, gpu, threads):
    """The same design iterations with the FP64 oracle on the host, timed
    around the oracle's whole top3d call (its grid set-up included; the GPU
    times exclude creating the grid), with all threads and, for the smallest
    case, one thread."""
    import ctypes
    try:
        gomp = ctypes.CDLL("libgomp.so.1")
    except OSError:
        gomp = None
    rows = []
    for name, shape, iters in SMALL_CONFIGS:
        for nthr in ((threads, 1) if shape[0] == 48 and gomp else (threads,)):
            if gomp:
                gomp.omp_set_num_threads(nthr)
            t0 = time.perf_counter()
            recs, _, _ = pyoracle.top3d(*shape, iters)
            sec = (time.perf_counter() - t0) / len(recs)
            g64, cg64 = gpu.get((name, "fp64"), (float("nan"), []))
            g32, _ = gpu.get((name, "mixed"), (float("nan"), []))
            rows.append({"workload": f"cantilever {name}", "design_iters": len(recs),
                         "threads": nthr, "cpu_s_per_design_iter": sec,
                         "gpu_fp64_s_per_design_iter": g64, "gpu_mixed_s_per_design_iter": g32,
                         "cg_iters_cpu": [r["cgiter"] + 1 for r in recs], "cg_iters_gpu_fp64": cg64})
    if gomp:
        gomp.omp_set_num_threads(threads)
    return rows


def first_iteration_gpu(top3d_amd, torch, g):
    """Whole first design iteration (uniform density, cold start) on a warm
    grid: the work the CPU baseline times."""
    st = top3d_amd.DesignState(g, volfrac=0.2, rmin=1.5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rec = top3d_amd.design_step(st, cgtol=1e-5)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, rec


def memory_record(top3d_amd, torch):
    m = top3d_amd.memory_info()
    return {"library_peak_GiB": m["library_peak_bytes"] / 2 ** 30,
            "torch_peak_GiB": torch.cuda.max_memory_allocated() / 2 ** 30,
            "device_used_GiB": m["device_used_bytes"] / 2 ** 30,
            "device_total_GiB": m["device_total_bytes"] / 2 ** 30}


def run_c5(top3d_amd, torch, cx, cy, cz, steps, warmup):
    """Largest single-GPU run: coarse 16x8x8, 6 levels (512x256x256), mixed
    form, OC fed by the autograd compliance and filter (design_step_autograd)."""
    nl = 6
    f = 1 << (nl - 1)
    nelx, nely, nelz = cx * f, cy * f, cz * f
    dofs = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    top3d_amd.memory_info(reset_peak=True)
    torch.cuda.reset_peak_memory_stats()
    g = top3d_amd.Grid(nelx, nely, nelz, nl, precision="mixed")
    st = top3d_amd.DesignState(g, volfrac=0.2, rmin=1.5)
    recs = []

    def step():
        t0 = time.perf_counter()
        rec = top3d_amd.design_step_autograd(st, cgtol=1e-5)
        torch.cuda.synchronize()
        stt = g.stats()
        rec["solve_s"] = (stt["solve_ms"] + stt["setup_ms"]) / 1e3
        rec["step_s"] = time.perf_counter() - t0
        return rec

    for _ in range(warmup):
        r = step()
        print(f"[bench] c5 warm-up step: solve {r['solve_s']:.4f}s cg {r['cgiter'] + 1}",
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    g.sweep_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        recs.append(step())
        print(f"[bench] c5 step {len(recs)}: solve {recs[-1]['solve_s']:.4f}s "
              f"cg {recs[-1]['cgiter'] + 1} c {recs[-1]['c']:.6e}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sweep_ms, sweeps, sweep_bytes = g.sweep_timing(False)
    solve_s = sum(r["solve_s"] for r in recs) / len(recs)
    gbs = sweep_bytes / (sweep_ms * 1e-3) / 1e9 if sweep_ms > 0 else 0.0
    out = {"workload": f"cantilever {nelx}x{nely}x{nelz}, coarse {cx}x{cy}x{cz}, {nl} levels, "
                       f"mixed MGCG, OC fed by the autograd compliance/filter gradient",
           "dofs": dofs, "steps": steps, "warmup": warmup,
           "s_per_design_iter": wall / steps, "solve_s_per_design_iter": solve_s,
           "dofs_per_s": dofs / solve_s, "cg_iters": [r["cgiter"] + 1 for r in recs],
           "compliance": [r["c"] for r in recs],
           "roofline": {"bound": "hbm", "kernel": "fine_sweep_f32", "achieved": gbs,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                        "algorithmic_bytes": sweep_bytes, "avg_launch_ms": sweep_ms,
                        "launches_timed": sweeps},
           "memory": memory_record(top3d_amd, torch)}
    g.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--coarse", default="16,8,8")
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--compare-fp64", type=int, default=1)
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--c5", type=int, default=1,
                    help="also run the 512x256x256 (6 levels) mixed design loop with the "
                         "autograd OC update for 5 design iterations")
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

    torch.cuda.reset_peak_memory_stats()
    top3d_amd.memory_info(reset_peak=True)
    g, st, recs, wall = run_form(top3d_amd, torch, nelx, nely, nelz, nl, "mixed",
                                 args.warmup, args.steps)
    solve_s = sum(r["solve_s"] for r in recs) / len(recs)
    cg_mixed = [r["cgiter"] + 1 for r in recs]

    # kernel timings on the mixed run's buffers (HIP events, library stream):
    # back to back on the same buffers (operand sets under 256 MiB then stay
    # in the memory-side cache), and "cold", each launch after a 1 GiB copy
    # that flushes that cache, as inside a solve
    kernels = {}
    for name in ("fine_sweep_f32", "level1_sweep_f32", "fine_apply_f64", "fine_sweep_f64"):
        ms, by = g.time_kernel(name, args.kernel_reps)
        cms, _ = g.time_kernel(name, args.kernel_reps, cold=True)
        kernels[name] = {"avg_ms": ms, "bytes": by, "GB_s": by / (ms * 1e-3) / 1e9,
                         "hbm_frac": by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "cold_avg_ms": cms,
                         "cold_hbm_frac": by / (cms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    # dominant kernel of the mixed solve: the FP32 fine sweep (7 per V-cycle),
    # timed live over the timed design iterations
    live = recs[0]["sweep_live"]
    dom = "fine_sweep_f32"
    traffic = measured_traffic(dom, f"{nelx}x{nely}x{nelz}")
    live_gbs = live["bytes"] / (live["avg_ms"] * 1e-3) / 1e9 if live["avg_ms"] > 0 else 0.0
    memory = memory_record(top3d_amd, torch)

    fp64 = None
    contract = None
    gpu_same = {}
    if args.compare_fp64:
        g64, st64, recs64, _ = run_form(top3d_amd, torch, nelx, nely, nelz, nl, "fp64",
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
        # precision contract on one density: the FP64 run's last design,
        # both forms from a cold start
        xl = st64.xPhys
        res = {}
        for name, gg in (("fp64", g64), ("mixed", g)):
            U = gg.zeros(0)
            it, _ = gg.solve(xl, st64.F, U)
            res[name] = (it + 1, float(st64.F @ U), U)
        (i64, c64, u64), (i32, c32, u32) = res["fp64"], res["mixed"]
        contract = {"density": f"last design of the FP64 run (design iteration "
                               f"{args.warmup + args.steps}), cold start",
                    "compliance_rel_err": abs(c32 - c64) / abs(c64),
                    "displacement_rel_l2": float(torch.linalg.norm(u32 - u64) /
                                                 torch.linalg.norm(u64)),
                    "cg_iters_fp64": i64, "cg_iters_mixed": i32, "cg_iter_change": i32 - i64}
        del u64, u32
        if args.cpu_baseline:
            gpu_same["mixed_s"], _ = first_iteration_gpu(top3d_amd, torch, g)
            gpu_same["fp64_s"], _ = first_iteration_gpu(top3d_amd, torch, g64)
        g64.close()
    g.close()
    small_gpu = small_config_gpu(top3d_amd, torch) if args.cpu_baseline else {}

    c5 = None
    if args.c5:
        print("[bench] 512x256x256, 6 levels, autograd OC ...", file=sys.stderr, flush=True)
        c5 = run_c5(top3d_amd, torch, cx, cy, cz, 5, 1)

    cpu = None
    if args.cpu_baseline:
        if not gpu_same:
            gpu_same = {"mixed_s": float("nan"), "fp64_s": float("nan")}
        print("[bench] CPU baseline (oracle) ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(nelx, nely, nelz, nl, gpu_same, small_gpu)

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
        "memory": memory,
        "kernels": kernels,
        "fp64": fp64,
        "precision_contract": contract,
        "c5": c5,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
