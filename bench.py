#!/usr/bin/env python3
"""Benchmark of the skew -> operator -> deskew -> reduce path on one MI355X.

Headline (default): batched total viewshed -- 32 fBm DEMs (Hurst U[0.7,0.9],
heights 0-3000 m, 10 m cells, observer 1.5 m) with sides drawn from
{1000, 2000, 3000, 4000}, all 180 sectors (both scan directions) of every DEM
per step.  Metric: Mpts*sectors/s = sum(n^2) * 180 / 1e6 / seconds per step.

One JSON line is printed.  `roofline` covers the viewshed operator kernel
(op_sdem), timed with HIP events on the engine streams; `cpu_baseline` is the
C/OpenMP oracle (a restatement of the reference CPU farm) timed on a bounded
sample on this host and scaled to the workload by the exact ray-step count.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X vector FP32 peak 157.3 TFLOP/s counts an FMA as 2 flops: 78.6e12
# FP32 VALU lane operations per second (256 CUs x 128 lanes/clk x 2.4 GHz).
# A VALU issue-rate microbenchmark on the box measures 6.6e13/s for chains
# of v_sub/v_max, consistent with that figure.
PEAK_VALU_TOPS = 157.3 / 2
OPS_PER_STEP = 4  # sub, divide, compare, max per ray step (src/kernelSDEM.h:191-196)
HBM_PEAK_GBS = 8000.0


def _traffic(name):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(name)
    except (OSError, ValueError):
        return None


def cpu_baseline_viewshed(total_steps, total_pts_sectors, seconds_budget):
    """Oracle farm (reference thread rule) on a 1000^2 DEM over a sector
    subset sized to ~seconds_budget, scaled by exact ray steps."""
    import oracle
    from skewengine_amd import synth
    n = 1000
    grid = synth.dem(n, 0.8, 12345, 3000.0)
    threads = oracle.default_threads()
    # probe one sector to size the sample
    _, st1 = oracle.work(n, n, 0, 1)
    _, t1 = oracle.execute(oracle.SDEM, grid, nthreads=1, s0=0, s1=1)
    rate1 = st1 / max(t1, 1e-9)
    s1 = 1
    while s1 < 180:
        _, st = oracle.work(n, n, 0, s1 + 1)
        if st / (rate1 * threads) > seconds_budget:
            break
        s1 += 1
    s1 = max(s1, min(threads, 180))
    _, steps = oracle.work(n, n, 0, s1)
    _, secs = oracle.execute(oracle.SDEM, grid, nthreads=threads, s0=0, s1=s1)
    rate = steps / secs
    cpu_time = total_steps / rate
    return {
        "value": total_pts_sectors / 1e6 / cpu_time,
        "unit": "Mpts*sectors/s",
        "cores": threads,
        "kind": "port",
        "sample": "1000x1000 fBm DEM, sectors 0-%d of 180, %d OpenMP threads (largest divisor of 180 <= %d "
                  "procs), %.1f s, %.3e ray steps/s; scaled to the workload by its exact ray-step count"
                  % (s1 - 1, threads, os.cpu_count() or 0, secs, rate),
    }


def bench_viewshed_batch(args, ske, synth):
    from skewengine_amd import _lib
    sides, hursts = synth.viewshed_batch_sizes(32, seed=4)
    dems = [synth.dem(n, h, 1000 + i, 3000.0) for i, (n, h) in enumerate(zip(sides, hursts))]
    batch = ske.SkewBatch(dems)  # inputs now resident in HBM
    del dems
    steps_total = 0.0
    pts_sectors = 0.0
    for n in sides:
        _, st = ske.work(n, n)
        steps_total += st
        pts_sectors += n * n * 180.0
    waves, pers, occ, _ = batch.autotune(ske.KERNEL_SDEM, ske.POVH_DEFAULT)
    batch.set_option(_lib.OPT_PROFILE, 1)  # HIP events around every launch on the batch stream

    for _ in range(args.warmup):
        batch.run(ske.KERNEL_SDEM, 0, 180, ske.POVH_DEFAULT)
    batch.synchronize()
    batch.profile_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run(ske.KERNEL_SDEM, 0, 180, ske.POVH_DEFAULT)
    batch.synchronize()
    t1 = time.perf_counter()
    sec_per_step = (t1 - t0) / args.steps
    k_ms, k_cnt = batch.profile("op_sdem")
    achieved = OPS_PER_STEP * steps_total * args.steps / (k_ms / 1e3) / 1e12
    result = {
        "metric": "Mpts*sectors/s (batched total viewshed)",
        "value": pts_sectors / 1e6 / sec_per_step,
        "unit": "Mpts*sectors/s",
        "ms_per_step": sec_per_step * 1e3,
        "dtype": "f32",
        "config": {"workload": "batched total viewshed: 32 fBm DEMs, sides %s, Hurst U[0.7,0.9], heights 0-3000 m, "
                               "10 m cells, observer 1.5 m, 180 sectors x 2 directions" % sorted(set(sides)),
                   "dems": len(sides), "sides": sides, "ray_steps_per_step": steps_total,
                   "pts_sectors_per_step": pts_sectors, "seed": 4},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_VALU_TOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_VALU_TOPS, "traffic": _traffic("op_sdem"),
                     "kernel": "op_sdem", "kernel_ms_total": k_ms, "launches": k_cnt,
                     "ray_steps_per_s": steps_total * args.steps / (k_ms / 1e3),
                     "autotuned": {"waves_per_workgroup": waves, "persistent_workgroups_per_cu": pers,
                                   "waves_per_simd_hint": occ}},
    }
    batch.close()
    return result, steps_total, pts_sectors


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs on one GPU only (the sector farm shards across GPUs only by design); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import skewengine_amd as ske
    from skewengine_amd import synth

    res, steps_total, pts_sectors = bench_viewshed_batch(args, ske, synth)
    line = {
        "metric": res["metric"], "value": res["value"], "unit": res["unit"], "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": res["dtype"],
        "data": "synthetic (seeded fBm spectral synthesis)", "config": res["config"],
        "roofline": res["roofline"],
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline_viewshed(steps_total, pts_sectors, args.cpu_seconds)
        line["speedup_vs_cpu"] = line["value"] / line["cpu_baseline"]["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
