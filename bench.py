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

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X vector FP32 peak 157.3 TFLOP/s counts an FMA as 2 flops: 78.6e12
# FP32 VALU lane operations per second (256 CUs x 128 lanes/clk x 2.4 GHz).
# A VALU issue-rate microbenchmark on the box measures 6.6e13/s for chains
# of v_sub/v_max, consistent with that figure.
PEAK_VALU_TOPS = 157.3 / 2
OPS_PER_STEP = 4  # sub, divide, compare, max per ray step (src/kernelSDEM.h:191-196)
# VALU wave instructions per 64 executed lane-steps of the viewshed kernel,
# measured with rocprofv3 --pmc SQ_INSTS_VALU over the bench's timed step
# (6 per step in the scan, then per 32-step block of a slot the popcount
# flush, the end-of-row mask, the early-exit bound and the retire/compaction
# bookkeeping, plus row staging and chunk tables).
VALU_PER_LANE_STEP = 8.72
HBM_PEAK_GBS = 8000.0


def _traffic(name):
    """HBM bytes per launch of `name` from the committed rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE passes (copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(name)
        return rec["bytes_per_launch"] if rec else None
    except (OSError, ValueError, KeyError, TypeError):
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


def _progress(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def bench_viewshed_batch(args, ske, synth):
    from skewengine_amd import _lib
    sides, hursts = synth.viewshed_batch_sizes(32, seed=4)
    dems = [synth.dem(n, h, 1000 + i, 3000.0) for i, (n, h) in enumerate(zip(sides, hursts))]
    _progress("synthesised %d DEMs" % len(dems))
    t_in = time.perf_counter()
    batch = ske.SkewBatch(dems)  # inputs now resident in HBM (host -> device copy + prepare)
    batch.synchronize()
    t_in = time.perf_counter() - t_in
    del dems
    steps_total = 0.0
    pts_sectors = 0.0
    for n in sides:
        _, st = ske.work(n, n)
        steps_total += st
        pts_sectors += n * n * 180.0
    waves, pers, occ, _ = batch.autotune(ske.KERNEL_SDEM, ske.POVH_DEFAULT)
    group = batch.get_option(_lib.OPT_SDEM_GROUP)
    _progress("autotuned: %d waves/workgroup, %d persistent workgroups/CU, occupancy hint %d, chunk group %d"
              % (waves, pers, occ, group))
    batch.set_option(_lib.OPT_PROFILE, 1)  # HIP events around every launch on the batch stream

    for _ in range(args.warmup):
        batch.run(ske.KERNEL_SDEM, 0, 180, ske.POVH_DEFAULT)
        batch.synchronize()
        _progress("warmup step done")
    batch.profile_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run(ske.KERNEL_SDEM, 0, 180, ske.POVH_DEFAULT)
    batch.synchronize()
    t1 = time.perf_counter()
    _progress("timed %d step(s): %.2f s" % (args.steps, t1 - t0))
    sec_per_step = (t1 - t0) / args.steps
    t_out = time.perf_counter()
    batch.outputs()  # device -> host copy of the 32 result maps
    t_out = time.perf_counter() - t_out
    k_ms, k_cnt = batch.profile("op_sdem")
    # One more (untimed) step with the kernels counting the ray steps their
    # lanes execute: the scans stop once no later target can be visible.
    batch.set_option(_lib.OPT_COUNT_STEPS, 1)
    batch.profile_reset()
    batch.run(ske.KERNEL_SDEM, 0, 180, ske.POVH_DEFAULT)
    lane_steps = batch.counter("op_sdem_lane_steps")
    batch.set_option(_lib.OPT_COUNT_STEPS, 0)
    _progress("counted %.4e executed lane-steps (%.3f of the reference's ray steps)"
              % (lane_steps, lane_steps / steps_total))
    # Roofline of the work the kernel performs: the scans stop once no later
    # target can be visible, so the steps executed (whole 32-step blocks of 64
    # lanes) replace the reference's sum L(L-1) as the algorithmic count.
    achieved = OPS_PER_STEP * lane_steps * args.steps / (k_ms / 1e3) / 1e12
    result = {
        "metric": "Mpts*sectors/s (batched total viewshed)",
        "value": pts_sectors / 1e6 / sec_per_step,
        "unit": "Mpts*sectors/s",
        "ms_per_step": sec_per_step * 1e3,
        "dtype": "f32",
        "config": {"workload": "batched total viewshed: 32 fBm DEMs, sides %s, Hurst U[0.7,0.9], heights 0-3000 m, "
                               "10 m cells, observer 1.5 m, 180 sectors x 2 directions" % sorted(set(sides)),
                   "dems": len(sides), "sides": sides, "ray_steps_per_step": steps_total,
                   "pts_sectors_per_step": pts_sectors, "seed": 4,
                   "end_to_end_s_incl_host_copies": t_in + sec_per_step + t_out,
                   "host_to_device_and_prepare_s": t_in, "device_to_host_s": t_out},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_VALU_TOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_VALU_TOPS, "traffic": _traffic("op_sdem"),
                     "kernel": "op_sdem", "kernel_ms_total": k_ms, "launches": k_cnt,
                     "ray_steps_per_s": steps_total * args.steps / (k_ms / 1e3),
                     "reference_step_equivalent_frac": OPS_PER_STEP * steps_total * args.steps / (k_ms / 1e3)
                                                       / 1e12 / PEAK_VALU_TOPS,
                     "executed_lane_steps_per_step": lane_steps,
                     "executed_over_reference_steps": lane_steps / steps_total,
                     "issued_valu_frac": VALU_PER_LANE_STEP * lane_steps * args.steps / (k_ms / 1e3) / 1e12
                                         / PEAK_VALU_TOPS,
                     "autotuned": {"waves_per_workgroup": waves, "persistent_workgroups_per_cu": pers,
                                   "waves_per_simd_hint": occ, "chunk_group": group}},
    }
    batch.close()
    return result, steps_total, pts_sectors


def _timed(fn, sync, reps):
    fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    sync()
    return (time.perf_counter() - t0) / reps


def extra_c1_viewshed(ske, synth):
    """Config 1: one 500^2 DEM (H=0.8, 0-1500 m), 180 sectors."""
    g = synth.dem(500, 0.8, 1, 1500.0)
    _, steps = ske.work(500, 500)
    with ske.SkewEngine(g) as e:
        t = _timed(lambda: e.run(ske.KERNEL_SDEM, copy_out=False), e.synchronize, 20)
    return {"workload": "500x500 fBm DEM, 180 sectors", "ms": t * 1e3,
            "Mpts_sectors_per_s": 500 * 500 * 180 / 1e6 / t, "ray_steps_per_s": steps / t}


def extra_c2_radon(ske, synth):
    """Config 2: 512^2 ellipse phantom, 180 angles -- the reference-semantics
    skew-space Radon/FBP operator, and the standard 725-bin projector + Ram-Lak
    FBP, in slices/s."""
    from skewengine_amd.fbp import FBP
    img = synth.ellipse_phantom(512, seed=2)
    res = {}
    with ske.SkewEngine(img) as e:
        t = _timed(lambda: e.run(ske.KERNEL_RADON, copy_out=False), e.synchronize, 20)
    res["skew_radon_fbp"] = {"workload": "512x512 phantom, 180 skew sectors (kernelRADN semantics)",
                             "ms_per_slice": t * 1e3, "slices_per_s": 1.0 / t}
    nb = 16
    with ske.SkewBatch([synth.ellipse_phantom(512, seed=200 + i) for i in range(nb)]) as b:
        t = _timed(lambda: b.run(ske.KERNEL_RADON), b.synchronize, 10)
    res["skew_radon_fbp_batch"] = {"workload": "%d x 512x512 phantoms, 180 skew sectors each, one batched run" % nb,
                                   "ms_per_slice": t * 1e3 / nb, "slices_per_s": nb / t}
    from skewengine_amd import _lib
    cells, _ = ske.work(512, 512)
    with ske.SkewEngine(img) as e:
        e.set_option(_lib.OPT_RADON_LOCAL, 1)
        t = _timed(lambda: e.run(ske.KERNEL_RADON, copy_out=False), e.synchronize, 3)
    res["local_radon"] = {"workload": "512x512 phantom, 180 skew sectors, local transform (pointRadon, LRT=true)",
                          "ms_per_slice": t * 1e3, "cells_per_s": cells / t,
                          "fp64_TFLOPs_if_6240_per_cell": 6240 * cells / t / 1e12}
    batch = 16
    imgs = np.stack([synth.ellipse_phantom(512, seed=100 + i) for i in range(batch)])
    with FBP(512, 180, 725, max_batch=batch) as f:
        f.upload(0, imgs)
        f.set_profile(True)

        def step():
            f.project(batch)
            f.reconstruct(batch)
        t = _timed(step, f.synchronize, 5)
        f.profile_reset()
        step()
        f.synchronize()
        k = {name: f.profile(name)[0] / batch for name in ("fbp_project", "fbp_filter", "fbp_backproject")}
    res["standard_radon_plus_fbp"] = {
        "workload": "512x512 phantoms, 180 angles x 725 bins, project + Ram-Lak filter + backproject, batch %d" % batch,
        "ms_per_slice": t * 1e3 / batch, "slices_per_s": batch / t, "kernel_ms_per_slice": k}
    return res


def extra_c3_cepstrum(ske, synth, count):
    """Config 3: batched cepstrum blur estimation over `count` 1024^2 images."""
    from skewengine_amd import _lib
    imgs = [synth.blurred_texture(1024, 0.6, seed=i)[0] for i in range(count)]
    cells, _ = ske.work(1024, 1024)
    with ske.SkewBatch(imgs) as b:
        del imgs
        b.set_option(_lib.OPT_PROFILE, 1)
        b.run(ske.KERNEL_CEPSTRUM)
        b.synchronize()
        b.profile_reset()
        t0 = time.perf_counter()
        b.run(ske.KERNEL_CEPSTRUM)
        b.synchronize()
        t = time.perf_counter() - t0
        k_ms, k_cnt = b.profile("op_blur")
    total = cells * count
    return {"workload": "%d x 1024^2 fBm (H=0.6) + motion blur + noise, 180 sectors, batched" % count,
            "s": t, "cell_sectors_per_s": total / t, "op_blur_ms": k_ms, "op_blur_launches": k_cnt,
            "op_blur_TFLOPs_fp64_at_2848_per_cell": 2848 * total / (k_ms / 1e3) / 1e12}


def extra_c5_learned(batch=64):
    """Config 5: learned sparse-view FBP on 512^2 phantoms, 60 angles, 725
    bins, Poisson noise at I0 = 1e5, batch 64: time of plain Ram-Lak FBP vs
    the learned reconstruction (filter + back-projection + CNN) on the same
    sinograms, and held-out PSNR of both with the committed parameters."""
    from skewengine_amd import _lib
    from skewengine_amd import learned as lm
    from skewengine_amd.fbp import FBP
    path = lm.default_weights(512, 60, 725)
    res = {"workload": "%d x 512^2 ellipse phantoms, 60 angles x 725 bins, Poisson I0=1e5, batch inference" % batch}
    with FBP(512, 60, 725, max_batch=batch) as f, lm.LearnedFBP(f) as m:
        if path:
            m.load(path)
        res["weights"] = os.path.basename(path) if path else None
        m.make_batch(batch, seed=lm.HELD_OUT_SEED)
        t_fbp = _timed(lambda: f.reconstruct(batch), f.synchronize, 10)
        t_l = _timed(lambda: m.reconstruct(batch), f.synchronize, 10)
        f.set_profile(True)
        f.profile_reset()
        m.reconstruct(batch)
        f.synchronize()
        cnn_ms = f.profile("cnn_forward")[0]
        f.set_profile(False)
        c = _lib.LEARNED_CHANNELS
        flop = 2 * 9 * (1 * c + c * c + c * c + c * 1) * 512 * 512 * batch
        res.update({"fbp_ms_per_slice": t_fbp * 1e3 / batch, "learned_ms_per_slice": t_l * 1e3 / batch,
                    "time_ratio": t_l / t_fbp, "cnn_ms_per_slice": cnn_ms / batch,
                    "cnn_TFLOPs": flop / (cnn_ms / 1e3) / 1e12})
        if path:
            p_fbp, p_l = m.evaluate(64, batch)
            res.update({"psnr_fbp_db": float(p_fbp.mean()), "psnr_learned_db": float(p_l.mean()),
                        "psnr_gain_db": float((p_l - p_fbp).mean()), "held_out": 64})
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the secondary configurations 1, 2, 3 and 5")
    ap.add_argument("--cepstrum-images", type=int, default=64)
    ap.add_argument("--only-extra", choices=["c1", "c2", "c3", "c5"], default=None,
                    help="run just one secondary configuration and print its JSON")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 4, "copies": 3}))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs on one GPU only (the sector farm shards across GPUs only by design); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import skewengine_amd as ske
    from skewengine_amd import synth

    if args.only_extra:
        fn = {"c1": lambda: extra_c1_viewshed(ske, synth), "c2": lambda: extra_c2_radon(ske, synth),
              "c3": lambda: extra_c3_cepstrum(ske, synth, args.cepstrum_images), "c5": extra_c5_learned}
        print(json.dumps({args.only_extra: fn[args.only_extra]()}))
        return

    res, steps_total, pts_sectors = bench_viewshed_batch(args, ske, synth)
    line = {
        "metric": res["metric"], "value": res["value"], "unit": res["unit"], "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": res["dtype"],
        "data": "synthetic (seeded fBm spectral synthesis)", "config": res["config"],
        "roofline": res["roofline"],
    }
    if not args.no_extra:
        line["extra"] = {"c1_viewshed_500": extra_c1_viewshed(ske, synth),
                         "c2_radon_fbp_512": extra_c2_radon(ske, synth),
                         "c3_cepstrum_batch": extra_c3_cepstrum(ske, synth, args.cepstrum_images),
                         "c5_learned_sparse_fbp": extra_c5_learned()}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline_viewshed(steps_total, pts_sectors, args.cpu_seconds)
        line["speedup_vs_cpu"] = line["value"] / line["cpu_baseline"]["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
