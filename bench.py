#!/usr/bin/env python3
"""Benchmark of the skew -> operator -> deskew -> reduce path on one MI355X.

Headline (default): batched total viewshed -- 32 fBm DEMs (Hurst U[0.7,0.9],
heights 0-3000 m, 10 m cells, observer 1.5 m) with sides drawn from
{1000, 2000, 3000, 4000}, all 180 sectors (both scan directions) of every DEM
per step.  Metric: Mpts*sectors/s = sum(n^2) * 180 / 1e6 / seconds per step.

The timed steps replay the step captured once in a hipGraph.  One JSON line
is printed.  `roofline` covers the viewshed operator kernel (op_sdem), timed
with HIP events around the graph replays on the batch stream; `cpu_baseline`
is the C/OpenMP oracle (a restatement of the reference CPU farm) timed on a
bounded sample on this host and scaled to the workload by the exact ray-step
count; `parity` compares the GPU map of that sample with the oracle's (and
both with the FP64 restatement).  `extra` holds the other configurations,
each with its CPU time and its error against the oracle.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published vector FP32 peak: 157.3 TFLOP/s, counting an FMA as two
# flops.  The viewshed's operations (sub, divide, compare, max) are not FMAs,
# so its issue ceiling is half of that: 78.6e12 FP32 VALU lane operations per
# second (256 CUs x 128 lanes/clk x 2.4 GHz); a VALU issue-rate
# microbenchmark on the box measures 6.6e13/s for chains of v_sub/v_max.
PEAK_FP32_PUBLISHED = 157.3
PEAK_VALU_TOPS = PEAK_FP32_PUBLISHED / 2
OPS_PER_STEP = 4  # sub, divide, compare, max per ray step (src/kernelSDEM.h:191-196)
# VALU wave instructions per 64 executed lane-steps of the viewshed kernel,
# measured with rocprofv3 --pmc SQ_INSTS_VALU over the bench's timed step
# (6 per step in the scan, then per 32-step block of a slot the popcount
# flush, the end-of-row mask, the early-exit bound and the retire/compaction
# bookkeeping, plus row staging and chunk tables).
VALU_PER_LANE_STEP = 8.26
HBM_PEAK_GBS = 8000.0


def _traffic(name):
    """HBM bytes per launch of `name`: a constant committed from the
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes (copy-calibrated) of the
    same command, not measured by this run; None if absent."""
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
    farm, secs = oracle.execute(oracle.SDEM, grid, nthreads=threads, s0=0, s1=s1)
    rate = steps / secs
    cpu_time = total_steps / rate
    sample = {"grid": grid, "s1": s1, "farm": farm, "threads": threads}
    return sample, {
        "value": total_pts_sectors / 1e6 / cpu_time,
        "unit": "Mpts*sectors/s",
        "cores": threads,
        "kind": "port",
        "sample": "1000x1000 fBm DEM, sectors 0-%d of 180, %d OpenMP threads (largest divisor of 180 <= %d "
                  "procs), %.1f s, %.3e ray steps/s; scaled to the workload by its exact ray-step count"
                  % (s1 - 1, threads, os.cpu_count() or 0, secs, rate),
    }


def _rel_stats(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    d = np.abs(got - ref)
    scale = max(float(np.max(np.abs(ref))), 1e-30)
    return {"max_abs_over_max": float(d.max() / scale), "mean_abs_over_max": float(d.mean() / scale),
            "cells_rel_above_1e-5": float(np.mean(d > 1e-5 * np.maximum(np.abs(ref), 1e-30)))}


def parity_viewshed(ske, grid, s1, farm, threads, fp64_sectors=12):
    """The GPU map of the CPU-baseline DEM and sectors against the oracle:
    relative error vs the timed multi-thread farm (whose sum order over
    threads varies), bit-exactness vs a one-thread oracle run on two sectors,
    and both maps against the FP64 restatement of kernelV3 on the first
    `fp64_sectors` sectors."""
    import oracle
    with ske.SkewEngine(grid) as e:
        gpu = e.run(ske.KERNEL_SDEM, 0, s1)
        gpu2 = e.run(ske.KERNEL_SDEM, 0, 2)
        k = min(fp64_sectors, s1)
        gpuk = e.run(ske.KERNEL_SDEM, 0, k)
    one, _ = oracle.execute(oracle.SDEM, grid, nthreads=1, s0=0, s1=2)
    f64 = oracle.execute_sdem_f64(grid, nthreads=threads, s0=0, s1=k)
    farmk, _ = oracle.execute(oracle.SDEM, grid, nthreads=threads, s0=0, s1=k)
    st = _rel_stats(gpu, farm)
    return {"sample": "%dx%d DEM of the CPU baseline, sectors 0-%d" % (grid.shape[0], grid.shape[1], s1 - 1),
            "max_rel_err_vs_oracle_farm": st["max_abs_over_max"],
            "cells_rel_err_above_1e-5_vs_oracle_farm": st["cells_rel_above_1e-5"],
            "bit_exact_vs_one_thread_oracle_sectors_0_1": bool(np.array_equal(gpu2, one)),
            "fp64_oracle": {"sectors": "0-%d" % (k - 1), "gpu": _rel_stats(gpuk, f64),
                            "fp32_oracle": _rel_stats(farmk, f64)}}


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
    # One eager step with HIP events around every launch: the viewshed
    # kernel's share of the step.
    batch.set_option(_lib.OPT_PROFILE, 1)
    batch.profile_reset()
    batch.run(ske.KERNEL_SDEM, 0, 180, ske.POVH_DEFAULT)
    eager_op_ms, eager_op_cnt = batch.profile("op_sdem")
    eager_step_ms = sum(batch.profile(k)[0] for k in ("op_sdem", "deskew", "reduce", "fill"))
    _progress("eager step: op_sdem %.1f ms over %d launches, %.1f ms of kernels" % (eager_op_ms, eager_op_cnt,
                                                                                    eager_step_ms))
    # The timed steps replay the whole step captured once in a hipGraph;
    # HIP events bracket each replay on the batch stream.
    batch.set_option(_lib.OPT_GRAPH, 1)

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
    g_ms, g_cnt = batch.profile("graph")
    # One more (untimed) step with the kernels counting the ray steps their
    # lanes execute: the scans stop once no later target can be visible.
    batch.set_option(_lib.OPT_GRAPH, 0)
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
    # The time is the graph replays' (HIP events over the timed region),
    # which also hold the deskew and reduce (1-2% of the step): the kernel
    # figure is a slight underestimate.
    k_ms = g_ms
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
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_PUBLISHED, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_PUBLISHED, "traffic": _traffic("op_sdem"),
                     "traffic_source": "constant from profiles/op_sdem_batch32_traffic.md (rocprofv3 --pmc "
                                       "FETCH_SIZE + WRITE_SIZE of this command, copy-calibrated), not measured "
                                       "by this run",
                     "peak_non_fma_issue": PEAK_VALU_TOPS, "frac_of_non_fma_issue_peak": achieved / PEAK_VALU_TOPS,
                     "kernel": "op_sdem", "timed": "hipGraph replays of the whole step (op_sdem is %.1f%% of the "
                               "eager step's kernel time)" % (100.0 * eager_op_ms / max(eager_step_ms, 1e-9)),
                     "kernel_ms_total": k_ms, "graph_launches": g_cnt,
                     "op_sdem_launches_per_step": eager_op_cnt, "op_sdem_eager_ms_per_step": eager_op_ms,
                     "ray_steps_per_s": steps_total * args.steps / (k_ms / 1e3),
                     "reference_step_equivalent_frac": OPS_PER_STEP * steps_total * args.steps / (k_ms / 1e3)
                                                       / 1e12 / PEAK_FP32_PUBLISHED,
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


def _cpu_threads():
    import oracle
    return oracle.default_threads()


def extra_c1_viewshed(ske, synth, cpu=True):
    """Config 1: one 500^2 DEM (H=0.8, 0-1500 m), 180 sectors.  CPU: the
    oracle farm's full run; errors vs that farm, vs a one-thread oracle run
    (bit-exact, three sectors) and vs the FP64 restatement."""
    g = synth.dem(500, 0.8, 1, 1500.0)
    _, steps = ske.work(500, 500)
    with ske.SkewEngine(g) as e:
        t = _timed(lambda: e.run(ske.KERNEL_SDEM, copy_out=False), e.synchronize, 20)
        out = e.run(ske.KERNEL_SDEM)
        out3 = e.run(ske.KERNEL_SDEM, 44, 47)
    res = {"workload": "500x500 fBm DEM, 180 sectors", "ms": t * 1e3,
           "Mpts_sectors_per_s": 500 * 500 * 180 / 1e6 / t, "ray_steps_per_s": steps / t}
    if cpu:
        import oracle
        thr = _cpu_threads()
        farm, secs = oracle.execute(oracle.SDEM, g, nthreads=thr)
        one, _ = oracle.execute(oracle.SDEM, g, nthreads=1, s0=44, s1=47)
        f64 = oracle.execute_sdem_f64(g, nthreads=thr)
        res["cpu_baseline"] = {"value": 500 * 500 * 180 / 1e6 / secs, "unit": "Mpts*sectors/s", "s": secs,
                               "cores": thr, "kind": "port", "sample": "full run"}
        res["speedup_vs_cpu"] = secs / t
        res["error_vs_oracle"] = {"max_rel_err_vs_oracle_farm": _rel_stats(out, farm)["max_abs_over_max"],
                                  "bit_exact_vs_one_thread_oracle_sectors_44_46": bool(np.array_equal(out3, one)),
                                  "fp64_oracle": {"gpu": _rel_stats(out, f64), "fp32_oracle": _rel_stats(farm, f64)}}
    return res


def extra_c2_radon(ske, synth, cpu=True):
    """Config 2: 512^2 ellipse phantom, 180 angles -- the reference-semantics
    skew-space Radon/FBP operator, and the standard 725-bin projector + Ram-Lak
    FBP, in slices/s.  CPU: the oracle farm (full run; local transform on a
    sector subset scaled by cells) and the FP64 numpy restatement of the
    standard geometry (projection on an angle subset scaled by angles)."""
    from skewengine_amd.fbp import FBP
    img = synth.ellipse_phantom(512, seed=2)
    res = {}
    with ske.SkewEngine(img) as e:
        t = _timed(lambda: e.run(ske.KERNEL_RADON, copy_out=False), e.synchronize, 20)
        out = e.run(ske.KERNEL_RADON)
    res["skew_radon_fbp"] = {"workload": "512x512 phantom, 180 skew sectors (kernelRADN semantics)",
                             "ms_per_slice": t * 1e3, "slices_per_s": 1.0 / t}
    if cpu:
        import oracle
        thr = _cpu_threads()
        ref, secs = oracle.execute(oracle.RADON, img, nthreads=thr)
        res["skew_radon_fbp"].update({
            "cpu_baseline": {"value": 1.0 / secs, "unit": "slices/s", "s": secs, "cores": thr, "kind": "port",
                             "sample": "full run"},
            "speedup_vs_cpu": secs / t,
            "max_rel_err_vs_oracle": _rel_stats(out, ref)["max_abs_over_max"]})
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
        lrt = e.run(ske.KERNEL_RADON, 0, 8)
    res["local_radon"] = {"workload": "512x512 phantom, 180 skew sectors, local transform (pointRadon, LRT=true)",
                          "ms_per_slice": t * 1e3, "cells_per_s": cells / t,
                          "fp64_TFLOPs_if_6240_per_cell": 6240 * cells / t / 1e12}
    if cpu:
        import oracle
        thr = _cpu_threads()
        ref, secs = oracle.execute(oracle.RADON_LOCAL, img, nthreads=thr, s0=0, s1=8)
        c8, _ = ske.work(512, 512, 0, 8)
        st = _rel_stats(lrt, ref)
        res["local_radon"].update({
            "cpu_baseline": {"value": c8 / secs, "unit": "cells/s", "s": secs, "cores": thr, "kind": "port",
                             "sample": "sectors 0-7 of 180 (%d cells)" % c8},
            "speedup_vs_cpu": (cells / t) / (c8 / secs),
            "error_vs_oracle_sectors_0_7": {"max_abs_over_max": st["max_abs_over_max"],
                                            "cells_rel_above_1e-5": st["cells_rel_above_1e-5"]}})
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
        sino = f.download(1, 1)[0]  # the first phantom's sinogram (SINO)
        f.reconstruct(1)
        rec = f.download(3, 1)[0]  # its reconstruction (RECON)
    res["standard_radon_plus_fbp"] = {
        "workload": "512x512 phantoms, 180 angles x 725 bins, project + Ram-Lak filter + backproject, batch %d" % batch,
        "ms_per_slice": t * 1e3 / batch, "slices_per_s": batch / t, "kernel_ms_per_slice": k}
    if cpu:
        from oracle import std_fbp
        idx = list(range(0, 180, 10))
        t0 = time.perf_counter()
        ref_sino = std_fbp.project(imgs[0], 180, 725, angle_idx=idx)
        t_proj = (time.perf_counter() - t0) * 180 / len(idx)
        t0 = time.perf_counter()
        ref_rec = std_fbp.fbp(sino.astype(np.float64), 512)
        t_fbp = time.perf_counter() - t0
        res["standard_radon_plus_fbp"].update({
            "cpu_baseline": {"value": 1.0 / (t_proj + t_fbp), "unit": "slices/s", "s": t_proj + t_fbp, "cores": 1,
                             "kind": "port", "sample": "FP64 numpy restatement, one slice; projection timed on "
                                                       "18 of 180 angles and scaled by 10"},
            "speedup_vs_cpu": (t_proj + t_fbp) / (t / batch),
            "max_rel_err_vs_fp64": {"projection_18_angles": _rel_stats(sino[idx], ref_sino)["max_abs_over_max"],
                                    "fbp": _rel_stats(rec, ref_rec)["max_abs_over_max"]}})
    return res


def extra_c3_cepstrum(ske, synth, count, cpu=True):
    """Config 3: batched cepstrum blur estimation over `count` 1024^2 images.
    CPU: the oracle farm on the first image over a sector subset, scaled by
    cells; the GPU map of the same image and sectors against it."""
    from skewengine_amd import _lib
    imgs = [synth.blurred_texture(1024, 0.6, seed=i)[0] for i in range(count)]
    img0 = imgs[0]
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
    res = {"workload": "%d x 1024^2 fBm (H=0.6) + motion blur + noise, 180 sectors, batched" % count,
           "s": t, "cell_sectors_per_s": total / t, "op_blur_ms": k_ms, "op_blur_launches": k_cnt,
           "op_blur_TFLOPs_fp64_at_2848_per_cell": 2848 * total / (k_ms / 1e3) / 1e12}
    if cpu:
        import oracle
        thr = _cpu_threads()
        k = 12
        ref, secs = oracle.execute(oracle.CEPSTRUM, img0, nthreads=thr, s0=0, s1=k)
        ck, _ = ske.work(1024, 1024, 0, k)
        with ske.SkewEngine(img0) as e:
            out = e.run(ske.KERNEL_CEPSTRUM, 0, k)
        res.update({"cpu_baseline": {"value": ck / secs, "unit": "cell-sectors/s", "s": secs, "cores": thr,
                                     "kind": "port", "sample": "image 0, sectors 0-%d of 180 (%d cell-sectors)"
                                                               % (k - 1, ck)},
                    "speedup_vs_cpu": (total / t) / (ck / secs),
                    "map_vs_oracle_sectors_0_%d" % (k - 1): {"cells_equal_frac": float(np.mean(out == ref)),
                                                              "cells_differing": int(np.sum(out != ref))}})
    return res


def extra_c5_learned(batch=64, cpu=True):
    """Config 5 at both channel counts: 32 hidden channels (the configured
    model) and 16 (the faster one); see c5_learned_once."""
    res = {"c32": c5_learned_once(batch, 32, cpu), "c16": c5_learned_once(batch, 16, False)}
    return res


def c5_learned_once(batch, channels, cpu):
    """Config 5: learned sparse-view FBP on 512^2 phantoms, 60 angles, 725
    bins, Poisson noise at I0 = 1e5, batch 64: time of plain Ram-Lak FBP vs
    the learned reconstruction (filter + back-projection + CNN) on the same
    sinograms, and held-out PSNR of both with the committed parameters.
    CPU: plain FBP of one sinogram by the FP64 numpy restatement, which also
    gives the GPU's plain-FBP error."""
    from skewengine_amd import _lib
    from skewengine_amd import learned as lm
    from skewengine_amd.fbp import FBP
    path = lm.default_weights(512, 60, 725, channels=channels)
    res = {"workload": "%d x 512^2 ellipse phantoms, 60 angles x 725 bins, Poisson I0=1e5, batch inference, "
                       "4-layer residual CNN with %d hidden channels" % (batch, channels), "channels": channels}
    with FBP(512, 60, 725, max_batch=batch) as f, lm.LearnedFBP(f, channels=channels) as m:
        if path:
            m.load(path)
        res["weights"] = os.path.basename(path) if path else None
        m.make_batch(batch, seed=lm.HELD_OUT_SEED)
        f.reconstruct(batch)
        f.synchronize()
        sino0 = f.download(1, 1)[0]  # SINO of the first phantom
        rec0 = f.download(3, 1)[0]  # its plain FBP (RECON)
        # both paths warmed up, then timed in alternating rounds of 10 calls
        # each (same clocks for both); the median round of each is reported
        for _ in range(10):
            f.reconstruct(batch)
            m.reconstruct(batch)
        f.synchronize()
        rounds_fbp, rounds_l = [], []
        for _ in range(5):
            rounds_fbp.append(_timed(lambda: f.reconstruct(batch), f.synchronize, 10))
            rounds_l.append(_timed(lambda: m.reconstruct(batch), f.synchronize, 10))
        t_fbp, t_l = float(np.median(rounds_fbp)), float(np.median(rounds_l))
        f.set_profile(True)
        f.profile_reset()
        for _ in range(10):
            m.reconstruct(batch)
        f.synchronize()
        cnn_ms = f.profile("cnn_forward")[0] / 10
        f.set_profile(False)
        c = channels
        flop = 2 * 9 * (1 * c + c * c + c * c + c * 1) * 512 * 512 * batch
        res.update({"timing": "median of 5 alternating rounds of 10 calls of each path, after 10 warm-up calls of each",
                    "fbp_ms_per_slice": t_fbp * 1e3 / batch, "learned_ms_per_slice": t_l * 1e3 / batch,
                    "time_ratio": t_l / t_fbp, "cnn_ms_per_slice": cnn_ms / batch,
                    "cnn_TFLOPs": flop / (cnn_ms / 1e3) / 1e12,
                    "cnn_frac_of_bf16_dense_peak": flop / (cnn_ms / 1e3) / 2.5e15})
        if path:
            p_fbp, p_l = m.evaluate(64, batch)
            res.update({"psnr_fbp_db": float(p_fbp.mean()), "psnr_learned_db": float(p_l.mean()),
                        "psnr_gain_db": float((p_l - p_fbp).mean()), "held_out": 64})
    if cpu:
        from oracle import std_fbp
        t0 = time.perf_counter()
        ref = std_fbp.fbp(sino0.astype(np.float64), 512)
        t_cpu = time.perf_counter() - t0
        res.update({"cpu_baseline_plain_fbp": {"value": 1.0 / t_cpu, "unit": "slices/s", "s": t_cpu, "cores": 1,
                                               "kind": "port", "sample": "FP64 numpy restatement, one slice"},
                    "plain_fbp_max_rel_err_vs_fp64": _rel_stats(rec0, ref)["max_abs_over_max"]})
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
        cpu = not args.no_cpu_baseline
        fn = {"c1": lambda: extra_c1_viewshed(ske, synth, cpu), "c2": lambda: extra_c2_radon(ske, synth, cpu),
              "c3": lambda: extra_c3_cepstrum(ske, synth, args.cepstrum_images, cpu),
              "c5": lambda: extra_c5_learned(cpu=cpu)}
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
        cpu = not args.no_cpu_baseline
        line["extra"] = {}
        for key, fn in (("c1_viewshed_500", lambda: extra_c1_viewshed(ske, synth, cpu)),
                        ("c2_radon_fbp_512", lambda: extra_c2_radon(ske, synth, cpu)),
                        ("c3_cepstrum_batch", lambda: extra_c3_cepstrum(ske, synth, args.cepstrum_images, cpu)),
                        ("c5_learned_sparse_fbp", lambda: extra_c5_learned(cpu=cpu))):
            line["extra"][key] = fn()
            _progress("extra %s done" % key)
    if not args.no_cpu_baseline:
        sample, line["cpu_baseline"] = cpu_baseline_viewshed(steps_total, pts_sectors, args.cpu_seconds)
        line["speedup_vs_cpu"] = line["value"] / line["cpu_baseline"]["value"]
        _progress("parity check on the CPU-baseline sample")
        line["parity"] = parity_viewshed(ske, sample["grid"], sample["s1"], sample["farm"], sample["threads"])
    print(json.dumps(line))


if __name__ == "__main__":
    main()
