#!/usr/bin/env python3
"""Throughput of the multi-level PCA block-matching denoise on one MI355X.

One step = one whole ``PyramidProcessor::denoise`` call (5-level pyramid,
5x5 patches -> 8 PCA components, 9x9 search window) over one synthetic 12 MP
frame already resident in HBM, at the iPhone 14 Wide high-ISO denoise preset
(chroma multiplier 0.25), where the denoise raises the frame's PSNR.  The
reported form is the precision form (FP16 MFMA projection and FP16 distance
products, FP32 accumulation); the plain FP32 form is timed beside it on the
same frame, and the PSNR loss of the reported form against the FP64 oracle is
measured on that frame.  Prints ONE JSON line with the throughput in Mpix/s of
the base level (K timed steps with no instrumentation; per-step median, min and
max beside the mean), the roofline of the dominant kernel (block matching,
timed over 5 further steps with HIP events the library records around each
stage on the stream it launches on), matrix-core utilisation, and the CPU
oracle timed on this host's cores on the same frame.  Every timed leg starts
after a time-based warm-up, and comparisons between forms use median steps.

    python bench.py [--steps 20] [--warmup 3] [--precision fp16] [--no-cpu-baseline]
    python bench.py --workload burst32      # only the 32-frame mixed-size burst (batched form)
    python bench.py --workload copy         # 256 MiB device copy: the HBM counter calibration

Beside the denoise the default run times the raw front end that produces its
inputs from the frame's Bayer mosaic (RawFrontEnd, "front_end"; --no-front-end
to skip), burst fusion (registerAndFuse of the frame, FindHomography on 1500
matches: "registration"; --no-registration), the reference-CPU configuration
(1024x1024, 4 levels, 7x7: both forms and the CPU oracle, "c1"; --no-c1), both
forms with the reference's own 21x21 window on the same frame ("reference_window";
--no-wide) and the 32-frame burst of the batched form ("burst32"; --no-burst).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

METRIC = "Mpix/s, 12MP 5-level PCA denoise; FP16-MFMA vs FP32 at ≤0.05dB PSNR loss"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (width, height, levels, radius, shot gain a, read sigma b)
    "12mp": (4032, 3024, 5, 4, 4e-3, 2e-3),
    "1mp": (1024, 1024, 4, 3, 2e-3, 1e-3),
    "48mp": (8064, 6048, 5, 4, 4e-3, 2e-3),
}
BURST_SIZES = [(3264, 2448), (4032, 3024), (5472, 3648), (8064, 6048)]
# HBM bytes per level pixel that the block-matching launch of each level moves on the 12 MP workload
# (9x9 window), levels 0..4: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by
# what the counter reads for a plain 256 MiB copy (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000).
MEASURED_BM_TRAFFIC = {
    ("12mp", "fp16"): [32.5, 33.0, 34.1, 36.7, 42.0],
    ("12mp", "fp32"): [84.2, 65.9, 67.8, 71.6, 80.2],
}
# Wave-level VALU instructions per block-matching dispatch on the same workload, levels 0..4:
# rocprofv3 --pmc SQ_INSTS_VALU, in millions.
MEASURED_BM_VALU = {
    ("12mp", "fp16"): [391.316, 97.987, 24.964, 6.781, 1.712],
    ("12mp", "fp32"): [531.124, 132.795, 33.733, 8.654, 2.175],
}
# Matrix-core work per dispatch on the same workload, levels 0..4 (rocprofv3 --pmc
# SQ_INSTS_VALU_MFMA_MOPS_F16 / _F64, which count flops / 512): the FP16 projection inside the precision
# form's block matching (v_mfma_f32_32x32x16_f16) and the FP64 patch moments (v_mfma_f64_16x16x4f64,
# levels 0-3 in the fused k_subtract_moments, level 4 in k_image_moments; levels 0-2 share a grid, so
# the counter pass gives their mean, 250172, and the per-level values are its tiles' groups x 3 MFMAs x 4).
MEASURED_MFMA_MOPS = {
    ("12mp", "fp16"): {"block_match_f16": [91445760, 22982400, 5898240, 1966080, 491520],
                       "moments_f64": [571536, 142884, 36096, 9024, 4608]},
}
MFMA_PEAK_TFLOPS = {"f16": 2500.0, "bf16": 2500.0, "f64": 78.6}
DTYPE = {"fp32": "f32", "fp16": "f16", "bf16": "bf16"}
# Vector ALU: 256 CUs x 4 SIMDs x 16 lanes at 2.4 GHz (a wave64 instruction takes 4 cycles).
VALU_PEAK_LANE_OPS = 256 * 4 * 16 * 2.4e9
# Issue slots per (pixel, window position) of the reference's inner loop: ~39 VALU operations and
# two exponentials at quarter rate.
MODEL_SLOTS_PER_PAIR = 47
# Host CPU threads one GPU's job may use on the benchmark machines (their CPU share per GPU).
BOX_CPU_SHARE = 16


def mfma_block(workload, precision, stages_mean5, levels):
    """Matrix-core utilisation: measured MFMA flops per dispatch over the live per-level stage durations
    (block matching for the FP16 projection; for the FP64 moments the subtract, which forms them below
    the top level, and the moment kernel of the top level), against the dense MFMA peaks."""
    m = MEASURED_MFMA_MOPS.get((workload, precision))
    if m is None:
        return None
    out = {}
    for name, key, dt in (("block_match_f16_projection", "block_match", "f16"), ("moments_f64", "moments", "f64")):
        mops = m["block_match_f16" if dt == "f16" else "moments_f64"]
        ms = list(stages_mean5.get(key, [0.0] * levels))
        if dt == "f64":
            sub = stages_mean5.get("subtract", [0.0] * levels)
            ms = [sub[i] if i < levels - 1 else ms[i] for i in range(levels)]
        flops = [512.0 * x for x in mops[:levels]]
        rates = [f / (t / 1e3) / 1e12 if t > 0 else None for f, t in zip(flops, ms)]
        tot_t = sum(ms[:levels])
        agg = sum(flops) / (tot_t / 1e3) / 1e12 if tot_t > 0 else None
        out[name] = {"tflops_per_level": [round(r, 3) if r else None for r in rates],
                     "tflops": round(agg, 3) if agg else None, "peak_tflops": MFMA_PEAK_TFLOPS[dt],
                     "frac": round(agg / MFMA_PEAK_TFLOPS[dt], 5) if agg else None,
                     "gflop_per_frame": round(sum(flops) / 1e9, 3)}
    return out


def algorithmic_bytes_per_base_pixel(levels: int) -> float:
    """Stage DAG of the denoise with every image 3 channels x FP32, read once and
    written once (features recomputed in-tile, gradient unused): per level pixel
    72.125 B on levels < L-1 and 27.125 B on the coarsest, weighted by 4^-i."""
    return sum(72.125 * 4.0 ** -i for i in range(levels - 1)) + 27.125 * 4.0 ** -(levels - 1)


def cpu_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def psnr(a, b) -> float:
    import numpy as np
    mse = float(np.mean((a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)) ** 2))
    return 10.0 * math.log10(1.0 / mse)


def step_stats(step_ms):
    return {"median_ms": round(statistics.median(step_ms), 4), "min_ms": round(min(step_ms), 4),
            "max_ms": round(max(step_ms), 4), "mean_ms": round(sum(step_ms) / len(step_ms), 4)}


def settle(fn, warmup, min_s=0.3):
    """W warmup calls, then more until at least min_s of wall time has passed with the GPU busy, so that
    a timed leg that follows an idle gap (allocation, host work) does not start at ramping clocks."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t = time.perf_counter()
    while time.perf_counter() - t < min_s:
        for _ in range(4):
            fn()
        torch.cuda.synchronize()


def time_single(gcam, torch, w, h, levels, radius, precision, bm_rows, ycc, params, nlf, steps, warmup, dev):
    """K timed steps of one processor, then a few steps with the library's stage
    events on (HIP events it records around every stage on its stream) for the
    per-stage breakdown; the timed steps carry no stage events."""
    stream = torch.cuda.current_stream()
    proc = gcam.PyramidProcessor(w, h, levels=levels, search_radius=radius, precision=precision, device=dev,
                                 bm_rows_per_thread=bm_rows)
    settle(lambda: proc.denoise(params, ycc, None, nlf), warmup)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record(stream)
    for k in range(steps):
        proc.denoise(params, ycc, None, nlf)
        evs[k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = [evs[k].elapsed_time(evs[k + 1]) for k in range(steps)]
    proc.set_profiling(True)
    runs = []
    for _ in range(5):
        proc.denoise(params, ycc, None, nlf)
        runs.append(proc.stage_times())
    stages = runs[-1]
    mean5 = {k: [sum(r[k][i] for r in runs) / len(runs) for i in range(levels)] for k in stages}
    proc.set_profiling(False)
    out = proc.denoisedImagePyramid[0].clone()
    proc.close()
    total_ms = evs[0].elapsed_time(evs[-1])
    return {"ms_per_step": total_ms / steps, "step_stats": step_stats(step_ms),
            "wall_ms_per_step": wall * 1e3 / steps, "stages": stages, "stages_mean5": mean5,
            "stage_sum_ms": sum(sum(v) for k, v in mean5.items() if k != "moments"), "out": out}


# HBM bytes per pixel of the front end as it runs, each pass's inputs read once
# and outputs written once (FP32 images, 16-bit raw): scale 2 + 4, gradient
# (Sobel fused into the blur) 4 + 8, demosaic through despeckle in one tiled
# pass 4 + 8 + 16.  The passes are bound by vector-ALU work and latency, not
# HBM; the fraction below is of the HBM roofline all the same.
FRONT_END_BYTES_PER_PX = 6 + 12 + 28


def front_end_params(gcam, a, b, despeckle_nlf):
    """The synthetic frame's raw model: 12-bit DN in 16-bit words, sRGB primaries
    (rgb_cam = I), white balance 1; raw noise var = a x + b^2 in [0, 1] units,
    mapped through scaleRawData's 0.9 x + 0.1."""
    return gcam.FrontEndParameters(
        bayerPattern=2, black_level=0.0, lensShadingCorrection=0.0, scale_mul=[65535.0 / 4095.0] * 4,
        rawNlf=((0.81 * b * b - 0.09 * a,) * 4, (0.9 * a,) * 4), rgb_cam=[[1, 0, 0], [0, 1, 0], [0, 0, 1]],
        highlightClip=1.0, despeckleNlf=despeckle_nlf)


def time_front_end(gcam, torch, raw, w, h, params, steps, warmup, cpu_baseline, threads):
    """RawFrontEnd.run (raw -> YCbCr + gradient for the denoise) timed like the denoise."""
    stream = torch.cuda.current_stream()
    fe = gcam.RawFrontEnd(w, h)
    ycc = torch.empty((h, w, 4), dtype=torch.float32, device=raw.device)
    grad = torch.empty((h, w, 2), dtype=torch.float32, device=raw.device)
    settle(lambda: fe.run(raw, params, ycc, grad), warmup)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(steps):
        fe.run(raw, params, ycc, grad)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    fe.close()
    gbs = FRONT_END_BYTES_PER_PX * w * h / (ms / 1e3) / 1e9
    out = {"ms_per_frame": round(ms, 4), "value": round(w * h / 1e6 / (ms / 1e3), 2), "unit": "Mpix/s",
           "roofline": {"bound": "valu", "bytes_per_px": FRONT_END_BYTES_PER_PX, "achieved": round(gbs, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)}}
    if cpu_baseline:
        import numpy as np

        import oracle
        cw, ch = min(w, 2048), min(h, 1536)
        crop = np.ascontiguousarray(raw[:ch, :cw].cpu().numpy().view(np.uint16))
        p = params
        t = time.perf_counter()
        oracle.raw_front_end(crop, p.bayerPattern, p.scale_mul, p.black_level, p.lensShadingCorrection,
                             p.rawNlf[0], p.rawNlf[1], np.array(p.rgb_cam, np.float32), p.highlightClip,
                             p.despeckleNlf)
        cpu_s = time.perf_counter() - t
        out["cpu_baseline"] = {"value": round(cw * ch / 1e6 / cpu_s, 3), "unit": "Mpix/s", "cores": 1, "kind": "port",
                               "sample": f"{cw}x{ch} crop of the frame, FP32 C++ oracle, 1 thread, {cpu_s:.2f} s"}
    return out


REGISTER_AND_FUSE_BYTES_PER_PX = 48  # fused RGBA32F read + warped input read (one texel per pixel) + write


def time_registration(gcam, torch, ycc, w, h, steps, warmup, cpu_baseline):
    """Burst fusion beside the denoise: registerAndFuse of the frame into a fused frame (in place,
    as pyramidFusion.cpp:345 calls it) under a small rotation + translation, and FindHomography on
    1500 matches (30% outliers) with the reference's threshold 1 and 2000 iterations."""
    import numpy as np

    stream = torch.cuda.current_stream()
    c, s_ = np.cos(0.002), np.sin(0.002)
    H = np.array([[c, -s_, 1.7], [s_, c, -2.3], [1e-7, -2e-7, 1.0]], np.float32)
    fused = ycc.clone()
    settle(lambda: gcam.registerAndFuse(fused, ycc, fused, H, 2), warmup)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(steps):
        gcam.registerAndFuse(fused, ycc, fused, H, 2)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    gbs = REGISTER_AND_FUSE_BYTES_PER_PX * w * h / (ms / 1e3) / 1e9
    rng = np.random.default_rng(77)
    n = 1500
    p = rng.uniform(0, w / 2, (n, 2))
    ph = np.c_[p, np.ones(n)] @ H.astype(np.float64).T
    q = ph[:, :2] / ph[:, 2:] + 0.05 * rng.standard_normal((n, 2))
    bad = rng.random(n) < 0.3
    q[bad] += rng.uniform(5, 50, (bad.sum(), 2))
    m = np.c_[p, q].astype(np.float32)
    md = torch.from_numpy(m).to(ycc.device)
    for _ in range(2):
        gcam.FindHomography(md, 1.0, 2000)
    t = time.perf_counter()
    fits = 10
    for _ in range(fits):
        gcam.FindHomography(md, 1.0, 2000)
    fit_ms = (time.perf_counter() - t) / fits * 1e3
    out = {"register_and_fuse": {
               "ms_per_frame": round(ms, 4), "value": round(w * h / 1e6 / (ms / 1e3), 2), "unit": "Mpix/s",
               "roofline": {"bound": "hbm", "bytes_per_px": REGISTER_AND_FUSE_BYTES_PER_PX, "achieved": round(gbs, 1),
                            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)}},
           "find_homography": {"ms_per_fit": round(fit_ms, 3), "matches": n, "iterations": 2000,
                               "note": "host-inclusive: scratch allocation, 2 launches, result copy"}}
    if cpu_baseline:
        import oracle
        cw, ch = min(w, 2048), min(h, 1536)
        img = np.ascontiguousarray(ycc[:ch, :cw].cpu().numpy())
        t = time.perf_counter()
        oracle.register_and_fuse(img, img, H, 2)
        cpu_s = time.perf_counter() - t
        out["register_and_fuse"]["cpu_baseline"] = {
            "value": round(cw * ch / 1e6 / cpu_s, 3), "unit": "Mpix/s", "cores": 1, "kind": "port",
            "sample": f"{cw}x{ch} crop of the frame, FP32 C++ oracle, 1 thread, {cpu_s:.2f} s"}
        t = time.perf_counter()
        oracle.find_homography(m, 1.0, 2000, 0)
        cpu_s = time.perf_counter() - t
        out["find_homography"]["cpu_baseline"] = {
            "ms_per_fit": round(cpu_s * 1e3, 1), "cores": 1, "kind": "port",
            "sample": "the same 1500 matches, 2000 hypotheses, C++ oracle (reference algorithm), 1 thread"}
    return out


def preset_params(gcam, levels):
    """iPhone 14 Wide denoise preset at high ISO (iPhone14WideCalibration.cpp:75-111 at nlf_alpha = 1: luma
    3, 1.5, 1, 1, 1 x 2; chromaBoost 8, 4, 2, 2, 2; gradientBoost 2, 1, 0..; gradientThreshold 2, 2, 1..;
    sharpening 1) with the chroma multiplier at 0.25: the synthetic scene's chroma carries the same
    high-frequency texture as its luma (independent simplex octaves per RGB channel), which the preset's
    chroma multiplier of 1 would average away.  At this point the denoise raises the frame's PSNR on all
    three channels (no gradient image: the gradient terms are inert)."""
    lm, cb = [3.0, 1.5, 1.0, 1.0, 1.0], [8.0, 4.0, 2.0, 2.0, 2.0]
    gb, gt = [2.0, 1.0, 0.0, 0.0, 0.0], [2.0, 2.0, 1.0, 1.0, 1.0]
    return [gcam.DenoiseParameters(luma=2.0 * lm[min(i, 4)], chroma=0.25, chromaBoost=cb[min(i, 4)],
                                   gradientBoost=gb[min(i, 4)], gradientThreshold=gt[min(i, 4)], sharpening=1.0)
            for i in range(levels)]


PRESET_NAME = ("iPhone 14 Wide high-ISO preset (luma 6, 3, 2, 2, 2; chromaBoost 8, 4, 2, 2, 2), chroma 0.25, "
               "no gradient image")


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def calibrate(gcam, torch, w, h, levels, radius, params, ycc, dev):
    """MeasureYCbCrNLF on every level of the frame (the in-pipeline calibration); returns the noise model
    and the fit's own time (a calibrating call minus a plain one)."""
    stream = torch.cuda.current_stream()
    nlf = [((1e-8,) * 3, (1e-8,) * 3) for _ in range(levels)]
    cal = gcam.PyramidProcessor(w, h, levels=levels, search_radius=radius, device=dev)
    cal.denoise(params, ycc, None, nlf, calibrateFromImage=True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    cal.denoise(params, ycc, None, list(nlf), calibrateFromImage=True)
    e1.record(stream)
    torch.cuda.synchronize()
    calib_ms = e0.elapsed_time(e1)
    e0.record(stream)
    cal.denoise(params, ycc, None, nlf)
    e1.record(stream)
    torch.cuda.synchronize()
    fit_ms = max(0.0, calib_ms - e0.elapsed_time(e1))
    cal.close()
    return nlf, fit_ms


def single_frame(args) -> dict:
    import numpy as np
    import torch

    import glasscamerapublic_amd as gcam

    w, h, levels, radius, a, b = WORKLOADS[args.workload]
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)

    ycc, raw, clean = gcam.synth_frame(w, h, a, b, scene_seed=1234, noise_seed=5678, device=dev, want_raw=True,
                                       want_clean=True)
    params = preset_params(gcam, levels)
    # Noise model: MeasureYCbCrNLF on every level, once per frame, outside the timed region.
    nlf, nlf_fit_ms = calibrate(gcam, torch, w, h, levels, radius, params, ycc, dev)

    prim = time_single(gcam, torch, w, h, levels, radius, args.precision, args.bm_rows, ycc, params, nlf,
                       args.steps, args.warmup, dev)
    ref32 = None
    if args.precision != "fp32" and not args.no_compare:
        ref32 = time_single(gcam, torch, w, h, levels, radius, "fp32", args.bm_rows, ycc, params, nlf, args.steps,
                            args.warmup, dev)

    mpix = w * h / 1e6
    ms_per_step = prim["ms_per_step"]
    value = mpix / (ms_per_step / 1e3)

    # Dominant kernel: block matching, one launch per level.  Algorithmic bytes
    # per launch = 24 B (3 ch x FP32 in + 3 ch x FP32 out) per level pixel.
    level_px = [(w >> i) * (h >> i) for i in range(levels)]
    bm_ms = prim["stages_mean5"]["block_match"]
    bm_avg_ms = sum(bm_ms) / levels
    bm_bytes_per_launch = 24.0 * sum(level_px) / levels
    bm_gbs = bm_bytes_per_launch / (bm_avg_ms / 1e3) / 1e9
    coef = MEASURED_BM_TRAFFIC.get((args.workload, args.precision))
    traffic = None if coef is None else round(sum(c * n for c, n in zip(coef, level_px)) / levels)
    pairs_per_launch = (2 * radius + 1) ** 2 * sum(level_px) / levels
    model_rate = MODEL_SLOTS_PER_PAIR * pairs_per_launch / (bm_avg_ms / 1e3)
    valu = {"kernel": f"k_block_match{'' if args.precision == 'fp32' else '_lp'}<{radius}>",
            "peak_lane_ops_per_s": VALU_PEAK_LANE_OPS,
            "model_slots_per_pair": MODEL_SLOTS_PER_PAIR,
            "model_frac": round(model_rate / VALU_PEAK_LANE_OPS, 4),
            "measured_lane_ops_per_pair": None, "issue_frac": None}
    insts = MEASURED_BM_VALU.get((args.workload, args.precision))
    if insts is not None:
        lane_ops = 64e6 * sum(insts) / levels
        valu["measured_lane_ops_per_pair"] = round(lane_ops / pairs_per_launch, 2)
        valu["issue_frac"] = round(lane_ops / (bm_avg_ms / 1e3) / VALU_PEAK_LANE_OPS, 4)
    b_alg = algorithmic_bytes_per_base_pixel(levels)
    pipeline_gbs = b_alg * w * h / (ms_per_step / 1e3) / 1e9

    cpu, quality = None, None
    if not args.no_cpu_baseline:
        import oracle
        host_img = ycc.cpu().numpy()
        host_clean = clean.cpu().numpy()
        host_nlf = np.array([[list(A), list(B)] for A, B in nlf], np.float64)
        pr = [dict(luma=p.luma, chroma=p.chroma, chromaBoost=p.chromaBoost, gradientBoost=p.gradientBoost,
                   gradientThreshold=p.gradientThreshold, sharpening=p.sharpening) for p in params]
        host_cpus = cpu_cores()
        threads = args.cpu_threads or min(host_cpus, BOX_CPU_SHARE)
        oracle.set_threads(threads)
        t = time.perf_counter()
        oracle.pyramid_denoise(host_img, pr, host_nlf, levels=levels, radius=radius)
        cpu_s = time.perf_counter() - t
        cpu = {"value": round(mpix / cpu_s, 4), "unit": "Mpix/s", "cores": threads, "kind": "port",
               "host_cpus": host_cpus, "cpu_model": cpu_model(),
               "sample": f"one full {w}x{h} frame (the benchmark frame), {levels} levels, "
                         f"{2 * radius + 1}x{2 * radius + 1} window, FP32 C++ oracle, {threads} threads "
                         f"(the host's CPU share for one GPU: at most {BOX_CPU_SHARE} of its {host_cpus} CPUs), "
                         f"{cpu_s:.1f} s"}
        if not args.no_quality:
            ref64 = oracle.pyramid_denoise(host_img, pr, host_nlf, levels=levels, radius=radius,
                                           dtype=np.float64).denoised[0]
            got = prim["out"].cpu().numpy()
            p64, pgpu = psnr(ref64, host_clean), psnr(got, host_clean)
            quality = {"params": PRESET_NAME, "psnr_noisy_db": round(psnr(host_img, host_clean), 4),
                       "psnr_fp64_oracle_db": round(p64, 4), f"psnr_{args.precision}_db": round(pgpu, 4),
                       "psnr_loss_db": round(p64 - pgpu, 4),
                       "max_abs_vs_fp64": float(np.abs(got[..., :3] - ref64[..., :3]).max())}
            if ref32 is not None:
                got32 = ref32["out"].cpu().numpy()
                quality["psnr_fp32_db"] = round(psnr(got32, host_clean), 4)
                quality["fp32_max_abs_vs_fp64"] = float(np.abs(got32[..., :3] - ref64[..., :3]).max())

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic",
        "config": {
            "workload": f"{w}x{h} RGGB 12-bit simplex scene, Poisson-Gaussian a={a:g} b={b:g}, bilinear demosaic "
                        f"-> YCbCr, {levels}-level pyramid, 5x5 patches -> 8 PCA comps, "
                        f"{2 * radius + 1}x{2 * radius + 1} window",
            "width": w, "height": h, "levels": levels, "search_radius": radius, "precision": args.precision,
            "denoise_parameters": PRESET_NAME, "parallelism": "single-gpu",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_block_match{'' if args.precision == 'fp32' else '_lp'}<{radius}>",
            "achieved": round(bm_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(bm_gbs / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "avg_launch_ms": round(bm_avg_ms, 4),
            "bytes_per_launch": int(bm_bytes_per_launch),
        },
        "valu": valu,
        "pipeline": {"B_alg_per_base_px": round(b_alg, 3), "achieved_GBs": round(pipeline_gbs, 2),
                     "hbm_frac": round(pipeline_gbs / HBM_PEAK_GBS, 5)},
        "step_stats": prim["step_stats"],
        "wall_ms_per_step": round(prim["wall_ms_per_step"], 4),
        "stage_sum_ms": round(prim["stage_sum_ms"], 4),
        "nlf_fit_ms": round(nlf_fit_ms, 3),
        "stage_ms_mean5": {k: [round(x, 4) for x in v] for k, v in prim["stages_mean5"].items()},
        "cpu_baseline": cpu,
    }
    mf = mfma_block(args.workload, args.precision, prim["stages_mean5"], levels)
    if mf is not None:
        out["mfma"] = mf
    if ref32 is not None:
        med16, med32 = prim["step_stats"]["median_ms"], ref32["step_stats"]["median_ms"]
        out["fp32_form"] = {"value": round(mpix / (ref32["ms_per_step"] / 1e3), 3),
                            "value_median": round(mpix / (med32 / 1e3), 3),
                            "ms_per_step": round(ref32["ms_per_step"], 4), "step_stats": ref32["step_stats"],
                            "stage_sum_ms": round(ref32["stage_sum_ms"], 4),
                            "speedup_of_reported_form": round(med32 / med16, 4),
                            "speedup_basis": "median step times",
                            "stage_ms_mean5": {k: [round(x, 4) for x in v] for k, v in ref32["stages_mean5"].items()}}
    if quality is not None:
        out["quality"] = quality
    if not args.no_wide and args.workload == "12mp":
        # The reference's own window (const int size = 10, demosaic.metal:853) on the same frame, both forms.
        wide = {"window": "21x21 (radius 10)", "unit": "Mpix/s"}
        for prec in ("fp16", "fp32"):
            r = time_single(gcam, torch, w, h, levels, 10, prec, 0, ycc, params, nlf, min(args.steps, 10),
                            args.warmup, dev)
            wide[prec] = {"value_median": round(mpix / (r["step_stats"]["median_ms"] / 1e3), 2),
                          "step_stats": r["step_stats"],
                          "block_match_ms": [round(x, 4) for x in r["stages_mean5"]["block_match"]]}
        out["reference_window"] = wide
    if not args.no_front_end:
        # The inputs of this denoise from the frame's raw mosaic (front end of RawConverter::demosaic and
        # its despeckle), reported beside the denoise; the despeckle uses the calibrated level-0 model.
        fe_params = front_end_params(gcam, a, b, nlf[0])
        out["front_end"] = time_front_end(gcam, torch, raw, w, h, fe_params, args.steps, args.warmup,
                                          not args.no_cpu_baseline, 1)
    if not args.no_registration:
        out["registration"] = time_registration(gcam, torch, ycc, w, h, args.steps, args.warmup,
                                                not args.no_cpu_baseline)
    del ycc, raw, clean
    torch.cuda.empty_cache()
    if not args.no_c1 and args.workload == "12mp":
        out["c1"] = c1_block(args)
    if not args.no_burst and args.workload == "12mp":
        bst = burst(args, steps=min(args.steps, 10), warmup=min(args.warmup, 2))
        bst.pop("metric")
        for k in ("n_gpus", "higher_is_better", "scaling", "vs_baseline", "data"):
            bst.pop(k, None)
        if ref32 is not None:
            bst["vs_fp32_form"] = round(bst["value_median"] / out["fp32_form"]["value_median"], 4)
            bst["vs_fp32_form_basis"] = "burst median Mpix/s over the FP32 form's median Mpix/s on the 12 MP frame"
        out["burst32"] = bst
    return out


def c1_block(args) -> dict:
    """The reference-CPU configuration: one 1024x1024 frame (a=2e-3, b=1e-3), 4 levels, 7x7 window, both GPU
    forms and the CPU oracle on the same frame (the whole frame, the host's CPU share)."""
    import numpy as np
    import torch

    import glasscamerapublic_amd as gcam

    w, h, levels, radius, a, b = WORKLOADS["1mp"]
    dev = torch.device("cuda", 0)
    ycc, clean = gcam.synth_frame(w, h, a, b, scene_seed=1234, noise_seed=5678, device=dev, want_clean=True)
    params = preset_params(gcam, levels)
    nlf, _ = calibrate(gcam, torch, w, h, levels, radius, params, ycc, dev)
    mpix = w * h / 1e6
    out = {"workload": f"{w}x{h} RGGB 12-bit simplex scene, Poisson-Gaussian a={a:g} b={b:g}, {levels} levels, "
                       f"{2 * radius + 1}x{2 * radius + 1} window", "unit": "Mpix/s"}
    for prec in ("fp16", "fp32"):
        r = time_single(gcam, torch, w, h, levels, radius, prec, 0, ycc, params, nlf, args.steps, args.warmup, dev)
        out[prec] = {"value_median": round(mpix / (r["step_stats"]["median_ms"] / 1e3), 2),
                     "step_stats": r["step_stats"]}
        if prec == "fp16":
            b_alg = algorithmic_bytes_per_base_pixel(levels)
            gbs = b_alg * w * h / (r["step_stats"]["median_ms"] / 1e3) / 1e9
            out["pipeline"] = {"B_alg_per_base_px": round(b_alg, 3), "achieved_GBs": round(gbs, 2),
                               "hbm_frac": round(gbs / HBM_PEAK_GBS, 5)}
    if not args.no_cpu_baseline:
        import oracle
        host_img = ycc.cpu().numpy()
        host_nlf = np.array([[list(A), list(B)] for A, B in nlf], np.float64)
        pr = [dict(luma=p.luma, chroma=p.chroma, chromaBoost=p.chromaBoost, gradientBoost=p.gradientBoost,
                   gradientThreshold=p.gradientThreshold, sharpening=p.sharpening) for p in params]
        host_cpus = cpu_cores()
        threads = args.cpu_threads or min(host_cpus, BOX_CPU_SHARE)
        oracle.set_threads(threads)
        reps, t = 0, time.perf_counter()
        while True:  # at least one frame and about a second of CPU work
            oracle.pyramid_denoise(host_img, pr, host_nlf, levels=levels, radius=radius)
            reps += 1
            if time.perf_counter() - t > 1.0:
                break
        cpu_s = (time.perf_counter() - t) / reps
        out["cpu_baseline"] = {"value": round(mpix / cpu_s, 3), "unit": "Mpix/s", "cores": threads, "kind": "port",
                               "host_cpus": host_cpus, "cpu_model": cpu_model(),
                               "sample": f"the whole {w}x{h} frame x {reps}, FP32 C++ oracle, {threads} threads"}
    return out


def burst(args, steps=None, warmup=None) -> dict:
    """The batched form: 32 frames, sizes drawn uniformly (seed 42) from four camera sizes, noise gain a
    log-uniform in [1e-3, 1e-2]; one step = one gcam_burst_denoise call (one graph launch)."""
    import numpy as np
    import torch

    import glasscamerapublic_amd as gcam

    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    levels, radius, b = 5, 4, 2e-3
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    sizes, gains = burst_plan(args.burst_frames)
    n = len(sizes)
    params = preset_params(gcam, levels)
    frames, nlfs = [], []
    cal = {}
    for f, ((w, h), a) in enumerate(zip(sizes, gains)):
        img = gcam.synth_frame(w, h, float(a), b, scene_seed=1234 + 7 * f, noise_seed=5678 + f, device=dev)
        # per-frame noise model from the frame itself (MeasureYCbCrNLF per level), outside the timed region
        if (w, h) not in cal:
            cal[(w, h)] = gcam.PyramidProcessor(w, h, levels=levels, search_radius=radius, precision=args.precision,
                                                device=dev)
        nl = [((1e-8,) * 3, (1e-8,) * 3) for _ in range(levels)]
        cal[(w, h)].denoise(params, img, None, nl, calibrateFromImage=True)
        frames.append(img)
        nlfs.append(nl)
    for p in cal.values():
        p.close()
    torch.cuda.synchronize()
    bp = gcam.BurstProcessor(sizes, levels=levels, search_radius=radius, precision=args.precision, device=dev)
    t = time.perf_counter()
    bp.denoise(params, frames, nlfs)  # autotune + graph capture + first run
    torch.cuda.synchronize()
    first_call_s = time.perf_counter() - t
    settle(lambda: bp.denoise(params, frames, nlfs), warmup)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record(stream)
    for k in range(steps):
        bp.denoise(params, frames, nlfs)
        evs[k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = [evs[k].elapsed_time(evs[k + 1]) for k in range(steps)]
    ms_per_step = evs[0].elapsed_time(evs[-1]) / steps
    mpix = sum(w * h for w, h in sizes) / 1e6
    value = mpix / (ms_per_step / 1e3)
    tun = bp.tuning
    b_alg = algorithmic_bytes_per_base_pixel(levels)
    pipeline_gbs = b_alg * mpix * 1e6 / (ms_per_step / 1e3) / 1e9
    level_px = [sum((w >> i) * (h >> i) for w, h in sizes) for i in range(levels)]
    bm_gbs = [24.0 * level_px[i] / (tun["level_ms"][i] / 1e3) / 1e9 if tun["level_ms"][i] > 0 else None
              for i in range(levels)]
    stats = step_stats(step_ms)
    out = {
        "metric": "Mpix/s, burst of 32 mixed-size frames, 5-level PCA denoise, one graph launch",
        "value": round(value, 3),
        "value_median": round(mpix / (stats["median_ms"] / 1e3), 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(ms_per_step, 3),
        "step_stats": stats,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic",
        "config": {"workload": f"burst of {n} frames, sizes uniform (seed 42) from "
                               f"{', '.join(f'{w}x{h}' for w, h in BURST_SIZES)}, a log-uniform [1e-3, 1e-2], "
                               f"b={b:g}, {levels} levels, {2 * radius + 1}x{2 * radius + 1} window",
                   "frames": n, "total_mpix": round(mpix, 3), "precision": args.precision,
                   "denoise_parameters": PRESET_NAME, "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "kernel": f"k_burst_block_match_lp<{radius}> (level 0)",
                     "achieved": round(bm_gbs[0], 2) if bm_gbs[0] else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bm_gbs[0] / HBM_PEAK_GBS, 5) if bm_gbs[0] else None, "traffic": None},
        "pipeline": {"B_alg_per_base_px": round(b_alg, 3), "achieved_GBs": round(pipeline_gbs, 2),
                     "hbm_frac": round(pipeline_gbs / HBM_PEAK_GBS, 5)},
        "tuning": tun,
        "first_call_s": round(first_call_s, 3),
        "wall_ms_per_step": round(wall * 1e3 / steps, 3),
        "cpu_baseline": None,
    }
    bp.close()
    del frames
    torch.cuda.empty_cache()
    return out


def burst_plan(n):
    """Frame sizes (uniform over the four camera sizes) and noise gains (log-uniform in [1e-3, 1e-2]) of the
    n-frame burst, seed 42."""
    import numpy as np
    rng = np.random.default_rng(42)
    sizes = [BURST_SIZES[int(k)] for k in rng.integers(0, len(BURST_SIZES), n)]
    gains = np.exp(rng.uniform(np.log(1e-3), np.log(1e-2), n))
    return sizes, gains


def copy_calibration(args) -> dict:
    """A plain device-to-device copy of 256 MiB (torch's copy kernel on the current stream), the reference
    point that scales the FETCH_SIZE / WRITE_SIZE counters of the denoise kernels to bytes: under
    rocprofv3 --pmc each dispatch reads and writes exactly 256 MiB.  Also reports the copy's bandwidth."""
    import torch
    n = 256 * (1 << 20) // 4
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream()
    settle(lambda: dst.copy_(src), args.warmup, min_s=0.1)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        dst.copy_(src)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    gbs = 2 * n * 4 / (ms / 1e3) / 1e9
    return {"metric": "GB/s, device copy of 256 MiB (counter calibration)", "value": round(gbs, 1), "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "copy 256 MiB", "bytes_read": n * 4, "bytes_written": n * 4},
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": 2 * n * 4}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="12mp", choices=sorted(WORKLOADS) + ["burst32", "copy"])
    ap.add_argument("--precision", default="fp16", choices=["fp32", "fp16", "bf16"])
    ap.add_argument("--bm-rows", type=int, default=0, help="block-matching output rows per thread (0 = default)")
    ap.add_argument("--burst-frames", type=int, default=32)
    ap.add_argument("--no-compare", action="store_true", help="skip timing the FP32 form beside the reported one")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-front-end", action="store_true", help="skip timing the raw front end beside the denoise")
    ap.add_argument("--no-registration", action="store_true",
                    help="skip timing burst registration (registerAndFuse, FindHomography) beside the denoise")
    ap.add_argument("--no-quality", action="store_true", help="skip the PSNR check against the FP64 oracle")
    ap.add_argument("--no-burst", action="store_true", help="skip the 32-frame burst leg of the default run")
    ap.add_argument("--no-wide", action="store_true",
                    help="skip timing both forms with the reference's 21x21 window on the same frame")
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the 1024x1024, 4-level, 7x7 leg (both forms and the CPU oracle) of the default run")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help=f"oracle threads (default: the host's CPUs, at most {BOX_CPU_SHARE})")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        return 2
    if args.workload == "copy":
        out = copy_calibration(args)
    elif args.workload == "burst32":
        if args.precision == "fp32":
            print("the burst form runs the precision form: --precision fp16 or bf16", file=sys.stderr)
            return 2
        out = burst(args)
    else:
        out = single_frame(args)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
