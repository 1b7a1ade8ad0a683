#!/usr/bin/env python3
"""Throughput of the multi-level PCA block-matching denoise on one MI355X.

One step = one whole ``PyramidProcessor::denoise`` call (5-level pyramid,
5x5 patches -> 8 PCA components, 9x9 search window) over one synthetic 12 MP
frame already resident in HBM.  The reported form is the precision form
(FP16 MFMA projection and FP16 distance products, FP32 accumulation); the
plain FP32 form is timed beside it on the same frame, and the PSNR loss of the
reported form against the FP64 oracle is measured on that frame.  Prints ONE
JSON line with the throughput in Mpix/s of the base level (K timed steps with
no instrumentation), the roofline of the dominant kernel (block matching,
timed over 5 further steps with HIP events the library records around each
stage on the stream it launches on) and the CPU oracle timed on this host's
cores on the same frame.

    python bench.py [--steps 20] [--warmup 3] [--precision fp16] [--no-cpu-baseline]
    python bench.py --workload burst32      # the 32-frame mixed-size burst (batched form)

The raw front end that produces the denoise's inputs from the frame's Bayer
mosaic (RawFrontEnd) is timed beside it under "front_end" (--no-front-end to
skip), and burst fusion (registerAndFuse of the frame, FindHomography on 1500
matches) under "registration" (--no-registration to skip).
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
    ("12mp", "fp16"): [58.1, 52.1, 40.3, 40.7, 43.9],
    ("12mp", "fp32"): [111.6, 103.7, 81.8, 79.7, 78.7],
}
# Wave-level VALU instructions per block-matching dispatch on the same workload, levels 0..4:
# rocprofv3 --pmc SQ_INSTS_VALU, in millions.
MEASURED_BM_VALU = {
    ("12mp", "fp16"): [395.661, 99.062, 25.232, 6.754, 1.704],
    ("12mp", "fp32"): [534.553, 133.651, 33.950, 8.639, 2.171],
}
DTYPE = {"fp32": "f32", "fp16": "f16", "bf16": "bf16"}
# Vector ALU: 256 CUs x 4 SIMDs x 16 lanes at 2.4 GHz (a wave64 instruction takes 4 cycles).
VALU_PEAK_LANE_OPS = 256 * 4 * 16 * 2.4e9
# Issue slots per (pixel, window position) of the reference's inner loop: ~39 VALU operations and
# two exponentials at quarter rate.
MODEL_SLOTS_PER_PAIR = 47


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


def time_single(gcam, torch, w, h, levels, radius, precision, bm_rows, ycc, params, nlf, steps, warmup, dev):
    """K timed steps of one processor, then a few steps with the library's stage
    events on (HIP events it records around every stage on its stream) for the
    per-stage breakdown; the timed steps carry no stage events."""
    stream = torch.cuda.current_stream()
    proc = gcam.PyramidProcessor(w, h, levels=levels, search_radius=radius, precision=precision, device=dev,
                                 bm_rows_per_thread=bm_rows)
    for _ in range(warmup):
        proc.denoise(params, ycc, None, nlf)
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
    bm = []
    for _ in range(5):
        proc.denoise(params, ycc, None, nlf)
        st = proc.stage_times()
        bm.append(st["block_match"])
    stages = proc.stage_times()
    stages["block_match_mean5"] = [sum(x[i] for x in bm) / len(bm) for i in range(levels)]
    proc.set_profiling(False)
    out = proc.denoisedImagePyramid[0].clone()
    proc.close()
    total_ms = evs[0].elapsed_time(evs[-1])
    return {"ms_per_step": total_ms / steps, "median_ms": statistics.median(step_ms),
            "wall_ms_per_step": wall * 1e3 / steps, "stages": stages, "out": out}


# HBM bytes per pixel of the front end's stage DAG, each stage's inputs read
# once and outputs written once (FP32 images, 16-bit raw): scale 2 + 4,
# gradient (Sobel fused) 4 + 8, green 4 + 8 + 4, red/blue 4 + 4 + 8 + 16,
# red/blue-at-green + highlights + YCbCr 16 + 8 + 16, despeckle 16 + 16.
FRONT_END_BYTES_PER_PX = 6 + 12 + 16 + 32 + 40 + 32


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
    for _ in range(warmup):
        fe.run(raw, params, ycc, grad)
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
           "roofline": {"bound": "hbm", "bytes_per_px": FRONT_END_BYTES_PER_PX, "achieved": round(gbs, 1),
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
    for _ in range(warmup):
        gcam.registerAndFuse(fused, ycc, fused, H, 2)
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


def single_frame(args) -> dict:
    import numpy as np
    import torch

    import glasscamerapublic_amd as gcam

    w, h, levels, radius, a, b = WORKLOADS[args.workload]
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()

    ycc, raw, clean = gcam.synth_frame(w, h, a, b, scene_seed=1234, noise_seed=5678, device=dev, want_raw=True,
                                       want_clean=True)
    params = [gcam.DenoiseParameters() for _ in range(levels)]
    nlf = [((1e-8,) * 3, (1e-8,) * 3) for _ in range(levels)]

    # Noise model: MeasureYCbCrNLF on every level, once per frame, outside the timed region.
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
    nlf_fit_ms = max(0.0, calib_ms - e0.elapsed_time(e1))
    cal.close()

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
    bm_ms = prim["stages"].pop("block_match_mean5")
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
        threads = args.cpu_threads or min(cpu_cores(), 16)
        oracle.set_threads(threads)
        t = time.perf_counter()
        oracle.pyramid_denoise(host_img, pr, host_nlf, levels=levels, radius=radius)
        cpu_s = time.perf_counter() - t
        cpu = {"value": round(mpix / cpu_s, 4), "unit": "Mpix/s", "cores": threads, "kind": "port",
               "sample": f"one full {w}x{h} frame (the benchmark frame), {levels} levels, "
                         f"{2 * radius + 1}x{2 * radius + 1} window, FP32 C++ oracle, {threads} threads, "
                         f"{cpu_s:.1f} s"}
        if not args.no_quality:
            ref64 = oracle.pyramid_denoise(host_img, pr, host_nlf, levels=levels, radius=radius,
                                           dtype=np.float64).denoised[0]
            got = prim["out"].cpu().numpy()
            p64, pgpu = psnr(ref64, host_clean), psnr(got, host_clean)
            quality = {"psnr_fp64_oracle_db": round(p64, 4), f"psnr_{args.precision}_db": round(pgpu, 4),
                       "psnr_loss_db": round(p64 - pgpu, 4), "psnr_noisy_db": round(psnr(host_img, host_clean), 4),
                       "max_abs_vs_fp64": float(np.abs(got[..., :3] - ref64[..., :3]).max())}

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
            "parallelism": "single-gpu",
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
        "median_ms": round(prim["median_ms"], 4),
        "wall_ms_per_step": round(prim["wall_ms_per_step"], 4),
        "nlf_fit_ms": round(nlf_fit_ms, 3),
        "stage_ms_last_step": {k: [round(x, 4) for x in v] for k, v in prim["stages"].items()},
        "cpu_baseline": cpu,
    }
    if ref32 is not None:
        v32 = mpix / (ref32["ms_per_step"] / 1e3)
        ref32["stages"].pop("block_match_mean5")
        out["fp32_form"] = {"value": round(v32, 3), "ms_per_step": round(ref32["ms_per_step"], 4),
                            "speedup_of_reported_form": round(value / v32, 4),
                            "stage_ms_last_step": {k: [round(x, 4) for x in v] for k, v in ref32["stages"].items()}}
    if quality is not None:
        out["quality"] = quality
    if not args.no_front_end:
        # The inputs of this denoise from the frame's raw mosaic (front end of RawConverter::demosaic and
        # its despeckle), reported beside the denoise; the despeckle uses the calibrated level-0 model.
        fe_params = front_end_params(gcam, a, b, nlf[0])
        out["front_end"] = time_front_end(gcam, torch, raw, w, h, fe_params, args.steps, args.warmup,
                                          not args.no_cpu_baseline, 1)
    if not args.no_registration:
        out["registration"] = time_registration(gcam, torch, ycc, w, h, args.steps, args.warmup,
                                                not args.no_cpu_baseline)
    return out


def burst(args) -> dict:
    """The batched form: 32 frames, sizes drawn uniformly (seed 42) from four camera sizes, noise gain a
    log-uniform in [1e-3, 1e-2]; one step = one gcam_burst_denoise call (one graph launch)."""
    import numpy as np
    import torch

    import glasscamerapublic_amd as gcam

    levels, radius, b = 5, 4, 2e-3
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    rng = np.random.default_rng(42)
    n = args.burst_frames
    sizes = [BURST_SIZES[int(k)] for k in rng.integers(0, len(BURST_SIZES), n)]
    gains = np.exp(rng.uniform(np.log(1e-3), np.log(1e-2), n))
    params = [gcam.DenoiseParameters() for _ in range(levels)]
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
    for _ in range(args.warmup):
        bp.denoise(params, frames, nlfs)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record(stream)
    for k in range(args.steps):
        bp.denoise(params, frames, nlfs)
        evs[k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = [evs[k].elapsed_time(evs[k + 1]) for k in range(args.steps)]
    ms_per_step = evs[0].elapsed_time(evs[-1]) / args.steps
    mpix = sum(w * h for w, h in sizes) / 1e6
    value = mpix / (ms_per_step / 1e3)
    tun = bp.tuning
    b_alg = algorithmic_bytes_per_base_pixel(levels)
    pipeline_gbs = b_alg * mpix * 1e6 / (ms_per_step / 1e3) / 1e9
    level_px = [sum((w >> i) * (h >> i) for w, h in sizes) for i in range(levels)]
    bm_gbs = [24.0 * level_px[i] / (tun["level_ms"][i] / 1e3) / 1e9 if tun["level_ms"][i] > 0 else None
              for i in range(levels)]
    out = {
        "metric": "Mpix/s, burst of 32 mixed-size frames, 5-level PCA denoise, one graph launch",
        "value": round(value, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic",
        "config": {"workload": f"burst of {n} frames, sizes uniform (seed 42) from "
                               f"{', '.join(f'{w}x{h}' for w, h in BURST_SIZES)}, a log-uniform [1e-3, 1e-2], "
                               f"b={b:g}, {levels} levels, {2 * radius + 1}x{2 * radius + 1} window",
                   "frames": n, "total_mpix": round(mpix, 3), "precision": args.precision,
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "kernel": f"k_burst_block_match_lp<{radius}> (level 0)",
                     "achieved": round(bm_gbs[0], 2) if bm_gbs[0] else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bm_gbs[0] / HBM_PEAK_GBS, 5) if bm_gbs[0] else None, "traffic": None},
        "pipeline": {"B_alg_per_base_px": round(b_alg, 3), "achieved_GBs": round(pipeline_gbs, 2),
                     "hbm_frac": round(pipeline_gbs / HBM_PEAK_GBS, 5)},
        "tuning": tun,
        "first_call_s": round(first_call_s, 3),
        "median_ms": round(statistics.median(step_ms), 3),
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 3),
        "cpu_baseline": None,
    }
    bp.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="12mp", choices=sorted(WORKLOADS) + ["burst32"])
    ap.add_argument("--precision", default="fp16", choices=["fp32", "fp16", "bf16"])
    ap.add_argument("--bm-rows", type=int, default=0, help="block-matching output rows per thread (0 = default)")
    ap.add_argument("--burst-frames", type=int, default=32)
    ap.add_argument("--no-compare", action="store_true", help="skip timing the FP32 form beside the reported one")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-front-end", action="store_true", help="skip timing the raw front end beside the denoise")
    ap.add_argument("--no-registration", action="store_true",
                    help="skip timing burst registration (registerAndFuse, FindHomography) beside the denoise")
    ap.add_argument("--no-quality", action="store_true", help="skip the PSNR check against the FP64 oracle")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (default: min(cores, 16))")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        return 2
    if args.workload == "burst32":
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
