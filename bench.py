#!/usr/bin/env python3
"""Throughput of the multi-level PCA block-matching denoise on one MI355X.

One step = one whole ``PyramidProcessor::denoise`` call (5-level pyramid,
5x5 patches -> 8 PCA components, 9x9 search window) over one synthetic 12 MP
frame already resident in HBM.  Prints ONE JSON line (rank 0) with the
throughput in Mpix/s of the base level, the roofline of the dominant kernel
(block matching, timed live with HIP events on its own stream) and the CPU
oracle timed on this host's cores on the same frame.

    python bench.py [--steps 20] [--warmup 3] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="12mp", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp16", "bf16"])
    ap.add_argument("--bm-rows", type=int, default=0, help="block-matching output rows per thread (0 = default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (default: min(cores, 16))")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import glasscamerapublic_amd as gcam

    w, h, levels, radius, a, b = WORKLOADS[args.workload]
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()

    ycc = gcam.synth_frame(w, h, a, b, scene_seed=1234, noise_seed=5678, device=dev)
    proc = gcam.PyramidProcessor(w, h, levels=levels, search_radius=radius, precision=args.precision, device=dev,
                                 bm_rows_per_thread=args.bm_rows)
    params = [gcam.DenoiseParameters() for _ in range(levels)]
    nlf = [((1e-8,) * 3, (1e-8,) * 3) for _ in range(levels)]

    # Noise model: MeasureYCbCrNLF on every level, once per frame, outside the timed region.
    proc.denoise(params, ycc, None, nlf, calibrateFromImage=True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    proc.denoise(params, ycc, None, list(nlf), calibrateFromImage=True)
    e1.record(stream)
    torch.cuda.synchronize()
    calib_ms = e0.elapsed_time(e1)

    for _ in range(args.warmup):
        proc.denoise(params, ycc, None, nlf)
    torch.cuda.synchronize()
    e0.record(stream)
    proc.denoise(params, ycc, None, nlf)
    e1.record(stream)
    torch.cuda.synchronize()
    nlf_fit_ms = max(0.0, calib_ms - e0.elapsed_time(e1))

    # Timed region: exactly K steps; per-step events give the median, stage events
    # (recorded on the same stream by the library) time the block-matching kernel.
    proc.set_profiling(True)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record(stream)
    for k in range(args.steps):
        proc.denoise(params, ycc, None, nlf)
        evs[k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = [evs[k].elapsed_time(evs[k + 1]) for k in range(args.steps)]
    total_ms = evs[0].elapsed_time(evs[-1])
    stages = proc.stage_times()
    proc.set_profiling(False)

    ms_per_step = total_ms / args.steps
    mpix = w * h / 1e6
    value = mpix / (ms_per_step / 1e3)

    # Dominant kernel: block matching, one launch per level.  Algorithmic bytes
    # per launch = 24 B (3 ch x FP32 in + 3 ch x FP32 out) per level pixel.
    level_px = [(w >> i) * (h >> i) for i in range(levels)]
    bm_ms = stages["block_match"]
    bm_avg_ms = sum(bm_ms) / levels
    bm_bytes_per_launch = 24.0 * sum(level_px) / levels
    bm_gbs = bm_bytes_per_launch / (bm_avg_ms / 1e3) / 1e9
    b_alg = algorithmic_bytes_per_base_pixel(levels)
    pipeline_gbs = b_alg * w * h / (ms_per_step / 1e3) / 1e9
    # VALU view of the same kernel: ~47 issue slots per (pixel, neighbour) incl. 2 quarter-rate exp
    pairs = sum(level_px) * (2 * radius + 1) ** 2
    valu_tops = pairs * 47 / (sum(bm_ms) / 1e3) / 1e12
    valu_peak_tops = 256 * 4 * 16 * 2.4e9 / 1e12

    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        threads = args.cpu_threads or min(cpu_cores(), 16)
        oracle.set_threads(threads)
        host_img = ycc.cpu().numpy()
        host_nlf = np.array([[list(A), list(B)] for A, B in nlf], np.float64)
        pr = [dict(luma=p.luma, chroma=p.chroma, chromaBoost=p.chromaBoost, gradientBoost=p.gradientBoost,
                   gradientThreshold=p.gradientThreshold, sharpening=p.sharpening) for p in params]
        t = time.perf_counter()
        oracle.pyramid_denoise(host_img, pr, host_nlf, levels=levels, radius=radius)
        cpu_s = time.perf_counter() - t
        cpu = {"value": round(mpix / cpu_s, 4), "unit": "Mpix/s", "cores": threads, "kind": "port",
               "sample": f"one full {w}x{h} frame (the benchmark frame), {levels} levels, "
                         f"{2 * radius + 1}x{2 * radius + 1} window, FP32 C++ oracle, {threads} threads, "
                         f"{cpu_s:.1f} s"}

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
        "dtype": {"fp32": "f32", "fp16": "f16", "bf16": "bf16"}[args.precision],
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
            "traffic": None,
            "avg_launch_ms": round(bm_avg_ms, 4),
            "bytes_per_launch": int(bm_bytes_per_launch),
        },
        "valu": {"kernel": f"k_block_match<{radius}>", "achieved_tops": round(valu_tops, 2),
                 "peak_tops": round(valu_peak_tops, 2), "frac": round(valu_tops / valu_peak_tops, 4)},
        "pipeline": {"B_alg_per_base_px": round(b_alg, 3), "achieved_GBs": round(pipeline_gbs, 2),
                     "hbm_frac": round(pipeline_gbs / HBM_PEAK_GBS, 5)},
        "median_ms": round(statistics.median(step_ms), 4),
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 4),
        "nlf_fit_ms": round(nlf_fit_ms, 3),
        "stage_ms_last_step": {k: [round(x, 4) for x in v] for k, v in stages.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
