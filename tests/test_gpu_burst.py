"""Batched (burst) form against the single-frame pyramid (GPU).

A burst runs PyramidProcessor::denoise on every frame through batched
launches, a persistent block-matching kernel and one captured graph.  With the
block-matching tile fixed to 32 x 16 outputs (two rows per thread), every output level of
every frame must equal the single-frame result bit for bit (same kernels, same
reduction order); with the autotuned tile the luma centring of the fused
projection moves with the tile, so outputs agree to FP32 rounding.  The graph is
replayed on new frames, with other GPU work between the replays, and every
replay is checked.
"""
import numpy as np
import pytest
import torch

import glasscamerapublic_amd as gcam
from tests.helpers import nlf_for

pytestmark = pytest.mark.gpu

SIZES = [(640, 480), (512, 384), (768, 512), (384, 256)]
LEVELS, RADIUS = 4, 4


def frames_for(seed, gains):
    return [gcam.synth_frame(w, h, a, 2e-3, scene_seed=seed + 10 * k, noise_seed=seed + 10 * k + 1)
            for k, ((w, h), a) in enumerate(zip(SIZES, gains))]


def frame_params(k):
    lm = [4.0, 2.0, 1.5, 1.0]
    return [gcam.DenoiseParameters(luma=lm[i] * (1 + 0.1 * k), chroma=1.0 + 0.2 * i, chromaBoost=[4, 2, 2, 1][i],
                                   sharpening=1.0 if k % 2 == 0 else 1.3) for i in range(LEVELS)]


def frame_nlf(a):
    return [(tuple(x[0]), tuple(x[1])) for x in nlf_for(LEVELS, a, 2e-3)]


def single_frame(frames, params, nlfs, precision="fp16"):
    outs = []
    for img, prm, nl, (w, h) in zip(frames, params, nlfs, SIZES):
        proc = gcam.PyramidProcessor(w, h, levels=LEVELS, search_radius=RADIUS, precision=precision,
                                     bm_rows_per_thread=2)
        proc.denoise(prm, img, None, list(nl))
        outs.append([t.clone() for t in proc.denoisedImagePyramid])
        proc.close()
    return outs


def test_burst_matches_single_frame_bit_exact():
    gains = [2e-3, 4e-3, 8e-3, 1e-3]
    frames = frames_for(100, gains)
    params = [frame_params(k) for k in range(len(SIZES))]
    nlfs = [frame_nlf(a) for a in gains]
    burst = gcam.BurstProcessor(SIZES, levels=LEVELS, search_radius=RADIUS, precision="fp16")
    burst.set_tuning([2] * LEVELS, [0] * LEVELS, [0] * LEVELS)
    burst.denoise(params, frames, nlfs)
    torch.cuda.synchronize()
    ref = single_frame(frames, params, nlfs)
    for k in range(len(SIZES)):
        for i in range(LEVELS):
            got = burst.output(k, i)
            assert torch.equal(got, ref[k][i]), f"frame {k} level {i}: max diff {(got - ref[k][i]).abs().max()}"


@pytest.mark.parametrize("grid,schedule", [(256, 0), (512, 1), (0, 1)])
def test_persistent_grid_and_schedule_are_bit_exact(grid, schedule):
    gains = [3e-3] * 4
    frames = frames_for(200, gains)
    params = [frame_params(k) for k in range(len(SIZES))]
    nlfs = [frame_nlf(a) for a in gains]
    burst = gcam.BurstProcessor(SIZES, levels=LEVELS, search_radius=RADIUS, precision="fp16")
    burst.set_tuning([2] * LEVELS, [0] * LEVELS, [0] * LEVELS)
    burst.denoise(params, frames, nlfs)
    torch.cuda.synchronize()
    base = [[burst.output(k, i).clone() for i in range(LEVELS)] for k in range(len(SIZES))]
    burst.set_tuning([2] * LEVELS, [grid] * LEVELS, [schedule] * LEVELS)
    burst.denoise(params, frames, nlfs)
    torch.cuda.synchronize()
    for k in range(len(SIZES)):
        for i in range(LEVELS):
            assert torch.equal(burst.output(k, i), base[k][i]), f"frame {k} level {i}"


def test_autotuned_burst_and_graph_replays():
    gains = [1e-3, 5e-3, 2e-3, 9e-3]
    params = [frame_params(k) for k in range(len(SIZES))]
    nlfs = [frame_nlf(a) for a in gains]
    burst = gcam.BurstProcessor(SIZES, levels=LEVELS, search_radius=RADIUS, precision="fp16")
    for rep, seed in enumerate([300, 400, 500]):
        frames = frames_for(seed, gains)
        burst.denoise(params, frames, nlfs)
        torch.cuda.synchronize()
        tun = burst.tuning
        assert tun["tuned"] and all(r in (1, 2, 4) for r in tun["rows_per_thread"])
        print("replay", rep, "tuning", tun)
        ref = single_frame(frames, params, nlfs)  # eager GPU work between the replays
        # The tuned tile may differ from the single-frame one; the FP16 features
        # are centred on each tile's first pixel, so the two differ by FP16
        # rounding of the features (the precision form's bar against FP64 is 5e-3).
        for k in range(len(SIZES)):
            for i in range(LEVELS):
                got = burst.output(k, i)
                err = (got[..., :3] - ref[k][i][..., :3]).abs().max().item()
                assert err < 1e-3, f"replay {rep} frame {k} level {i}: {err}"


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_64_wide_tiles(precision):
    """64-wide block-matching tiles (a tuner candidate for two rows per thread):
    the persistent XCD-chunked launch equals the one-workgroup-per-tile launch
    bit for bit, and both stay within the FP16/BF16 feature rounding of the
    32-wide tiles (features are centred on each tile's first pixel)."""
    gains = [2e-3, 7e-3, 4e-3, 3e-3]
    frames = frames_for(250, gains)
    params = [frame_params(k) for k in range(len(SIZES))]
    nlfs = [frame_nlf(a) for a in gains]
    burst = gcam.BurstProcessor(SIZES, levels=LEVELS, search_radius=RADIUS, precision=precision)
    burst.set_tuning([2] * LEVELS, [0] * LEVELS, [0] * LEVELS)
    burst.denoise(params, frames, nlfs)
    torch.cuda.synchronize()
    narrow = [[burst.output(k, i).clone() for i in range(LEVELS)] for k in range(len(SIZES))]
    burst.set_tuning([2] * LEVELS, [0] * LEVELS, [0] * LEVELS, [64] * LEVELS)
    assert burst.tuning["tile_width"] == [64] * LEVELS
    burst.denoise(params, frames, nlfs)
    torch.cuda.synchronize()
    wide = [[burst.output(k, i).clone() for i in range(LEVELS)] for k in range(len(SIZES))]
    burst.set_tuning([2] * LEVELS, [512] * LEVELS, [1] * LEVELS, [64] * LEVELS)
    burst.denoise(params, frames, nlfs)
    torch.cuda.synchronize()
    tol = 1e-3 if precision == "fp16" else 3e-3
    for k in range(len(SIZES)):
        for i in range(LEVELS):
            assert torch.equal(burst.output(k, i), wide[k][i]), f"frame {k} level {i}"
            err = (wide[k][i][..., :3] - narrow[k][i][..., :3]).abs().max().item()
            assert err < tol, f"frame {k} level {i}: {err}"


@pytest.mark.parametrize("precision,tol", [("fp16", 1e-3), ("bf16", 3e-3)])
def test_reference_window_burst(precision, tol):
    """Radius 10 (the reference's 21 x 21 window): with two rows per thread the
    burst equals the single frame bit for bit; with one row (32 x 16-output
    tiles in the burst) the persistent grid equals one workgroup per tile bit
    for bit, and both agree with the single frame's default tiling to the
    feature rounding of the form."""
    radius = 10
    gains = [3e-3, 6e-3, 2e-3, 5e-3]
    frames = frames_for(800, gains)
    params = [frame_params(k) for k in range(len(SIZES))]
    nlfs = [frame_nlf(a) for a in gains]
    ref2, ref0 = [], []
    for img, prm, nl, (w, h) in zip(frames, params, nlfs, SIZES):
        for rows, dst in ((2, ref2), (0, ref0)):
            proc = gcam.PyramidProcessor(w, h, levels=LEVELS, search_radius=radius, precision=precision,
                                         bm_rows_per_thread=rows)
            proc.denoise(prm, img, None, list(nl))
            dst.append([t.clone() for t in proc.denoisedImagePyramid])
            proc.close()
    burst = gcam.BurstProcessor(SIZES, levels=LEVELS, search_radius=radius, precision=precision)
    burst.set_tuning([2] * LEVELS, [0] * LEVELS, [0] * LEVELS)
    burst.denoise(params, frames, nlfs)
    torch.cuda.synchronize()
    for k in range(len(SIZES)):
        for i in range(LEVELS):
            assert torch.equal(burst.output(k, i), ref2[k][i]), f"rows 2, frame {k} level {i}"
    burst.set_tuning([1] * LEVELS, [0] * LEVELS, [0] * LEVELS)
    burst.denoise(params, frames, nlfs)
    torch.cuda.synchronize()
    one = [[burst.output(k, i).clone() for i in range(LEVELS)] for k in range(len(SIZES))]
    burst.set_tuning([1] * LEVELS, [512] * LEVELS, [1] * LEVELS)
    burst.denoise(params, frames, nlfs)
    torch.cuda.synchronize()
    for k in range(len(SIZES)):
        for i in range(LEVELS):
            assert torch.isfinite(one[k][i]).all(), f"frame {k} level {i}"
            assert torch.equal(burst.output(k, i), one[k][i]), f"persistent, frame {k} level {i}"
            err = (one[k][i][..., :3] - ref0[k][i][..., :3]).abs().max().item()
            assert err < tol, f"rows 1, frame {k} level {i}: {err}"
    burst.close()


def test_bf16_burst_matches_single_frame():
    gains = [4e-3] * 4
    frames = frames_for(600, gains)
    params = [frame_params(k) for k in range(len(SIZES))]
    nlfs = [frame_nlf(a) for a in gains]
    burst = gcam.BurstProcessor(SIZES, levels=LEVELS, search_radius=RADIUS, precision="bf16")
    burst.set_tuning([2] * LEVELS, [512] * LEVELS, [1] * LEVELS)
    burst.denoise(params, frames, nlfs)
    torch.cuda.synchronize()
    ref = single_frame(frames, params, nlfs, precision="bf16")
    for k in range(len(SIZES)):
        assert torch.equal(burst.output(k, 0), ref[k][0])


def test_burst_argument_errors():
    with pytest.raises(gcam.GcamError):
        gcam.BurstProcessor(SIZES, levels=LEVELS, search_radius=RADIUS, precision="fp32")
    burst = gcam.BurstProcessor(SIZES, levels=LEVELS, search_radius=RADIUS, precision="fp16")
    gains = [4e-3] * 4
    frames = frames_for(700, gains)
    params = [frame_params(k) for k in range(len(SIZES))]
    nlfs = [frame_nlf(a) for a in gains]
    bad = list(frames)
    bad[1] = frames[0]  # wrong size for frame 1
    with pytest.raises(gcam.GcamError, match="frames\\[1\\]"):
        burst.denoise(params, bad, nlfs)
    with pytest.raises(gcam.GcamError):
        burst.set_tuning([3] * LEVELS, [0] * LEVELS, [0] * LEVELS)
    with pytest.raises(gcam.GcamError, match="tile_width"):
        burst.set_tuning([2] * LEVELS, [0] * LEVELS, [0] * LEVELS, [48] * LEVELS)
    with pytest.raises(gcam.GcamError, match="64-wide"):
        burst.set_tuning([1] * LEVELS, [0] * LEVELS, [0] * LEVELS, [64] * LEVELS)
    with pytest.raises(ValueError):
        burst.denoise(params[:2], frames, nlfs)


def test_burst_gradient_sharpening_lens_odd_sizes():
    """Odd frame sizes (no multiple of the tiles), gradients on some frames,
    sharpening > 1 and lens shading: bit-equal to the single-frame path."""
    sizes = [(517, 389), (333, 250), (640, 481)]
    levels, radius = 3, 3
    gains = [3e-3, 6e-3, 2e-3]
    frames = [gcam.synth_frame(w, h, a, 2e-3, scene_seed=900 + k, noise_seed=950 + k)
              for k, ((w, h), a) in enumerate(zip(sizes, gains))]
    grads = [torch.rand((h, w, 2), device="cuda") * 0.05 if k != 1 else None for k, (w, h) in enumerate(sizes)]
    params = [[gcam.DenoiseParameters(luma=2.0, chroma=1.5, chromaBoost=2.0, gradientBoost=2.0 if k != 1 else 0.0,
                                      gradientThreshold=1.5, sharpening=1.4) for _ in range(levels)]
              for k in range(len(sizes))]
    nlfs = [[(tuple(x[0]), tuple(x[1])) for x in nlf_for(levels, a, 2e-3)] for a in gains]
    burst = gcam.BurstProcessor(sizes, levels=levels, search_radius=radius, precision="fp16")
    burst.set_tuning([2] * levels, [0] * levels, [0] * levels)
    burst.denoise(params, frames, nlfs, gradients=grads, lensShadingCorrection=0.3)
    torch.cuda.synchronize()
    for k, (w, h) in enumerate(sizes):
        proc = gcam.PyramidProcessor(w, h, levels=levels, search_radius=radius, precision="fp16",
                                     bm_rows_per_thread=2)
        proc.denoise(params[k], frames[k], grads[k], list(nlfs[k]), lensShadingCorrection=0.3)
        for i in range(levels):
            assert torch.equal(burst.output(k, i), proc.denoisedImagePyramid[i]), f"frame {k} level {i}"
        proc.close()


def test_single_level_burst():
    sizes = [(256, 192), (320, 200)]
    frames = [gcam.synth_frame(w, h, 4e-3, 2e-3, scene_seed=31 + k, noise_seed=41 + k) for k, (w, h) in enumerate(sizes)]
    params = [gcam.DenoiseParameters(luma=3.0)]
    nlf = [(tuple(x[0]), tuple(x[1])) for x in nlf_for(1)]
    burst = gcam.BurstProcessor(sizes, levels=1, search_radius=2, precision="fp16")
    burst.set_tuning([1], [0], [0])
    burst.denoise(params, frames, nlf)
    torch.cuda.synchronize()
    for k, (w, h) in enumerate(sizes):
        proc = gcam.PyramidProcessor(w, h, levels=1, search_radius=2, precision="fp16", bm_rows_per_thread=1)
        ref = proc.denoise(params, frames[k], None, list(nlf))
        assert torch.equal(burst.output(k, 0), ref)
        proc.close()
