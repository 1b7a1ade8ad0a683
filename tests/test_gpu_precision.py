"""Precision form (FP16 / BF16 contractions, FP32 accumulation) against the
FP64 oracle (GPU).

The bar of the precision form is image quality, not bit parity: the PSNR of
its output against the clean synthetic scene may be at most 0.05 dB below the
PSNR of the FP64 oracle's output (FP16).  BF16's 8-bit mantissa is reported,
with a looser bar.  Errors against the FP64 oracle are printed per level.
"""
import numpy as np
import pytest
import torch

import glasscamerapublic_amd as gcam
import oracle
from tests.helpers import nlf_for

pytestmark = pytest.mark.gpu


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def psnr(a, b):
    mse = np.mean((a[..., :3] - b[..., :3]) ** 2)
    return 10.0 * np.log10(1.0 / mse)


def camera_params(levels):
    lm = [6.0, 3.0, 2.0, 2.0, 2.0]
    return [dict(luma=lm[i], chroma=1.0, chromaBoost=[8, 4, 2, 2, 2][i], gradientBoost=0.0, gradientThreshold=1.0,
                 sharpening=1.0) for i in range(levels)]


@pytest.mark.parametrize("radius", [1, 4, 5, 7, 10])
@pytest.mark.parametrize("precision,tol", [("fp16", 5e-3), ("bf16", 3e-2)])
def test_block_match_lp_close_to_fp64(precision, tol, radius):
    w, h = 96, 64
    ycc = gcam.synth_frame(w, h, 4e-3, 2e-3, scene_seed=3, noise_seed=4)
    img = host(ycc)
    basis, _ = oracle.build_pca_space(oracle.collect_patches(img))
    feat = oracle.pca_project(img, basis)
    args = dict(var_a=[4e-6, 1e-6, 1e-6], var_b=[2.4e-3, 1e-3, 1e-3], thresholdMultipliers=[4.0, 2.0, 2.0],
                chromaBoost=2.0)
    ref = oracle.block_match(img, feat, args["var_a"], args["var_b"], args["thresholdMultipliers"],
                             args["chromaBoost"], 0.0, 1.0, 0.0, radius)
    got = host(gcam.block_matching_denoise_image_lp(ycc, torch.from_numpy(basis.astype(np.float32)),
                                                    search_radius=radius, precision=precision, **args))
    err = np.abs(got[..., :3] - ref[..., :3])
    print(f"{precision} r{radius}: max-abs {err.max():.3g} mean-abs {err.mean():.3g}")
    assert err.max() < tol
    assert err.mean() < tol / 10


@pytest.mark.parametrize("rows", [1, 2, 4])
@pytest.mark.parametrize("radius", [1, 5, 7, 10])
@pytest.mark.parametrize("precision,tol", [("fp16", 5e-3), ("bf16", 3e-2)])
def test_lp_tile_shapes_vs_fp64(precision, tol, radius, rows, oracle_threads):
    """Every tile shape of the precision-form kernel at this size (one, two or four output rows per
    thread for radius <= 6; one or two rows for radius 7..10, where four runs as one) through the
    whole two-level pyramid, against the FP64 oracle on both levels."""
    w, h, levels = 288, 200, 2
    ycc = gcam.synth_frame(w, h, 4e-3, 2e-3, scene_seed=11 + radius, noise_seed=12 + rows)
    prm = camera_params(levels)
    nlf = [(tuple(x[0]), tuple(x[1])) for x in nlf_for(levels)]
    proc = gcam.PyramidProcessor(w, h, levels=levels, search_radius=radius, precision=precision,
                                 bm_rows_per_thread=rows)
    proc.denoise([gcam.DenoiseParameters(**p) for p in prm], ycc, None, list(nlf))
    got = [host(t) for t in proc.denoisedImagePyramid]
    proc.close()
    ref = oracle.pyramid_denoise(host(ycc), prm, np.array(nlf), levels=levels, radius=radius, dtype=np.float64)
    for i in range(levels):
        err = np.abs(got[i][..., :3] - ref.denoised[i][..., :3])
        print(f"{precision} r{radius} rows {rows} level {i}: max-abs {err.max():.3g} mean-abs {err.mean():.3g}")
        assert err.max() < tol
        assert err.mean() < tol / 10


@pytest.mark.parametrize("radius,rows", [(10, 0), (7, 0), (5, 1)])
def test_one_row_32x16_tiles_vs_fp64(radius, rows, oracle_threads):
    """One output row per thread on 32 x 16 threads, the shape of levels with at least 1024 such
    tiles (the default above radius 6, and an explicit one row per thread at any radius): level 0
    of 1024 x 512 takes it, level 1 the 32 x 8 threads of the small levels; FP16 against the FP64
    oracle on both levels."""
    w, h, levels = 1024, 512, 2
    ycc = gcam.synth_frame(w, h, 4e-3, 2e-3, scene_seed=21 + radius, noise_seed=22)
    prm = camera_params(levels)
    nlf = [(tuple(x[0]), tuple(x[1])) for x in nlf_for(levels)]
    proc = gcam.PyramidProcessor(w, h, levels=levels, search_radius=radius, precision="fp16",
                                 bm_rows_per_thread=rows)
    proc.denoise([gcam.DenoiseParameters(**p) for p in prm], ycc, None, list(nlf))
    got = [host(t) for t in proc.denoisedImagePyramid]
    proc.close()
    ref = oracle.pyramid_denoise(host(ycc), prm, np.array(nlf), levels=levels, radius=radius, dtype=np.float64)
    for i in range(levels):
        err = np.abs(got[i][..., :3] - ref.denoised[i][..., :3])
        print(f"r{radius} rows {rows} level {i}: max-abs {err.max():.3g} mean-abs {err.mean():.3g}")
        assert err.max() < 5e-3
        assert err.mean() < 5e-4


@pytest.mark.parametrize("radius", [4, 10])
@pytest.mark.parametrize("precision,max_loss_db", [("fp16", 0.05), ("bf16", 0.5)])
def test_pyramid_psnr_loss_vs_fp64(precision, max_loss_db, radius, oracle_threads):
    w, h, levels = 1024, 768, 5
    ycc, clean = gcam.synth_frame(w, h, 4e-3, 2e-3, scene_seed=61, noise_seed=62, want_clean=True)
    prm = camera_params(levels)
    nlf = [(tuple(x[0]), tuple(x[1])) for x in nlf_for(levels)]
    proc = gcam.PyramidProcessor(w, h, levels=levels, search_radius=radius, precision=precision)
    proc.denoise([gcam.DenoiseParameters(**p) for p in prm], ycc, None, list(nlf))
    lp_levels = [host(t) for t in proc.denoisedImagePyramid]
    ref = oracle.pyramid_denoise(host(ycc), prm, np.array(nlf), levels=levels, radius=radius, dtype=np.float64)
    c = host(clean)
    p_ref, p_lp = psnr(ref.denoised[0], c), psnr(lp_levels[0], c)
    for i in range(levels):
        e = np.abs(lp_levels[i][..., :3] - ref.denoised[i][..., :3])
        print(f"{precision} level {i}: max-abs {e.max():.3g} mean-abs {e.mean():.3g} vs FP64")
    print(f"{precision}: PSNR fp64 {p_ref:.4f} dB, {precision} {p_lp:.4f} dB, loss {p_ref - p_lp:.4f} dB, "
          f"noisy {psnr(host(ycc), c):.4f} dB")
    assert p_ref - p_lp <= max_loss_db


def test_fp32_form_psnr_matches_fp64(oracle_threads):
    w, h, levels, radius = 1024, 768, 5, 4
    ycc, clean = gcam.synth_frame(w, h, 4e-3, 2e-3, scene_seed=61, noise_seed=62, want_clean=True)
    prm = camera_params(levels)
    nlf = [(tuple(x[0]), tuple(x[1])) for x in nlf_for(levels)]
    proc = gcam.PyramidProcessor(w, h, levels=levels, search_radius=radius)
    out = host(proc.denoise([gcam.DenoiseParameters(**p) for p in prm], ycc, None, list(nlf)))
    ref = oracle.pyramid_denoise(host(ycc), prm, np.array(nlf), levels=levels, radius=radius, dtype=np.float64)
    c = host(clean)
    assert abs(psnr(out, c) - psnr(ref.denoised[0], c)) < 1e-3
