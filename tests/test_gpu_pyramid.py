"""Whole-pyramid parity of PyramidProcessor::denoise on gfx950 against the CPU
oracle (GPU), on the synthetic workloads the benchmark uses.

Bar: every denoised level within max-abs 1e-4 (YCbCr channels, full scale
~1) of the FP32 oracle; .w (the luma weight sum, up to ~(2R+1)^2) within 1e-4
relative.  Every configuration keeps at least 9 sampled patches on its
coarsest level so that the 8-dimensional PCA subspace is well defined.
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
    return t.detach().cpu().numpy()


def check_levels(proc, ref, tag):
    for i, lvl in enumerate(proc.denoisedImagePyramid):
        got = host(lvl)
        exp = ref.denoised[i]
        err = np.abs(got[..., :3] - exp[..., :3]).max()
        assert err <= 1e-4, f"{tag} level {i}: max-abs {err:.3g}"
        np.testing.assert_allclose(got[..., 3], exp[..., 3], rtol=1e-4, err_msg=f"{tag} level {i} .w")


def run_case(w, h, levels, radius, params, a, b, seed, calibrate=False, gradient=False, lens=0.0, exposure=1.0):
    ycc = gcam.synth_frame(w, h, a, b, scene_seed=seed, noise_seed=seed + 1)
    grad = None
    if gradient:
        g = torch.Generator(device="cpu").manual_seed(seed)
        grad = (torch.randn((h, w, 2), generator=g) * 0.02).to("cuda")
    proc = gcam.PyramidProcessor(w, h, levels=levels, search_radius=radius)
    nlf = [(tuple(x[0]), tuple(x[1])) for x in nlf_for(levels, a, b)]
    nlf_in = list(nlf)
    proc.denoise([gcam.DenoiseParameters(**p) for p in params], ycc, grad, nlf, exposure, lens, calibrate)
    ref = oracle.pyramid_denoise(host(ycc), params, np.array(nlf_in) if not calibrate else np.zeros((levels, 2, 3)),
                                 levels=levels, radius=radius, gradient=None if grad is None else host(grad),
                                 exposure_multiplier=exposure, lens_shading=lens, calibrate=calibrate)
    return proc, ref, nlf


def params_default(levels):
    return [dict(luma=1.0, chroma=1.0, chromaBoost=1.0, gradientBoost=0.0, gradientThreshold=1.0, sharpening=1.0)
            for _ in range(levels)]


def params_camera(levels):
    # iPhone-14-like presets: strong luma thresholds, gradient boost and sharpening on the finest levels
    lm = [6.0, 3.0, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0]
    out = []
    for i in range(levels):
        out.append(dict(luma=lm[i], chroma=1.0, chromaBoost=[8, 4, 2, 2, 2, 2, 2, 2][i],
                        gradientBoost=[4.0, 2.0, 0, 0, 0, 0, 0, 0][i], gradientThreshold=2.0 if i < 2 else 1.0,
                        sharpening=1.5 if i == 0 else 1.0))
    return out


def test_pyramid_1024_l4_r3_default_params(oracle_threads):
    """1024x1024 RGGB, 4 levels, 7x7 window, a=2e-3 b=1e-3."""
    proc, ref, _ = run_case(1024, 1024, 4, 3, params_default(4), 2e-3, 1e-3, seed=11)
    check_levels(proc, ref, "1024/L4/r3")


def test_pyramid_camera_params_gradient_lens(oracle_threads):
    proc, ref, _ = run_case(640, 480, 5, 4, params_camera(5), 5e-3, 2e-3, seed=21, gradient=True, lens=1.6)
    check_levels(proc, ref, "640x480/camera")


def test_pyramid_512_l5_r4(oracle_threads):
    """512x512 crop size of the differentiation workload: 16 patches on level 4."""
    proc, ref, _ = run_case(512, 512, 5, 4, params_camera(5), 5e-3, 2e-3, seed=31)
    check_levels(proc, ref, "512/L5/r4")


def test_pyramid_reference_radius_10(oracle_threads):
    """The reference's own 21x21 window."""
    proc, ref, _ = run_case(384, 256, 3, 10, params_camera(3), 4e-3, 2e-3, seed=41)
    check_levels(proc, ref, "384x256/r10")


def test_pyramid_calibrate_from_image(oracle_threads):
    proc, ref, nlf = run_case(768, 512, 4, 3, params_camera(4), 4e-3, 2e-3, seed=51, calibrate=True, exposure=1.3)
    for i in range(4):
        np.testing.assert_allclose(nlf[i][0], ref.nlf[i, 0], rtol=1e-4, err_msg=f"A level {i}")
        np.testing.assert_allclose(nlf[i][1], ref.nlf[i, 1], rtol=1e-4, err_msg=f"B level {i}")
    check_levels(proc, ref, "calibrated")


def test_pyramid_12mp_l5_r4(oracle_threads):
    """4032x3024 (12 MP), 5 levels, 9x9 window, a=4e-3 b=2e-3: the benchmark workload."""
    proc, ref, _ = run_case(4032, 3024, 5, 4, params_default(5), 4e-3, 2e-3, seed=1234)
    check_levels(proc, ref, "12MP")


def test_repeat_calls_are_bit_identical_and_strided_input_matches():
    w, h = 512, 384
    ycc = gcam.synth_frame(w, h, 4e-3, 2e-3, scene_seed=5, noise_seed=6)
    proc = gcam.PyramidProcessor(w, h, levels=4, search_radius=4)
    nlf = [(tuple(x[0]), tuple(x[1])) for x in nlf_for(4)]
    prm = [gcam.DenoiseParameters(**p) for p in params_camera(4)]
    a = proc.denoise(prm, ycc, None, nlf).clone()
    b = proc.denoise(prm, ycc, None, nlf).clone()
    assert torch.equal(a, b)
    padded = torch.zeros((h, w + 40, 4), device="cuda")
    padded[:, :w] = ycc
    c = proc.denoise(prm, padded[:, :w], None, nlf).clone()  # row pitch w + 40 pixels
    assert torch.equal(a, c)
    # the returned view is level 0 of denoisedImagePyramid
    assert torch.equal(proc.denoisedImagePyramid[0], c)


def test_denoise_reduces_error_against_clean_scene():
    w, h = 1024, 768
    ycc, clean = gcam.synth_frame(w, h, 8e-3, 2e-3, scene_seed=7, noise_seed=8, want_clean=True)
    proc = gcam.PyramidProcessor(w, h, levels=5, search_radius=4)
    nlf = [(tuple(x[0]), tuple(x[1])) for x in nlf_for(5)]
    prm = [gcam.DenoiseParameters(**p) for p in params_camera(5)]
    proc.denoise(prm, ycc, None, nlf, calibrateFromImage=True)
    out = proc.denoise(prm, ycc, None, nlf)
    e_in = torch.sqrt(((ycc[..., 0] - clean[..., 0]) ** 2).mean()).item()
    e_out = torch.sqrt(((out[..., 0] - clean[..., 0]) ** 2).mean()).item()
    assert e_out < 0.8 * e_in, (e_in, e_out)
