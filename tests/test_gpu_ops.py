"""Per-kernel parity of the gfx950 kernels against the CPU oracle (GPU).

Inputs are seeded; each kernel gets the same inputs as the oracle function it
is checked against.  FP32 bar: max-abs 1e-4 of full scale (values here are
O(1)), tighter where the arithmetic is short; integer/index work (patch slots)
is bit-exact.
"""
import numpy as np
import pytest
import torch

import glasscamerapublic_amd as gcam
import oracle

pytestmark = pytest.mark.gpu

DEV = "cuda"


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(DEV)


def to_host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def rand_img(h, w, seed, ch=4):
    a = np.random.default_rng(seed).uniform(0.05, 0.95, (h, w, ch)).astype(np.float32)
    if ch == 4:
        a[..., 1:3] -= 0.5
    return a


def test_synth_frame_is_seeded_and_plausible():
    a = gcam.synth_frame(256, 192, 4e-3, 2e-3, scene_seed=1, noise_seed=2, want_raw=True, want_clean=True)
    b = gcam.synth_frame(256, 192, 4e-3, 2e-3, scene_seed=1, noise_seed=2, want_raw=True, want_clean=True)
    c = gcam.synth_frame(256, 192, 4e-3, 2e-3, scene_seed=1, noise_seed=3)
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    assert not torch.equal(a[0], c)
    ycc, raw, clean = (to_host(t) for t in a)
    assert np.isfinite(ycc).all() and raw.min() >= 0 and raw.max() <= 4095
    # noiseless scene lies in [0.02, 0.9] per RGB channel: Y of it too
    assert clean[..., 0].min() >= 0.02 - 1e-6 and clean[..., 0].max() <= 0.9 + 1e-6
    # the noise is unbiased and of the right size: DN/4095 - scene at red sites
    err = ycc[..., 0] - clean[..., 0]
    assert abs(err.mean()) < 2e-3
    assert 0.01 < err.std() < 0.06


@pytest.mark.parametrize("hin,win,hout,wout", [(64, 96, 32, 48), (189, 252, 94, 126), (37, 51, 17, 23)])
def test_downsample_parity(hin, win, hout, wout):
    img = rand_img(hin, win, 1)
    out = to_host(gcam.downsample_image(to_dev(img), torch.empty((hout, wout, 4), device=DEV)))
    np.testing.assert_allclose(out, oracle.downsample(img, wout, hout), atol=1e-6)
    g = rand_img(hin, win, 2, ch=2)
    outg = to_host(gcam.downsample_image(to_dev(g), torch.empty((hout, wout, 2), device=DEV)))
    np.testing.assert_allclose(outg, oracle.downsample(g, wout, hout), atol=1e-6)


@pytest.mark.parametrize("sharpening,with_grad", [(1.0, False), (1.5, True), (1.5, False)])
@pytest.mark.parametrize("h,w,h1,w1", [(64, 96, 32, 48), (47, 33, 23, 16)])
def test_subtract_noise_parity(h, w, h1, w1, sharpening, with_grad):
    img, img1, den1 = rand_img(h, w, 3), rand_img(h1, w1, 4), rand_img(h1, w1, 5)
    grad = np.random.default_rng(6).normal(0, 0.05, (h, w, 2)).astype(np.float32) if with_grad else None
    out = gcam.subtract_noise_image(to_dev(img), to_dev(img1), to_dev(den1),
                                    to_dev(grad) if grad is not None else None, 0.8, sharpening, (2e-4, 3e-3))
    ref = oracle.subtract_noise(img, img1, den1, grad, 0.8, sharpening, (2e-4, 3e-3))
    np.testing.assert_allclose(to_host(out), ref, atol=2e-6)


@pytest.mark.parametrize("h,w", [(64, 64), (189, 252), (100, 37)])
def test_collect_patches_bit_exact(h, w):
    img = rand_img(h, w, 7) + 1.0
    got = to_host(gcam.collect_patches(to_dev(img)))
    np.testing.assert_array_equal(got, oracle.collect_patches(img))


@pytest.mark.parametrize("n", [16, 744, 20000])
def test_build_pca_space_parity(n):
    rng = np.random.default_rng(n)
    pts = (rng.normal(size=(n, 25)) @ rng.normal(size=(25, 25)) * 0.05 + 0.4).astype(np.float32)
    basis, ev = gcam.build_pca_space(to_dev(pts))
    basis, ev = to_host(basis).astype(np.float64), to_host(ev)
    rb, rev = oracle.build_pca_space(pts)
    np.testing.assert_allclose(ev[:8], rev[:8], rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(basis @ basis.T, rb @ rb.T, atol=1e-6)
    np.testing.assert_allclose(basis, rb, atol=1e-6)  # same sign convention
    np.testing.assert_allclose(basis.T @ basis, np.eye(8), atol=1e-6)


@pytest.mark.parametrize("n", [744, 20000])
def test_build_pca_space_all_eigenvalues(n):
    """Every eigenvalue (the leading eight and the other seventeen) against the oracle's Jacobi."""
    rng = np.random.default_rng(n + 1)
    pts = (rng.normal(size=(n, 25)) @ rng.normal(size=(25, 25)) * 0.05 + 0.4).astype(np.float32)
    _, ev = gcam.build_pca_space(to_dev(pts))
    _, rev = oracle.build_pca_space(pts)
    np.testing.assert_allclose(to_host(ev), rev, rtol=0, atol=1e-12 * abs(rev[0]))


@pytest.mark.parametrize("kind", ["equal", "zero", "near"])
def test_build_pca_space_degenerate_spectra(kind):
    """An exactly diagonal covariance with an 8-fold leading eigenvalue (inverse iteration
    cannot separate the eigenvectors, so the solve falls back to the Jacobi sweeps, which
    keep index order for ties like the oracle), the zero matrix, and a near-degenerate
    spectrum (eigenvalues equal up to rounding: only the spanned subspace is defined)."""
    if kind == "equal":
        pts = np.zeros((17, 25), np.float32)
        for i in range(8):
            pts[2 * i, i], pts[2 * i + 1, i] = 1.0, -1.0
    elif kind == "zero":
        pts = np.full((64, 25), 0.25, np.float32)
    else:
        pts = np.full((17, 25), 0.4, np.float32)
        for i in range(8):
            pts[2 * i, i] += 2.0
            pts[2 * i + 1, i] -= 2.0
    basis, ev = gcam.build_pca_space(to_dev(pts))
    basis, ev = to_host(basis).astype(np.float64), to_host(ev)
    rb, rev = oracle.build_pca_space(pts)
    np.testing.assert_allclose(ev, rev, rtol=0, atol=1e-12 * max(abs(rev[0]), 1e-30))
    np.testing.assert_allclose(basis.T @ basis, np.eye(8), atol=1e-6)
    np.testing.assert_allclose(basis @ basis.T, rb @ rb.T, atol=1e-6)
    if kind != "near":
        np.testing.assert_array_equal(basis, rb)


def test_image_moments_match_collected_patches():
    """The fused path (moments straight from the image) equals collect + build_pca_space."""
    img = rand_img(189, 252, 8)
    p = gcam.PyramidProcessor(252, 189, levels=1, search_radius=2)
    nlf = [((1e-5,) * 3, (1e-3,) * 3)]
    p.denoise([gcam.DenoiseParameters()], to_dev(img), None, nlf)
    fused_b, fused_e = p.pcaSpace(0)
    rb, rev = oracle.build_pca_space(oracle.collect_patches(img))
    np.testing.assert_allclose(fused_e.numpy()[:8], rev[:8], rtol=1e-9)
    np.testing.assert_allclose(fused_b.numpy(), rb, atol=1e-6)


@pytest.mark.parametrize("h,w", [(256, 384), (211, 301), (100, 70)])
def test_subtracted_level_moments(h, w):
    """Below the top level the PCA moments are formed by the subtract kernel
    as it produces the level (tiles that hold whole patches, the sample never
    written): each level's eigenvalues and basis equal collect_patches +
    build_pca_space of the subtracted level to FP64 rounding (the same
    patches, summed in another order), and the oracle's.  Sizes: multiples of
    8, odd sizes (a partial patch column and row, the right and bottom tiles
    past the image), and a level with fewer tiles than XCDs."""
    L = 3
    img = rand_img(h, w, 12)
    p = gcam.PyramidProcessor(w, h, levels=L, search_radius=2)
    nlf = [((1e-5,) * 3, (1e-3,) * 3)] * L
    p.denoise([gcam.DenoiseParameters() for _ in range(L)], to_dev(img), None, nlf)
    subs = p.subtractedImagePyramid
    for i in range(L - 1):
        fb, fe = (t.numpy() for t in p.pcaSpace(i))
        sub = subs[i].contiguous()
        cb, ce = (to_host(t) for t in gcam.build_pca_space(gcam.collect_patches(sub)))
        np.testing.assert_allclose(fe, ce, rtol=1e-12, atol=1e-14 * abs(ce).max())
        np.testing.assert_allclose(fb, cb, atol=1e-6)
        rb, rev = oracle.build_pca_space(oracle.collect_patches(to_host(sub)))
        np.testing.assert_allclose(fe[:8], rev[:8], rtol=1e-9)
        np.testing.assert_allclose(fb, rb, atol=1e-6)
    p.close()


@pytest.mark.parametrize("h,w", [(48, 64), (37, 29)])
def test_pca_projection_parity(h, w):
    img = rand_img(h, w, 9)
    basis = np.random.default_rng(10).normal(size=(25, 8)).astype(np.float32)
    got = to_host(gcam.pca_projection(to_dev(img), to_dev(basis)))
    np.testing.assert_allclose(got, oracle.pca_project(img, basis), atol=2e-6)


def bm_case(h, w, seed):
    img = rand_img(h, w, seed)
    img[..., 0] = np.clip(img[..., 0], 0.05, 1)
    basis, _ = oracle.build_pca_space(oracle.collect_patches(img))
    feat = oracle.pca_project(img, basis.astype(np.float32))
    return img, feat


@pytest.mark.parametrize("radius", [1, 3, 4, 7, 10])
def test_block_match_parity(radius):
    img, feat = bm_case(67, 45, 11 + radius)
    grad = np.random.default_rng(12).normal(0, 0.02, img.shape[:2] + (2,)).astype(np.float32)
    args = dict(var_a=[2e-4, 1e-4, 1e-4], var_b=[2e-3, 1e-3, 1.5e-3], thresholdMultipliers=[3.0, 2.0, 2.5],
                chromaBoost=1.7, gradientBoost=1.5, gradientThreshold=2.0, lensShadingCorrection=0.8)
    got = to_host(gcam.block_matching_denoise_image(to_dev(img), to_dev(feat), search_radius=radius,
                                                    gradient=to_dev(grad), **args))
    ref = oracle.block_match(img, feat, args["var_a"], args["var_b"], args["thresholdMultipliers"],
                             args["chromaBoost"], args["gradientBoost"], args["gradientThreshold"],
                             args["lensShadingCorrection"], radius, grad=grad)
    np.testing.assert_allclose(got[..., :3], ref[..., :3], atol=1e-5)
    np.testing.assert_allclose(got[..., 3], ref[..., 3], rtol=1e-5)


def test_block_match_no_gradient_equals_zero_gradient():
    img, feat = bm_case(40, 40, 13)
    args = dict(var_a=[1e-4] * 3, var_b=[1e-3] * 3, thresholdMultipliers=[2.0] * 3, gradientBoost=2.0,
                search_radius=4)
    a = gcam.block_matching_denoise_image(to_dev(img), to_dev(feat), **args)
    b = gcam.block_matching_denoise_image(to_dev(img), to_dev(feat), gradient=torch.zeros((40, 40, 2), device=DEV),
                                          **args)
    assert torch.equal(a, b)


def test_noise_statistics_parity():
    img = rand_img(57, 83, 14)
    got = to_host(gcam.basic_noise_statistics(to_dev(img)))
    np.testing.assert_allclose(got, oracle.noise_stats(img), atol=1e-6)


def test_measure_ycbcr_nlf_parity():
    ycc = gcam.synth_frame(512, 384, 4e-3, 2e-3, scene_seed=3, noise_seed=4)
    A, B = gcam.measure_ycbcr_nlf(ycc, exposure_multiplier=1.5)
    fit = oracle.fit_ycbcr_nlf(oracle.noise_stats(to_host(ycc)), exposure_multiplier=1.5)
    np.testing.assert_allclose(A, fit.A, rtol=1e-4)
    np.testing.assert_allclose(B, fit.B, rtol=1e-4)


def test_errors_are_raised_not_silent():
    img = torch.zeros((32, 32, 4), device=DEV)
    with pytest.raises(gcam.GcamError, match="expected 16x16"):
        gcam.subtract_noise_image(img, img[:16, :16].contiguous(), img[:8, :8].contiguous())
    with pytest.raises(gcam.GcamError, match="search radius"):
        gcam.block_matching_denoise_image(img, torch.zeros((32, 32, 8), device=DEV), [1] * 3, [1] * 3, [1] * 3,
                                          search_radius=11)


@pytest.mark.parametrize("gradient_boost", [0.0, 1.5])
def test_denoise_image_parity(gradient_boost):
    img = rand_img(53, 71, 15)
    img[..., 0] = np.clip(img[..., 0], 0.05, 1)
    grad = np.random.default_rng(16).normal(0, 0.05, img.shape[:2] + (2,)).astype(np.float32)
    args = ([1e-4, 2e-4, 2e-4], [1e-2, 5e-3, 5e-3], [1.5, 1.0, 1.2], 1.3, gradient_boost, 30.0)
    got = to_host(gcam.denoise_image(to_dev(img), *args, gradient=to_dev(grad)))
    ref = oracle.denoise_image(img, *args, grad=grad)
    np.testing.assert_allclose(got[..., :3], ref[..., :3], atol=1e-5)
    np.testing.assert_allclose(got[..., 3], ref[..., 3], atol=1e-6)


def test_pyramid_bilateral_mode_parity(oracle_threads):
    from tests.helpers import nlf_for
    ycc = gcam.synth_frame(320, 240, 4e-3, 2e-3, scene_seed=17, noise_seed=18)
    nlf = [(tuple(x[0]), tuple(x[1])) for x in nlf_for(3)]
    params = [gcam.DenoiseParameters(luma=2.0, chromaBoost=1.5) for _ in range(3)]
    proc = gcam.PyramidProcessor(320, 240, levels=3, usePatchSimiliarity=False)
    proc.denoise(params, ycc, None, list(nlf))
    ref = oracle.pyramid_denoise(to_host(ycc), params, np.array(nlf), levels=3, patch_similarity=False)
    for i, lvl in enumerate(proc.denoisedImagePyramid):
        np.testing.assert_allclose(to_host(lvl), ref.denoised[i], atol=1e-4, err_msg=f"level {i}")
