"""Raw front end on the GPU against the FP32 oracle.

Scale, Sobel, the highlight blend, the transform and the despeckle evaluate
the same FP32 operations in the same order as the oracle (no contraction,
correctly rounded division and square root) and are compared bit for bit.  The
gradient blur evaluates its narrow kernel (which decides per pixel between the
two radii) tap by tap in the oracle's order, so the decision and every
narrow-kernel pixel are bit-exact; the wide kernel is evaluated separably
(one horizontal pass per Sobel row, one vertical pass, a corner correction),
which equals the tap evaluation to FP32 rounding: magnitudes within 1e-5
relative, and the sign (that of the blurred signed gradient) may differ only
where that sum is within rounding of zero, which the demosaic never sees (it
reads |gx|, |gy|).  The demosaic takes the gradient direction from atan2f,
whose device and host implementations differ in the last bits, and its
quotients and square roots from the hardware reciprocal and square root
(1 ulp); both enter continuously (through mix weights and smoothstep ramps), so
the demosaic is held to 1e-5 absolute on values in [0.1, 1].
"""
import numpy as np
import pytest
import torch

import glasscamerapublic_amd as gcam
import oracle
from tests.helpers import nlf_for

pytestmark = pytest.mark.gpu

DEMOSAIC_ATOL = 1e-5
SCALE_12BIT = 65535.0 / 4095.0


def raw_frame(w, h, seed, a=4e-3, b=2e-3):
    """A 12-bit RGGB synthetic frame (device int16 storage) and its uint16 host copy."""
    _, raw = gcam.synth_frame(w, h, a, b, scene_seed=seed, noise_seed=seed + 1, want_raw=True)
    return raw, raw.cpu().numpy().view(np.uint16)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def check_gradient(got, want, narrow_px=None):
    """The blurred gradient against the oracle's: bit-exact on the narrow-kernel pixels, magnitudes of the
    wide-kernel ones within 1e-5 relative (+1e-7 of the largest), at most a handful of sign flips."""
    if narrow_px is not None:
        np.testing.assert_array_equal(got[narrow_px], want[narrow_px])
    scale = float(np.abs(want).max())
    np.testing.assert_allclose(np.abs(got), np.abs(want), rtol=1e-5, atol=1e-7 * scale)
    flips = int(np.sum(np.sign(got) != np.sign(want)))
    assert flips <= 2 + got.size // 1_000_000, f"{flips} sign flips"
    err = float(np.abs(np.abs(got) - np.abs(want)).max())
    print(f"gradient: max |magnitude| difference {err:.3g} (scale {scale:.3g}), sign flips {flips}")


@pytest.mark.parametrize("bayer,lens", [(0, 0.0), (1, 0.4), (2, 0.0), (3, 0.25)])
def test_scale_raw_is_bit_exact(bayer, lens):
    rng = np.random.default_rng(bayer)
    h, w = 130, 250
    raw = rng.integers(0, 65536, (h, w)).astype(np.uint16)
    sm = np.array([2.0, 1.0, 1.6, 1.02], np.float32) * SCALE_12BIT
    got = gcam.scaleRawData(dev(raw.view(np.int16)), bayer, sm, 0.003, lens).cpu().numpy()
    want = oracle.scale_raw(raw, bayer, sm, 0.003, lens)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("w,h", [(256, 192), (250, 130), (66, 18), (64, 32)])
def test_sobel_bit_exact_and_gradient_blur(w, h):
    raw_d, raw = raw_frame(w, h, 11)
    scaled = oracle.scale_raw(raw, 2, [SCALE_12BIT] * 4)
    sob = oracle.raw_sobel(scaled)
    got_sob = gcam.rawImageSobel(dev(scaled)).cpu().numpy()
    np.testing.assert_array_equal(got_sob, sob)
    rv = oracle.raw_variance((2e-6,) * 4, (4e-4,) * 4)
    want = oracle.blur_sobel(scaled, sob, rv[1])
    given = gcam.gaussianBlurSobelImage(dev(scaled), dev(sob), rv[1]).cpu().numpy()
    fused = gcam.gaussianBlurSobelImage(dev(scaled), None, rv[1]).cpu().numpy()
    narrow = oracle.blur_sobel(scaled, sob, (0.0, 0.0))  # sigma 0: every pixel keeps the narrow kernel
    narrow_px = np.all(want == narrow, axis=-1)
    np.testing.assert_array_equal(given, fused)  # Sobel tile read from an image or computed in the tile
    check_gradient(fused, want, narrow_px)
    wide_frac = float(np.mean(~narrow_px))
    print(f"{w}x{h}: wide-kernel pixels {wide_frac:.3f}")
    assert 0.05 < wide_frac < 0.995  # both branches of the blur are exercised


@pytest.mark.parametrize("bayer", [0, 1, 2, 3])
def test_demosaic_matches_the_oracle(bayer):
    h, w = 192, 256
    _, raw = raw_frame(w, h, 21 + bayer)
    scaled = oracle.scale_raw(raw, bayer, [SCALE_12BIT * g for g in (1.9, 1.0, 1.5, 1.0)])
    rv = oracle.raw_variance((2e-6,) * 4, (4e-4,) * 4)
    grad = oracle.blur_sobel(scaled, oracle.raw_sobel(scaled), rv[1])
    g_o, t_o, rgb_o = oracle.demosaic(scaled, grad, bayer, rv)
    g_d, t_d, rgb_d = (x.cpu().numpy() for x in gcam.demosaicImage(dev(scaled), dev(grad), bayer, rv))
    for name, got, want in (("green", g_d, g_o), ("red/blue", t_d, t_o), ("rgb", rgb_d, rgb_o)):
        err = float(np.abs(got - want).max())
        print(f"bayer {bayer} {name}: max abs {err:.3g}")
        assert err <= DEMOSAIC_ATOL, name


def test_highlights_transform_despeckle_are_bit_exact():
    rng = np.random.default_rng(31)
    h, w = 96, 130
    img = np.zeros((h, w, 4), np.float32)
    img[..., :3] = rng.uniform(0.05, 1.8, (h, w, 3))
    got = gcam.blendHighlightsImage(dev(img), 1.0).cpu().numpy()
    np.testing.assert_array_equal(got, oracle.blend_highlights(img, 1.0))
    m = oracle.cam_ycbcr(np.array([[1.6, -0.4, -0.2], [-0.2, 1.4, -0.2], [0.0, -0.5, 1.5]]))
    ycc = oracle.transform(img, m)
    np.testing.assert_array_equal(gcam.transformImage(dev(img), m).cpu().numpy(), ycc)
    va, vb = (2e-5, 1e-5, 1e-5), (1.5e-3, 5e-4, 5e-4)
    np.testing.assert_array_equal(gcam.despeckleImage(dev(ycc), va, vb).cpu().numpy(),
                                  oracle.despeckle(ycc, va, vb))
    # luma extremes with NaN neighbours (ignored by the running extremes), exact
    # zeros of both signs, negative luma and ties
    odd = ycc.copy()
    odd[10, 10:14, 0] = np.nan
    odd[20:23, 30:33, 0] = 0.0
    odd[21, 31, 0] = -0.0
    odd[40:43, 50:53, 0] = -0.0
    odd[60:63, 70:73, 0] = -0.25
    odd[80:83, 90:93, 0] = 0.5
    got = gcam.despeckleImage(dev(odd), va, vb).cpu().numpy()
    np.testing.assert_array_equal(got, oracle.despeckle(odd, va, vb))


def params_for(bayer=2, despeckle=True):
    rgb_cam = [[1.7, -0.5, -0.2], [-0.25, 1.45, -0.2], [0.05, -0.6, 1.55]]
    nl = nlf_for(1)[0]
    return gcam.FrontEndParameters(
        bayerPattern=bayer, black_level=0.001, lensShadingCorrection=0.2,
        scale_mul=[SCALE_12BIT * g for g in (1.8, 1.0, 1.4, 1.0)],
        rawNlf=((2e-6, 1e-6, 3e-6, 1e-6), (5e-4, 3e-4, 6e-4, 3e-4)), rgb_cam=rgb_cam, highlightClip=1.0,
        despeckleNlf=(tuple(nl[0]), tuple(nl[1])) if despeckle else None)


def oracle_front(raw, p):
    return oracle.raw_front_end(raw, p.bayerPattern, p.scale_mul, p.black_level, p.lensShadingCorrection,
                                p.rawNlf[0], p.rawNlf[1], np.array(p.rgb_cam), p.highlightClip, p.despeckleNlf)


@pytest.mark.parametrize("w,h,despeckle", [(512, 384, True), (258, 130, False)])
def test_front_end_matches_the_oracle(w, h, despeckle):
    raw_d, raw = raw_frame(w, h, 41)
    p = params_for(despeckle=despeckle)
    fe = gcam.RawFrontEnd(w, h)
    ycc, grad = fe.run(raw_d, p)
    torch.cuda.synchronize()
    ycc_o, grad_o, _ = oracle_front(raw, p)
    check_gradient(grad.cpu().numpy(), grad_o)
    err = float(np.abs(ycc.cpu().numpy() - ycc_o).max())
    print(f"{w}x{h} despeckle={despeckle}: YCbCr max abs {err:.3g}")
    assert err <= 2 * DEMOSAIC_ATOL  # the transform's rows sum to at most ~1.5 in magnitude
    check_per_stage_chain(raw_d, p, ycc, grad)
    fe.close()


def check_per_stage_chain(raw_d, p, ycc, grad):
    """The per-operation entry points compose to the front end's result bit for bit."""
    scaled = gcam.scaleRawData(raw_d, p.bayerPattern, p.scale_mul, p.black_level, p.lensShadingCorrection)
    rv = gcam.getRawVariance(p.rawNlf)
    g2 = gcam.gaussianBlurSobelImage(scaled, gcam.rawImageSobel(scaled), rv[1])
    assert torch.equal(g2, grad)
    _, _, rgb = gcam.demosaicImage(scaled, g2, p.bayerPattern, rv)
    y2 = gcam.transformImage(gcam.blendHighlightsImage(rgb, p.highlightClip), gcam.cam_ycbcr(p.rgb_cam))
    if p.despeckleNlf is not None:
        y2 = gcam.despeckleImage(y2, *p.despeckleNlf)
    assert torch.equal(y2, ycc)


# The front end's demosaic runs as one pass over 48 x 32 tiles with every
# intermediate in LDS; these sizes cover images smaller than a tile, exact and
# partial tiles at the right and bottom edges and every Bayer pattern.
@pytest.mark.parametrize("w,h,bayer,despeckle", [
    (2, 2, 0, True), (6, 4, 1, False), (50, 34, 3, True), (48, 32, 2, False), (96, 64, 1, True),
    (130, 98, 0, False), (1030, 770, 1, True), (4032, 96, 2, True), (200, 1000, 3, False)])
def test_tiled_demosaic_equals_the_per_stage_chain(w, h, bayer, despeckle):
    print(f"tiled demosaic {w}x{h} bayer {bayer} despeckle {despeckle}", flush=True)
    raw_d, _ = raw_frame(w, h, 7 + bayer)
    p = params_for(bayer=bayer, despeckle=despeckle)
    fe = gcam.RawFrontEnd(w, h)
    ycc, grad = fe.run(raw_d, p)
    check_per_stage_chain(raw_d, p, ycc, grad)
    fe.close()


def test_raw_converter_chain(oracle_threads):
    """Raw frame -> front end -> FP32 pyramid denoise (with the front end's
    gradient) -> camera RGB, against the same chain on the oracle.  The
    pyramid's own bar is 1e-4; the return to RGB multiplies by up to ~2."""
    w, h, levels, radius = 384, 256, 4, 3
    raw_d, raw = raw_frame(w, h, 51)
    p = params_for()
    nl = [(tuple(a), tuple(b)) for a, b in nlf_for(levels)]
    prm = [gcam.DenoiseParameters(luma=1.5, chroma=1.2, chromaBoost=2.0, gradientBoost=2.0, gradientThreshold=1.5,
                                  sharpening=1.2) for _ in range(levels)]
    conv = gcam.RawConverter(w, h, levels=levels, search_radius=radius, precision="fp32")
    rgb = conv.demosaic(raw_d, p, prm, list(nl)).cpu().numpy()
    ycc_o, grad_o, _ = oracle_front(raw, p)
    pr = [dict(luma=q.luma, chroma=q.chroma, chromaBoost=q.chromaBoost, gradientBoost=q.gradientBoost,
               gradientThreshold=q.gradientThreshold, sharpening=q.sharpening) for q in prm]
    den = oracle.pyramid_denoise(ycc_o, pr, np.array([[a, b] for a, b in nl]), levels=levels, radius=radius,
                                 gradient=grad_o, lens_shading=p.lensShadingCorrection).denoised[0]
    back = np.linalg.inv(oracle.cam_ycbcr(np.array(p.rgb_cam)).astype(np.float64)).astype(np.float32)
    want = oracle.transform(den, back)
    err = float(np.abs(rgb[..., :3] - want[..., :3]).max())
    print(f"camera RGB max abs {err:.3g}")
    assert err <= 5e-4
    conv.close()


def test_front_end_argument_errors():
    with pytest.raises(gcam.GcamError):
        gcam.RawFrontEnd(251, 130)
    fe = gcam.RawFrontEnd(64, 32)
    with pytest.raises(ValueError):
        fe.run(torch.zeros((32, 66), dtype=torch.int16, device="cuda"), params_for())
    bad = params_for()
    bad.bayerPattern = 7
    with pytest.raises(ValueError):
        fe.run(torch.zeros((32, 64), dtype=torch.int16, device="cuda"), bad)
    with pytest.raises(gcam.GcamError):  # a blur radius beyond the tile apron
        gcam.gaussianBlurSobelImage(torch.zeros((32, 64), device="cuda"), None, (1e-5, 1e-3), radius2=8.0)
    fe.close()


@pytest.mark.parametrize("bayer", [0, 1, 2, 3])
def test_raw_noise_statistics_are_bit_exact(bayer):
    _, raw = raw_frame(258, 130, 61 + bayer)
    scaled = oracle.scale_raw(raw, bayer, [SCALE_12BIT] * 4)
    mean, var = gcam.rawNoiseStatistics(dev(scaled), bayer)
    m_o, v_o = oracle.raw_noise_stats(scaled, bayer)
    np.testing.assert_array_equal(mean.cpu().numpy(), m_o)
    np.testing.assert_array_equal(var.cpu().numpy(), v_o)


def test_measure_raw_nlf_matches_the_oracle():
    """The FP64 fit on the device (block sums in a fixed order) against the
    oracle's sequential FP64 sums: the same quads are selected, so the models
    agree to the rounding of the sums."""
    w, h = 1024, 768
    _, raw = raw_frame(w, h, 71, a=6e-3, b=3e-3)
    scaled = oracle.scale_raw(raw, 2, [SCALE_12BIT * g for g in (1.9, 1.0, 1.5, 1.0)])
    a, b = gcam.measureRawNLF(dev(scaled), 2, exposure_multiplier=1.5)
    fit = oracle.measure_raw_nlf(scaled, 2, exposure_multiplier=1.5)
    print("device", a, b, "oracle", fit.a, fit.b, "n1", fit.n1, "n2", fit.n2, "accepted", fit.accepted)
    np.testing.assert_allclose(np.array(a), fit.a, rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(np.array(b), fit.b, rtol=1e-5)


def test_front_end_calibrating_the_raw_model():
    w, h = 512, 384
    raw_d, raw = raw_frame(w, h, 81)
    p = params_for()
    p.calibrateRawNlf = True
    fe = gcam.RawFrontEnd(w, h)
    ycc, grad = fe.run(raw_d, p)
    fitted = fe.rawNlf
    scaled = gcam.scaleRawData(raw_d, p.bayerPattern, p.scale_mul, p.black_level, p.lensShadingCorrection)
    assert fitted == gcam.measureRawNLF(scaled, p.bayerPattern)
    q = params_for()
    q.rawNlf = fitted
    ycc2, grad2 = fe.run(raw_d, q)
    assert fe.rawNlf == fitted
    assert torch.equal(grad, grad2) and torch.equal(ycc, ycc2)
    fit = oracle.measure_raw_nlf(oracle.scale_raw(raw, p.bayerPattern, p.scale_mul, p.black_level,
                                                  p.lensShadingCorrection), p.bayerPattern)
    np.testing.assert_allclose(np.array(fitted[1]), fit.b, rtol=1e-5)
    fe.close()


def test_histogram_is_exact():
    rng = np.random.default_rng(91)
    img = np.zeros((378, 504, 4), np.float32)
    img[..., 0] = rng.uniform(-0.02, 1.4, (378, 504)).astype(np.float32)
    got = gcam.histogramImage(dev(img))
    want = oracle.histogram(img)
    np.testing.assert_array_equal(got["histogram"], want["histogram"])
    np.testing.assert_array_equal(got["bands"], want["bands"])
    for k in ("black_level", "white_level", "shadows", "highlights", "mean", "median"):
        assert got[k] == want[k], k


def test_raw_converter_histogram_of_level_3():
    w, h = 512, 384
    raw_d, _ = raw_frame(w, h, 95)
    p = params_for()
    nl = [(tuple(a), tuple(b)) for a, b in nlf_for(5)]
    conv = gcam.RawConverter(w, h, levels=5, search_radius=4, precision="fp16")
    conv.demosaic(raw_d, p, [gcam.DenoiseParameters() for _ in range(5)], list(nl))
    lvl3 = conv.pyramid.denoisedImagePyramid[3].cpu().numpy()
    want = oracle.histogram(lvl3)
    np.testing.assert_array_equal(conv.histogramData["histogram"], want["histogram"])
    assert conv.histogramData["median"] == want["median"]
    conv.close()


def test_ltm_filters_bit_exact_and_mask():
    rng = np.random.default_rng(101)
    h, w = 128, 192
    img = np.zeros((h, w, 4), np.float32)
    img[..., 0] = (0.1 + 0.7 * rng.random((h, w))).astype(np.float32)
    guides = [np.ascontiguousarray(img[::16, ::16]), np.ascontiguousarray(img[::4, ::4]), img]
    ab = [gcam.guidedFilterABImage(dev(g), 0.01) for g in guides]
    means = [gcam.boxFilterGFImage(a) for a in ab]
    grad = (0.02 * rng.standard_normal((h, w, 2))).astype(np.float32)
    mask_o, means_o = oracle.ltm_mask(img, grad, guides, nlf=(1e-5, 1e-3), hist_shadows=0.75, hist_highlights=1.4)
    for a_d, g in zip(ab, guides):
        np.testing.assert_array_equal(a_d.cpu().numpy(), oracle.guided_filter_ab(g, 0.01))
    for m_d, m_o in zip(means, means_o):
        np.testing.assert_array_equal(m_d.cpu().numpy(), m_o)
    hist = torch.zeros(270, dtype=torch.int32, device="cuda")
    hist.view(torch.float32)[264 + 2] = 0.75  # shadows
    hist.view(torch.float32)[264 + 3] = 1.4   # highlights
    ltm = gcam.LocalToneMapping(w, h)
    ltm.ab_mean = means
    import ctypes
    from glasscamerapublic_amd._lib import check
    from glasscamerapublic_amd.front_end import _img
    pc = gcam.LTMParameters().to_c()
    img_d, grad_d = dev(img), dev(grad)  # held until the launch (the image structs keep only pointers)
    check(gcam.lib().gcam_local_tone_mapping_mask_image(
        ctypes.byref(_img(img_d)), ctypes.byref(_img(grad_d)), ctypes.byref(_img(means[0])),
        ctypes.byref(_img(means[1])), ctypes.byref(_img(means[2])), ctypes.byref(pc), 1e-5, 1e-3,
        ctypes.c_void_p(hist.data_ptr()), ctypes.byref(_img(ltm.mask)), None))
    torch.cuda.synchronize()
    np.testing.assert_allclose(ltm.mask.cpu().numpy(), mask_o, rtol=1e-5)


def test_raw_converter_ltm_mask():
    w, h = 512, 384
    raw_d, _ = raw_frame(w, h, 111)
    p = params_for()
    nl = [(tuple(a), tuple(b)) for a, b in nlf_for(5)]
    conv = gcam.RawConverter(w, h, levels=5, search_radius=4, precision="fp16")
    conv.demosaic(raw_d, p, [gcam.DenoiseParameters() for _ in range(5)], list(nl),
                  ltmParameters=gcam.LTMParameters())
    d = [t.cpu().numpy() for t in conv.pyramid.denoisedImagePyramid]
    grad = conv.front.run(raw_d, p)[1].cpu().numpy()
    hd = conv.histogramData
    mask_o, _ = oracle.ltm_mask(d[0], grad, [d[4], d[2], d[0]], nlf=(nl[0][0][0], nl[0][1][0]),
                                hist_shadows=hd["shadows"], hist_highlights=hd["highlights"])
    np.testing.assert_allclose(conv.ltm.getMask().cpu().numpy(), mask_o, rtol=1e-5)
    conv.close()
