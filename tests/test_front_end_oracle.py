"""The raw front-end oracle pinned by the reference's own tables and by what
the mathematics fixes (CPU only).

The reference holds no test vectors for the front end; what it does hold are
the sRGB YCbCr matrices (MetalPipeline/demosaic_utils.cpp:728-740), which
cam_ycbcr with an identity camera matrix and its inverse must reproduce.  The
rest is pinned by independent computations: a direct Gaussian convolution
against the folded bilinear taps, closed-form gradients of ramps, exact
demosaicing of grey ramps, solid colours and straight edges, invariants of the
highlight blend, and sort-based medians and extremes for the despeckle.
"""
import ctypes

import numpy as np
import pytest

import glasscamerapublic_amd as gcam
import oracle

# MetalPipeline/demosaic_utils.cpp:728-740
SRGB_YCBCR = np.array([[0.2126, 0.7152, 0.0722], [-0.1146, -0.3854, 0.5], [0.5, -0.4542, -0.0458]])
YCBCR_SRGB = np.array([[1, 0, 1.5748], [1, -0.1873, -0.4681], [1, 1.8556, 0]])


def test_cam_ycbcr_reproduces_the_srgb_tables():
    m = oracle.cam_ycbcr(np.eye(3))
    np.testing.assert_allclose(m, SRGB_YCBCR, atol=6e-5)
    np.testing.assert_allclose(np.linalg.inv(m.astype(np.float64)), YCBCR_SRGB, atol=6e-5)
    # the library's host-side versions: same matrix, and its inverse
    np.testing.assert_array_equal(gcam.cam_ycbcr(np.eye(3)), m)
    np.testing.assert_allclose(gcam.inverse3(m), YCBCR_SRGB, atol=6e-5)
    rgb_cam = np.array([[1.6, -0.4, -0.2], [-0.2, 1.4, -0.2], [0.0, -0.5, 1.5]])
    np.testing.assert_array_equal(gcam.cam_ycbcr(rgb_cam), oracle.cam_ycbcr(rgb_cam))
    np.testing.assert_allclose(gcam.inverse3(rgb_cam) @ rgb_cam, np.eye(3), atol=1e-6)


def test_raw_variance_layout():
    a, b = (1.0, 2.0, 3.0, 4.0), (5.0, 6.0, 7.0, 8.0)
    want = np.array([[1, 5], [3, 7], [3, 7]], np.float32)
    np.testing.assert_array_equal(oracle.raw_variance(a, b), want)
    np.testing.assert_array_equal(gcam.getRawVariance((a, b)), want)


@pytest.mark.parametrize("radius", [0.5, 1.5, 2.5, 4.5, 5.5])
def test_bilinear_taps_reproduce_the_gaussian(radius):
    taps = oracle.gaussian_bilinear_weights(radius)
    np.testing.assert_array_equal(taps, gcam.gaussianKernelBilinearWeights(radius))  # library host code
    k = max(int(np.ceil(2 * radius)), 2)
    k += 1 - k % 2
    half = k // 2
    assert taps.shape == ((half + 1) ** 2, 3)
    d = np.arange(-half, half + 1, dtype=np.float64)
    g = np.exp(-(d[None, :] ** 2 + d[:, None] ** 2) / (2 * radius * radius))
    np.testing.assert_allclose(taps[:, 0].sum(), g.sum(), rtol=1e-5)
    # Blur a random 4-channel image with the oracle (wide kernel forced by a
    # huge noise floor) and directly: every weight at its own texel, except the
    # corner weight, which the reference places half a texel beyond the corner
    # (offset k/2 + 1/2): there it averages the four texels around that point.
    rng = np.random.default_rng(3)
    h, w = 40, 48
    img = rng.standard_normal((h, w, 4)).astype(np.float32)
    out = oracle.blur_sobel(np.zeros((h, w), np.float32), img, (1e30, 0.0), radius1=0.5, radius2=radius)
    pad = np.pad(img.astype(np.float64), ((half + 2, half + 2), (half + 2, half + 2), (0, 0)), mode="edge")
    direct = np.zeros((h, w, 4))
    for dy in range(-half, half + 1):
        for dx in range(-half, half + 1):
            wt = g[dy + half, dx + half]
            if dx == half and dy == half:
                s = 0.25 * sum(pad[half + 2 + dy + j:half + 2 + dy + j + h, half + 2 + dx + i:half + 2 + dx + i + w]
                               for j in (0, 1) for i in (0, 1))
            else:
                s = pad[half + 2 + dy:half + 2 + dy + h, half + 2 + dx:half + 2 + dx + w]
            direct += wt * s
    direct /= g.sum()
    want = np.stack([np.copysign(np.abs(direct[..., 2]), direct[..., 0]),
                     np.copysign(np.abs(direct[..., 3]), direct[..., 1])], -1)
    # channels z, w of a Sobel image are >= 0; here they are signed, so compare |.|
    np.testing.assert_allclose(np.abs(out), np.abs(want), atol=2e-5)
    np.testing.assert_array_equal(np.sign(out), np.sign(want))


def test_blur_switches_to_the_wide_kernel_below_four_sigma():
    rng = np.random.default_rng(4)
    h, w = 24, 24
    sob = np.abs(rng.standard_normal((h, w, 4))).astype(np.float32)
    sob[..., 2:] = sob[..., :2]  # like a Sobel image of non-negative gradients: (x, y, |x|, |y|)
    raw = np.full((h, w), 0.5, np.float32)
    narrow = oracle.blur_sobel(raw, sob, (0.0, 0.0))   # sigma 0: never below
    wide = oracle.blur_sobel(raw, sob, (1e30, 0.0))     # sigma huge: always below
    assert not np.allclose(narrow, wide)
    res1 = oracle.blur_sobel(raw, sob, (0.0, 0.0), radius1=1.5, radius2=1.5)
    np.testing.assert_array_equal(res1, narrow)
    # a threshold between: pixels take the narrow result where its gradient is >= 4 sigma
    nx = oracle.blur_sobel(raw, sob, (0.0, 0.0), radius2=1.5)
    mag = np.hypot(*[np.abs(nx[..., c]) for c in (0, 1)])
    sigma2 = (np.median(mag) / 4) ** 2
    mixed = oracle.blur_sobel(raw, sob, (sigma2, 0.0))
    # the narrow blur's (x, y) magnitudes are those of channels z, w here (inputs >= 0)
    take_narrow = mag >= 4 * np.sqrt(np.float32(sigma2))
    np.testing.assert_array_equal(mixed[take_narrow], narrow[take_narrow])
    np.testing.assert_array_equal(mixed[~take_narrow], wide[~take_narrow])


def test_sobel_of_a_ramp_and_against_a_direct_convolution():
    h, w = 20, 24
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    s = oracle.raw_sobel(0.01 * x + 0.02 * y)
    inner = (slice(1, -1), slice(1, -1))
    np.testing.assert_allclose(s[inner][..., 0], -8 * 0.01 / np.sqrt(4.5), rtol=1e-4)
    np.testing.assert_allclose(s[inner][..., 1], -8 * 0.02 / np.sqrt(4.5), rtol=1e-4)
    rng = np.random.default_rng(5)
    img = rng.random((h, w)).astype(np.float32)
    p = np.pad(img.astype(np.float64), 1, mode="edge")
    sh = lambda dx, dy: p[1 + dy:1 + dy + h, 1 + dx:1 + dx + w]  # noqa: E731
    gx = (sh(-1, -1) + 2 * sh(-1, 0) + sh(-1, 1) - sh(1, -1) - 2 * sh(1, 0) - sh(1, 1)) / np.sqrt(4.5)
    gy = (sh(-1, -1) + 2 * sh(0, -1) + sh(1, -1) - sh(-1, 1) - 2 * sh(0, 1) - sh(1, 1)) / np.sqrt(4.5)
    out = oracle.raw_sobel(img)
    np.testing.assert_allclose(out[..., 0], gx, atol=1e-6)
    np.testing.assert_allclose(out[..., 1], gy, atol=1e-6)
    np.testing.assert_array_equal(out[..., 2], np.abs(out[..., 0]))
    np.testing.assert_array_equal(out[..., 3], np.abs(out[..., 1]))


@pytest.mark.parametrize("bayer", [0, 1, 2, 3])
def test_scale_raw_closed_form(bayer):
    rng = np.random.default_rng(6 + bayer)
    h, w = 12, 16
    raw = rng.integers(0, 4096, (h, w)).astype(np.uint16)
    sm = np.array([2.1, 1.0, 1.7, 1.05], np.float32) * 16
    out = oracle.scale_raw(raw, bayer, sm, black_level=0.002, lens_correction=0.3)
    offs = [((1, 0), (0, 0), (0, 1), (1, 1)), ((0, 1), (0, 0), (1, 0), (1, 1)),
            ((0, 0), (0, 1), (1, 1), (1, 0)), ((1, 1), (0, 1), (0, 0), (1, 0))][bayer]
    cx, cy = w // 2, h // 2
    for yy in range(h):
        for xx in range(w):
            c = [i for i, (ox, oy) in enumerate(offs) if ox == xx % 2 and oy == yy % 2][0]
            qx, qy = xx - xx % 2, yy - yy % 2
            d2 = ((qx - cx) ** 2 + (qy - cy) ** 2) / (cx * cx + cy * cy)
            v = (1 + 0.3 * d2) * sm[c] * (raw[yy, xx] / 65535 - 0.002) * 0.9 + 0.1
            assert abs(out[yy, xx] - max(v, 0.0)) < 2e-6


def mosaic(rgb, bayer=2):
    """Bayer mosaic of an (H, W, 3) image."""
    offs = [((1, 0), (0, 0), (0, 1), (1, 1)), ((0, 1), (0, 0), (1, 0), (1, 1)),
            ((0, 0), (0, 1), (1, 1), (1, 0)), ((1, 1), (0, 1), (0, 0), (1, 0))][bayer]
    out = np.empty(rgb.shape[:2], np.float32)
    for c, (ox, oy) in enumerate(offs):
        out[oy::2, ox::2] = rgb[oy::2, ox::2, [0, 1, 2, 1][c]]
    return out


RV = np.array([[1e-5, 2e-4], [1e-5, 2e-4], [1e-5, 2e-4]], np.float32)


@pytest.mark.parametrize("bayer", [0, 2, 3])
def test_demosaic_is_exact_on_grey_ramps_and_solid_colours(bayer):
    h, w = 32, 40
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    # Clamp-to-edge breaks the mosaic's parity at the border; the passes read
    # +-2, +-2 and +-3 pixels, so border effects reach 7 pixels in.
    inner = (slice(8, -8), slice(8, -8))
    ramp = 0.2 + 0.01 * x + 0.005 * y
    grad = np.zeros((h, w, 2), np.float32)
    _, _, rgb = oracle.demosaic(mosaic(np.stack([ramp] * 3, -1), bayer), grad, bayer, RV)
    for c in range(3):
        np.testing.assert_allclose(rgb[inner][..., c], ramp[inner], atol=2e-6)
    solid = np.empty((h, w, 3), np.float32)
    solid[...] = (0.6, 0.35, 0.2)
    _, _, rgb = oracle.demosaic(mosaic(solid, bayer), grad, bayer, RV)
    np.testing.assert_allclose(rgb[inner][..., :3], solid[inner], atol=2e-6)


def test_green_follows_the_edge_direction():
    """A grey vertical step: at red and blue sites beside it the gradient is
    horizontal, so green is interpolated vertically (along the edge) and is
    exact; a horizontal step likewise uses the horizontal neighbours."""
    h, w = 32, 32
    for vertical in (True, False):
        img = np.full((h, w), 0.2, np.float32)
        if vertical:
            img[:, 16:] = 0.7
        else:
            img[16:, :] = 0.7
        raw = mosaic(np.stack([img] * 3, -1), 2)
        grad = oracle.blur_sobel(raw, oracle.raw_sobel(raw), (1e-6, 1e-4))
        green, _, _ = oracle.demosaic(raw, grad, 2, RV)
        band = (slice(6, -6), slice(14, 18)) if vertical else (slice(14, 18), slice(6, -6))
        np.testing.assert_allclose(green[band], img[band], atol=1e-6)


def test_highlight_blend_invariants():
    rng = np.random.default_rng(8)
    px = np.concatenate([rng.uniform(0, 1, (50, 3)), rng.uniform(0.5, 2.0, (200, 3)),
                         np.full((5, 3), 1.4)]).astype(np.float32)
    img = np.concatenate([px, np.zeros((len(px), 1), np.float32)], 1).reshape(1, -1, 4)
    out = oracle.blend_highlights(img, 1.0)[0, :, :3].astype(np.float64)
    p = px.astype(np.float64)
    low = (p <= 1.0).all(1)
    np.testing.assert_array_equal(out[low], p[low])
    T = np.array([[1, 1, 1], [np.sqrt(3), -np.sqrt(3), 0], [-1, -1, 2]])
    hi = ~low
    np.testing.assert_allclose(out[hi].sum(1), p[hi].sum(1), rtol=1e-5)      # luma kept
    a, b, c = (T @ out[hi].T)[1:], (T @ p[hi].T)[1:], (T @ np.minimum(p[hi], 1.0).T)[1:]
    cross = a[0] * b[1] - a[1] * b[0]
    np.testing.assert_allclose(cross, 0, atol=1e-4)                           # hue kept
    np.testing.assert_allclose(np.hypot(*a), np.hypot(*c), rtol=1e-4, atol=1e-5)  # clipped chroma length
    np.testing.assert_allclose(out[-5:], 1.4, rtol=1e-6)                      # grey stays grey


def test_transform_is_a_matrix_product():
    rng = np.random.default_rng(9)
    img = rng.random((6, 7, 4)).astype(np.float32)
    m = rng.standard_normal((3, 3)).astype(np.float32)
    out = oracle.transform(img, m)
    np.testing.assert_allclose(out[..., :3], img[..., :3] @ m.T, atol=1e-6)
    assert (out[..., 3] == 0).all()


def test_despeckle_against_sort_based_statistics():
    rng = np.random.default_rng(10)
    h, w = 16, 18
    img = rng.uniform(0.1, 0.9, (h, w, 4)).astype(np.float32)
    img[4:7, 6:9, 0] = 0.5
    img[5, 7, 0] = 0.56  # a speckle 1.7 sigma above a flat neighbourhood (sigma ~ 0.035)
    va, vb = (1e-4, 1e-4, 1e-4), (2e-3, 2e-3, 2e-3)
    out = oracle.despeckle(img, va, vb)
    p = np.pad(img, ((1, 1), (1, 1), (0, 0)), mode="edge")
    for y in range(h):
        for x in range(w):
            nb = p[y:y + 3, x:x + 3].reshape(9, 4)
            assert out[y, x, 1] == np.median(nb[:, 1]) and out[y, x, 2] == np.median(nb[:, 2])
            lum = np.sort(np.concatenate([nb[:, 0], [0.0, 0.0]]))      # maxima start at 0
            lmin = np.sort(np.concatenate([nb[:, 0], [65535.0, 65535.0]]))
            max1, max2, min1, min2 = lum[-1], lum[-2], lmin[0], lmin[1]
            c = img[y, x, 0]
            sig = np.sqrt(np.float32(va[0]) + np.float32(vb[0]) * c)

            def ss(e0, e1, v):
                t = np.clip((v - e0) / (e1 - e0), 0, 1)
                return t * t * (3 - 2 * t)
            lo = min2 + (min1 - min2) * ss(sig, 4 * sig, min2 - min1)
            hi = max2 + (max1 - max2) * ss(sig, 4 * sig, max1 - max2)
            assert abs(out[y, x, 0] - min(max(c, lo), hi)) < 1e-6
    assert 0.5 < out[5, 7, 0] < 0.52  # the speckle is pulled most of the way in
    big = img.copy()
    big[5, 7, 0] = 0.95  # far above 4 sigma: a feature, kept
    assert oracle.despeckle(big, va, vb)[5, 7, 0] == np.float32(0.95)


def test_front_end_entry_points_validate_arguments_without_a_device():
    L = gcam.lib()
    f = ctypes.c_void_p()
    assert L.gcam_front_end_create(7, 8, ctypes.byref(f)) == 1   # odd width
    assert L.gcam_front_end_create(8, 8, None) == 1
    assert L.gcam_front_end_run(None, None, None, None, None, None) == 1
    sm = (ctypes.c_float * 4)(1, 1, 1, 1)
    assert L.gcam_scale_raw_data(None, None, 2, sm, 0.0, 0.0, None) == 1
    assert L.gcam_raw_image_sobel(None, None, None) == 1
    assert L.gcam_transform_image(None, None, None, None) == 1
    assert L.gcam_gaussian_kernel_bilinear_weights(-1.0, None, 0) == -1
    bad = (ctypes.c_float * 9)(*([0.0] * 9))
    out = (ctypes.c_float * 9)()
    assert L.gcam_matrix3_inverse(bad, out) == 1


def test_despeckle_with_a_negative_variance_clamps_to_the_second_extremes():
    """var_a + var_b * Y < 0 makes sigma a NaN; min, max and clamp follow fmin /
    fmax (Metal's definition of clamp), so both blends take the second extremes."""
    img = np.zeros((3, 3, 4), np.float32)
    img[..., 0] = [[-0.30, -0.10, -0.20], [-0.25, -0.05, -0.15], [-0.12, -0.22, -0.28]]
    out = oracle.despeckle(img, (1e-4,) * 3, (1e-2,) * 3)
    lum = np.sort(np.concatenate([img[..., 0].ravel(), [0.0, 0.0]]))
    lmin = np.sort(img[..., 0].ravel())
    assert out[1, 1, 0] == np.float32(min(max(-0.05, lmin[1]), lum[-2]))


@pytest.mark.parametrize("bayer", [0, 2])
def test_raw_noise_stats_closed_form_on_a_solid_mosaic(bayer):
    """The statistics window steps one pixel at a time, so a channel's 81 reads
    are its own colour where both offsets are even (25), the horizontally and
    vertically adjacent colours where one is odd (20 + 20) and the diagonal
    colour where both are odd (16)."""
    col = np.array([0.6, 0.3, 0.15, 0.33], np.float32)  # red, green, blue, green2
    solid = np.empty((24, 32, 3), np.float32)
    solid[...] = (col[0], col[1], col[2])
    raw = mosaic(solid, bayer)
    offs = [((1, 0), (0, 0), (0, 1), (1, 1)), ((0, 1), (0, 0), (1, 0), (1, 1)),
            ((0, 0), (0, 1), (1, 1), (1, 0)), ((1, 1), (0, 1), (0, 0), (1, 0))][bayer]
    g2 = offs[3]
    raw[g2[1]::2, g2[0]::2] = col[3]  # second green differs
    mean, var = oracle.raw_noise_stats(raw, bayer)
    inner = mean[3:-3, 3:-3], var[3:-3, 3:-3]
    colour_at = {}
    for c, (ox, oy) in enumerate(offs):
        colour_at[(ox, oy)] = col[c]
    for c, (ox, oy) in enumerate(offs):
        vals = np.array([colour_at[((ox + x) % 2, (oy + y) % 2)] for y in range(-4, 5) for x in range(-4, 5)], np.float64)
        np.testing.assert_allclose(inner[0][..., c], vals.mean(), rtol=1e-6)
        np.testing.assert_allclose(inner[1][..., c], vals.var(), rtol=1e-3, atol=1e-7)


def linear_stats(n, A, B, rng, lo=0.01, hi=0.8):
    m = rng.uniform(lo, hi, (n, 1, 4)).astype(np.float32)
    v = (np.asarray(A, np.float32) + np.asarray(B, np.float32) * m).astype(np.float32)
    return m, v


def test_raw_nlf_fit_recovers_an_exact_linear_model():
    rng = np.random.default_rng(21)
    A, B = np.array([2e-6, 1e-6, 3e-6, 1.5e-6]), np.array([4e-4, 3e-4, 5e-4, 3.2e-4])
    m, v = linear_stats(5000, A, B, rng)
    fit = oracle.fit_raw_nlf(m, v, 100, 200)
    np.testing.assert_allclose(fit.b, v[..., :].reshape(-1, 4).astype(np.float64).std(0) / m.reshape(-1, 4).std(0),
                               rtol=1e-4)  # slope of the data itself (float32-rounded)
    np.testing.assert_allclose(fit.b, B, rtol=1e-5)
    np.testing.assert_allclose(fit.a, A, rtol=2e-3, atol=1e-9)
    # largest variances ~4e-4: log10 -3.4 truncates to -3 (the reference's (int)), bin 3 -> ceiling 1e-2
    assert fit.n1 == 5000 and fit.variance_max == 1e-2
    fit2 = oracle.fit_raw_nlf(m, v, 100, 200, exposure_multiplier=2.0)
    np.testing.assert_allclose(fit2.a, 4 * fit.a, rtol=1e-12)
    np.testing.assert_allclose(fit2.b, 4 * fit.b, rtol=1e-12)


def test_raw_nlf_fit_ceiling_and_refit():
    """Quads above the histogram's variance ceiling are ignored; quads far from
    the first fit are dropped by the refit, which is then accepted."""
    rng = np.random.default_rng(22)
    A, B = np.array([2e-6] * 4), np.array([4e-4] * 4)
    m, v = linear_stats(6000, A, B, rng)
    v = v * (1 + 0.02 * rng.standard_normal(v.shape)).astype(np.float32)  # scatter around the line
    out_m, out_v = linear_stats(300, A, B, rng)
    out_v = out_v * 2.5                                                     # outliers below the ceiling
    hi_m, hi_v = linear_stats(200, A, B, rng)
    hi_v = hi_v + np.float32(0.05)                                          # above the 1e-2 ceiling
    M = np.concatenate([m, out_m, hi_m]).astype(np.float32)
    V = np.concatenate([v, out_v, hi_v]).astype(np.float32)
    fit = oracle.fit_raw_nlf(M, V, 200, 200)
    assert fit.n1 == 6300           # the 200 above the ceiling are never used
    assert fit.accepted and fit.n2 < 6000
    np.testing.assert_allclose(fit.b, B, rtol=0.05)
    plain = oracle.fit_raw_nlf(np.concatenate([m, out_m]), np.concatenate([v, out_v]), 200, 200)
    np.testing.assert_allclose(plain.b, fit.b, rtol=1e-12)   # identical: the ceiling excluded them exactly


def test_histogram_bins_and_statistics():
    rng = np.random.default_rng(30)
    img = np.zeros((40, 50, 4), np.float32)
    img[..., 0] = rng.uniform(-0.05, 2.5, (40, 50)).astype(np.float32)
    img[0, 0, 0] = np.nan
    d = oracle.histogram(img)
    y = img[..., 0].ravel()
    with np.errstate(invalid="ignore"):
        v = np.sqrt(y / np.float32(2)) * np.float32(255)
    idx = np.where(v >= 0, np.minimum(np.nan_to_num(v, nan=0).astype(np.int64), 255), 0)
    np.testing.assert_array_equal(d["histogram"], np.bincount(idx, minlength=256))
    np.testing.assert_array_equal(d["bands"], d["histogram"].reshape(8, 32).sum(1))
    for k in (3, 100, 200):   # every pixel in bin k: median, black and white all at k
        one = np.zeros((10, 20, 4), np.float32)
        one[..., 0] = 2 * ((k + 0.5) / 255) ** 2
        s = oracle.histogram(one)
        assert s["histogram"][k] == 200
        black = 2 * ((k - 1) / 255) ** 2
        assert abs(s["black_level"] - black) < 1e-6 and abs(s["white_level"] - 2 * (k / 255) ** 2) < 1e-6
        assert abs(s["median"] - (k / 255 - black)) < 1e-6 and abs(s["mean"] - ((k + 1) / 256 - black)) < 1e-6
        assert s["highlights"] == (3.0 if k >= 160 else 1.0)
        assert abs(s["shadows"] - (0.3 if k < 64 else 0.8)) < 1e-6


def box5(a):
    """5x5 box mean with edge padding, channel-wise."""
    p = np.pad(a.astype(np.float64), ((2, 2), (2, 2), (0, 0)), mode="edge")
    h, w = a.shape[:2]
    return sum(p[j:j + h, i:i + w] for j in range(5) for i in range(5)) / 25


def test_guided_filter_and_box_filter():
    rng = np.random.default_rng(40)
    g = np.zeros((20, 28, 4), np.float32)
    g[..., 0] = rng.uniform(0.05, 0.9, (20, 28))
    ab = oracle.guided_filter_ab(g, 0.01)
    y = g[..., :1]
    mean, sq = box5(y)[..., 0], box5(y * y)[..., 0]
    var = sq - mean * mean
    a = var / (var + 0.01)
    np.testing.assert_allclose(ab[..., 0], a, atol=2e-6)
    np.testing.assert_allclose(ab[..., 1], mean * (1 - a), atol=2e-6)
    two = rng.standard_normal((20, 28, 2)).astype(np.float32)
    np.testing.assert_allclose(oracle.box_filter_gf(two), box5(two), atol=2e-6)


@pytest.mark.parametrize("c", [0.12, 0.22, 0.5])
def test_ltm_mask_of_a_flat_image(c):
    """A flat image: zero variance, so a = 0 and b = the level, the illuminance
    is c in every band and the reflectance 1; each band's multiplier is its
    tone curve at c over c (mid and high bands: highlights 1)."""
    h, w = 32, 48
    img = np.zeros((h, w, 4), np.float32)
    img[..., 0] = c
    guides = [img[::16, ::16].copy(), img[::4, ::4].copy(), img]
    mask, _ = oracle.ltm_mask(img, np.zeros((h, w, 2), np.float32), guides, hist_shadows=0.7, hist_highlights=1.3)

    def curve(ill, shadows, highlights):
        mid = 0.22
        return mid * (ill / mid) ** shadows if ill <= mid else (1 - mid) * ((ill - mid) / (1 - mid)) ** highlights + mid
    want = curve(c, 0.7, 1.3) / c * (curve(c, 0.7, 1.0) / c) ** 2
    np.testing.assert_allclose(mask, want, rtol=2e-5)
