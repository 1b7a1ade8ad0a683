"""Pins the CPU oracle (CPU only).

The reference ships no tests, fixtures or golden vectors, and its own code
cannot run here (Metal and Eigen are absent), so each oracle function is
pinned by what the mathematics fixes: an independent formulation (hat-function
resampling, numpy.cov, LAPACK's eigh, im2col projection, windowed statistics),
closed forms and invariants chosen so that a dropped term, a wrong sign or
index, or a transposed operand fails.
"""
import numpy as np
import pytest

import oracle
from tests.helpers import default_params, nlf_for, noisy_ycc


def hat_matrix(n_out, n_in, positions):
    """Linear-filter weights with clamp-to-edge: value(s) = sum_i T[clamp(i)] max(0, 1 - |s - (i + 1/2)|)."""
    M = np.zeros((n_out, n_in))
    for o, plist in enumerate(positions):
        for s, wgt in plist:
            for i in range(int(np.floor(s)) - 2, int(np.ceil(s)) + 2):
                h = max(0.0, 1.0 - abs(s - (i + 0.5)))
                if h > 0:
                    M[o, min(max(i, 0), n_in - 1)] += wgt * h
    return M


def down_matrix(n_in, n_out):
    # downsampleImage*: samples at (o + 1/2 -+ 1/2) / n_out (normalized) -> texel units (o + k) n_in / n_out
    return hat_matrix(n_out, n_in, [[((o + k) * n_in / n_out, 0.5) for k in (0, 1)] for o in range(n_out)])


def up_matrix(n_in, n_out):
    # subtractNoiseImage: one sample at (o + 1/2) / n_out -> texel units (o + 1/2) n_in / n_out
    return hat_matrix(n_out, n_in, [[((o + 0.5) * n_in / n_out, 1.0)] for o in range(n_out)])


def rand_img(h, w, seed, ch=4, dtype=np.float64):
    return np.random.default_rng(seed).uniform(0.05, 0.95, (h, w, ch)).astype(dtype)


# ---------------- downsample (demosaic.metal:886-920) ----------------

@pytest.mark.parametrize("hin,win,hout,wout", [(16, 16, 8, 8), (12, 20, 6, 10), (17, 15, 8, 7), (20, 9, 7, 4)])
@pytest.mark.parametrize("ch", [4, 2])
def test_downsample_matches_hat_filter(hin, win, hout, wout, ch):
    img = rand_img(hin, win, 1, ch)
    out = oracle.downsample(img, wout, hout)
    My, Mx = down_matrix(hin, hout), down_matrix(win, wout)
    nc = 3 if ch == 4 else 2
    for c in range(nc):
        np.testing.assert_allclose(out[..., c], My @ img[..., c] @ Mx.T, rtol=0, atol=1e-13)
    if ch == 4:
        assert np.all(out[..., 3] == 0)


def test_downsample_2x_is_4x4_box():
    img = rand_img(16, 16, 2)
    out = oracle.downsample(img)
    pad = np.pad(img, ((1, 1), (1, 1), (0, 0)), mode="edge")
    for oy in range(8):
        for ox in range(8):
            box = pad[2 * oy:2 * oy + 4, 2 * ox:2 * ox + 4, :3].mean(axis=(0, 1))
            np.testing.assert_allclose(out[oy, ox, :3], box, atol=1e-14)


def test_downsample_constant_and_f32():
    img = np.full((10, 14, 4), 0.37)
    np.testing.assert_allclose(oracle.downsample(img)[..., :3], 0.37, atol=1e-15)
    f = rand_img(32, 32, 3, dtype=np.float32)
    np.testing.assert_allclose(oracle.downsample(f), oracle.downsample(f.astype(np.float64)), atol=1e-6)


# ---------------- subtractNoiseImage (demosaic.metal:922-959) ----------------

@pytest.mark.parametrize("h,w,h1,w1", [(16, 16, 8, 8), (17, 13, 8, 6), (10, 12, 4, 7)])
def test_subtract_upsample_matches_hat_filter(h, w, h1, w1):
    img, img1, den1 = rand_img(h, w, 4), rand_img(h1, w1, 5), rand_img(h1, w1, 6)
    out = oracle.subtract_noise(img, img1, den1, luma_weight=0.7, sharpening=1.0)
    My, Mx = up_matrix(h1, h), up_matrix(w1, w)
    lw = [0.7, 1.0, 1.0]
    for c in range(3):
        up1 = My @ img1[..., c] @ Mx.T
        upd = My @ den1[..., c] @ Mx.T
        ref = img[..., c] - lw[c] * (up1 - upd)
        if c == 0:
            ref = np.maximum(ref, 0.0)
        np.testing.assert_allclose(out[..., c], ref, atol=1e-13)
    np.testing.assert_array_equal(out[..., 3], img[..., 3])


def test_subtract_2x_weights_quarter_three_quarters():
    img1 = np.zeros((4, 4, 4))
    img1[1, 1, 0] = 1.0  # unit impulse in the coarse noisy image, denoised = 0
    out = oracle.subtract_noise(np.zeros((8, 8, 4)) + 1.0, img1, np.zeros((4, 4, 4)))
    # Y = 1 - up(impulse); output rows/cols 1..4 see the impulse with 1-D weights 1/4, 3/4, 3/4, 1/4
    wts = np.array([0.25, 0.75, 0.75, 0.25])
    np.testing.assert_allclose(1.0 - out[1:5, 1:5, 0], np.outer(wts, wts), atol=1e-15)


def test_subtract_identity_when_denoised_equals_noisy():
    img, img1 = rand_img(12, 12, 7), rand_img(6, 6, 8)
    out = oracle.subtract_noise(img, img1, img1, sharpening=1.0)
    np.testing.assert_allclose(out, img, atol=1e-15)


def test_subtract_sharpening_limits():
    img, den1 = rand_img(8, 8, 9), np.full((4, 4, 4), 0.4)
    img1 = den1 + np.random.default_rng(10).uniform(-0.02, 0.02, (4, 4, 4))
    base = oracle.subtract_noise(img, img1, den1, sharpening=1.0)
    # zero gradient -> detail = 0 -> alpha = 1 whatever the sharpening
    np.testing.assert_allclose(oracle.subtract_noise(img, img1, den1, sharpening=2.0, nlf=(1e-4, 1e-4)), base,
                               atol=1e-15)
    # strong gradient, mid-tone denoised luma -> detail = 1 -> alpha = sharpening: mix(d1, d, s)
    grad = np.full((8, 8, 2), 10.0)
    sh = oracle.subtract_noise(img, img1, den1, grad=grad, sharpening=1.5, nlf=(1e-4, 1e-4))
    raw = oracle.subtract_noise(img, img1, den1, sharpening=1.0)  # = the unsharpened reconstruct (Y > 0 here)
    assert raw[..., 0].min() > 0
    expect = 0.4 + 1.5 * (raw[..., :3] - 0.4)
    expect[..., 0] = np.maximum(expect[..., 0], 0)
    np.testing.assert_allclose(sh[..., :3], expect, atol=1e-13)


# ---------------- collectPatches (demosaic.metal:755-772) ----------------

def test_collect_patches_content_and_slots_dense():
    img = rand_img(24, 32, 11)
    p = oracle.collect_patches(img)
    assert p.shape == (24 * 32 // 64, 25)
    pad = np.pad(img[..., 0], 2, mode="edge")
    for y in range(3):
        for x in range(4):
            np.testing.assert_array_equal(p[y * 4 + x], pad[8 * y:8 * y + 5, 8 * x:8 * x + 5].ravel())


def test_collect_patches_slot_table_width_252():
    # 252 x 189 (the 12 MP level-4 size): grid 31 x 23 patches, n = 252*189//64 = 744 slots.
    # Hand-derived: row y starts at slot floor(31.5 y); slot 62 is skipped, slot 63 is patch (0, 2),
    # the last written slot is floor(31.5*22) + 30 = 723 and 744 - 713 = 31 slots stay zero.
    img = rand_img(189, 252, 12) + 1.0
    p = oracle.collect_patches(img)
    assert p.shape == (744, 25)
    written = np.flatnonzero(np.any(p != 0, axis=1))
    assert written.size == 31 * 23
    assert written.max() == 723
    assert 62 not in written and 61 in written and 63 in written
    pad = np.pad(img[..., 0], 2, mode="edge")
    np.testing.assert_array_equal(p[63], pad[16:21, 0:5].ravel())
    np.testing.assert_array_equal(p[723], pad[176:181, 240:245].ravel())


# ---------------- build_pca_space (PCA.cpp:23-56) ----------------

def test_covariance_matches_numpy_cov():
    pts = np.random.default_rng(13).normal(size=(500, 25)) @ np.random.default_rng(14).normal(size=(25, 25))
    np.testing.assert_allclose(oracle.patch_covariance(pts), np.cov(pts, rowvar=False), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(oracle.patch_covariance(pts.astype(np.float32)),
                               np.cov(pts.astype(np.float32).astype(np.float64), rowvar=False), rtol=1e-11,
                               atol=1e-11)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_jacobi_matches_lapack(seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(25, 25))
    C = X @ X.T + np.diag(np.linspace(0, 3, 25))
    vals, vecs = oracle.eigh_jacobi(C)
    lv, lV = np.linalg.eigh(C)
    np.testing.assert_allclose(vals, lv[::-1], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(vecs.T @ vecs, np.eye(25), atol=1e-12)
    np.testing.assert_allclose(vecs.T @ C @ vecs, np.diag(vals), atol=1e-10)
    for j in range(25):  # same eigenvectors up to sign
        assert abs(abs(vecs[:, j] @ lV[:, ::-1][:, j]) - 1.0) < 1e-10


def test_pca_space_known_answers():
    # diagonal covariance with distinct variances -> unit vectors in decreasing-variance order
    rng = np.random.default_rng(15)
    scales = np.linspace(3.0, 0.5, 25)[rng.permutation(25)]
    pts = rng.normal(size=(20000, 25))
    pts = (pts - pts.mean(0)) / pts.std(0)
    Q, _ = np.linalg.qr(pts)
    pts = Q * np.sqrt(pts.shape[0] - 1) * scales  # exactly decorrelated columns with the given scales
    basis, ev = oracle.build_pca_space(pts)
    order = np.argsort(-scales)[:8]
    for c in range(8):
        e = np.zeros(25)
        e[order[c]] = 1.0
        np.testing.assert_allclose(basis[:, c], e, atol=1e-9)
    np.testing.assert_allclose(ev[:8], scales[order] ** 2, rtol=1e-9)
    # rank-1 covariance: the leading vector is the generator, the rest have zero eigenvalue
    v = rng.normal(size=25)
    v /= np.linalg.norm(v)
    basis, ev = oracle.build_pca_space(np.outer(rng.normal(size=300), v))
    assert abs(abs(basis[:, 0] @ v) - 1.0) < 1e-12
    assert np.all(np.abs(ev[1:]) < 1e-12 * ev[0])


def test_pca_space_rank_deficient_and_sign():
    # 16 patches (the 512x512 level-4 case): rank <= 15, leading subspace still matches LAPACK
    rng = np.random.default_rng(16)
    pts = rng.normal(size=(16, 25)) @ rng.normal(size=(25, 25))
    basis, ev = oracle.build_pca_space(pts)
    lv, lV = np.linalg.eigh(np.cov(pts, rowvar=False))
    P = lV[:, ::-1][:, :8]
    np.testing.assert_allclose(basis @ basis.T, P @ P.T, atol=1e-10)
    np.testing.assert_allclose(ev[:8], lv[::-1][:8], rtol=1e-10)
    assert np.all(np.abs(ev[15:]) < 1e-10 * ev[0])
    # sign convention: non-negative dot with the fixed reference vector
    ref = np.array([1.0 + ((r * r * 7 + r * 3 + 5) % 23) / 16.0 for r in range(25)])
    assert np.all(basis.T @ ref >= 0)


# ---------------- pcaProjection (demosaic.metal:774-796) ----------------

def test_projection_matches_im2col():
    img = rand_img(20, 28, 17)
    basis = np.random.default_rng(18).normal(size=(25, 8))
    feat = oracle.pca_project(img, basis)
    pad = np.pad(img[..., 0], 2, mode="edge")
    win = np.lib.stride_tricks.sliding_window_view(pad, (5, 5))  # (20, 28, 5, 5), [j][i] row-major
    ref = win.reshape(20, 28, 25) @ basis
    np.testing.assert_allclose(feat, ref, atol=1e-13)
    # a transposed patch layout would differ on this anisotropic basis
    ref_t = np.transpose(win, (0, 1, 3, 2)).reshape(20, 28, 25) @ basis
    assert np.abs(feat - ref_t).max() > 1e-3


# ---------------- blockMatchingDenoiseImage (demosaic.metal:806-884) ----------------

def _bm_inputs(h, w, seed):
    img = rand_img(h, w, seed)
    img[..., 1:3] -= 0.5
    basis, _ = oracle.build_pca_space(oracle.collect_patches(np.pad(img, ((0, 8), (0, 8), (0, 0)), mode="edge")))
    return img, oracle.pca_project(img, basis)


def test_bm_constant_image_is_fixed_point():
    img = np.zeros((12, 12, 4))
    img[..., :3] = [0.4, -0.05, 0.1]
    feat = oracle.pca_project(img, np.random.default_rng(19).normal(size=(25, 8)))
    out = oracle.block_match(img, feat, [1e-5] * 3, [1e-4] * 3, [1, 1, 1], radius=3)
    np.testing.assert_allclose(out[..., :3], img[..., :3], atol=1e-15)
    wd = np.exp(-2 * (0.1 * np.sqrt(np.add.outer(np.arange(-3, 4) ** 2, np.arange(-3, 4) ** 2))) ** 2).sum()
    np.testing.assert_allclose(out[..., 3], wd, rtol=1e-13)


def test_bm_tiny_threshold_returns_centre():
    img, feat = _bm_inputs(10, 10, 20)
    out = oracle.block_match(img, feat, [1e-5] * 3, [1e-4] * 3, [1e-9] * 3, radius=2)
    np.testing.assert_allclose(out[..., :3], img[..., :3], atol=1e-12)


def test_bm_huge_threshold_is_direction_weighted_mean():
    img, feat = _bm_inputs(11, 9, 21)
    r = 3
    out = oracle.block_match(img, feat, [1e-5] * 3, [1e-4] * 3, [1e12] * 3, radius=r)
    pad = np.pad(img, ((r, r), (r, r), (0, 0)), mode="edge")
    off = np.arange(-r, r + 1)
    wd = np.exp(-0.02 * np.add.outer(off ** 2, off ** 2))
    ref = np.zeros((11, 9, 3))
    for dy in range(2 * r + 1):
        for dx in range(2 * r + 1):
            ref += wd[dy, dx] * pad[dy:dy + 11, dx:dx + 9, :3]
    ref /= wd.sum()
    np.testing.assert_allclose(out[..., :3], ref, atol=1e-12)


def test_bm_invariant_to_orthogonal_feature_rotation():
    img, feat = _bm_inputs(10, 12, 22)
    Q, _ = np.linalg.qr(np.random.default_rng(23).normal(size=(8, 8)))
    args = ([2e-4, 1e-4, 1e-4], [4e-3, 1e-3, 1e-3], [1.0, 1.2, 1.2])
    a = oracle.block_match(img, feat, *args, chroma_boost=2.0, radius=2)
    b = oracle.block_match(img, feat @ Q, *args, chroma_boost=2.0, radius=2)
    np.testing.assert_allclose(a, b, atol=1e-12)


def test_bm_brute_force_small():
    img, feat = _bm_inputs(9, 8, 24)
    va, vb, th = np.array([2e-4, 1e-4, 1e-4]), np.array([4e-3, 1e-3, 2e-3]), np.array([1.0, 1.3, 0.8])
    cB, gB, gT, ls, r = 1.7, 0.6, 2.0, 0.8, 2
    grad = np.random.default_rng(25).normal(0, 0.05, (9, 8, 2))
    out = oracle.block_match(img, feat, va, vb, th, cB, gB, gT, ls, r, grad=grad)
    h, w = 9, 8
    cy, cx = h // 2, w // 2
    for py in range(h):
        for px in range(w):
            d = np.hypot(px - cx, py - cy) / np.hypot(cx, cy)
            lens = 1 + ls * d * d
            sig = np.sqrt(va + vb * lens * img[py, px, 0])
            dm = 1 / (th * sig)
            edge_t = np.clip((gT * np.hypot(*grad[py, px]) / sig[0] - 2) / 14, 0, 1)
            edge = edge_t * edge_t * (3 - 2 * edge_t)
            num, den = np.zeros(3), np.zeros(3)
            for y in range(-r, r + 1):
                for x in range(-r, r + 1):
                    qy, qx = min(max(py + y, 0), h - 1), min(max(px + x, 0), w - 1)
                    dist = np.linalg.norm(feat[qy, qx] - feat[py, px]) * dm[0]
                    wl = np.exp(-2 * (dist / (1 + gB * edge)) ** 2)
                    cd = (img[qy, qx, 1:3] - img[py, px, 1:3]) * dm[1:] / cB
                    wc = np.exp(-2 * (wl ** 2 + cd @ cd))
                    wdir = np.exp(-2 * (0.1 * np.hypot(x, y)) ** 2)
                    sw = wdir * np.array([wl, wc, wc])
                    num += sw * img[qy, qx, :3]
                    den += sw
            np.testing.assert_allclose(out[py, px, :3], num / den, atol=1e-12)
            np.testing.assert_allclose(out[py, px, 3], den[0], rtol=1e-12)


# ---------------- basicNoiseStatistics (demosaic.metal:1138-1159) ----------------

def test_noise_stats_matches_windowed_moments():
    img = rand_img(13, 17, 26)
    st = oracle.noise_stats(img)
    pad = np.pad(img[..., :3], ((2, 2), (2, 2), (0, 0)), mode="edge")
    win = np.lib.stride_tricks.sliding_window_view(pad, (5, 5), axis=(0, 1))  # (13, 17, 3, 5, 5)
    np.testing.assert_allclose(st[..., 0], win[:, :, 0].mean(axis=(-1, -2)), atol=1e-14)
    np.testing.assert_allclose(st[..., 1:], win.var(axis=(-1, -2)), atol=1e-14)


# ---------------- MeasureYCbCrNLF fit (pyramid_processor.cpp:174-324) ----------------

def test_nlf_fit_recovers_exact_linear_model():
    rng = np.random.default_rng(27)
    m = rng.uniform(0.01, 0.85, (64, 64))
    A, B = np.array([2e-6, 1e-6, 3e-6]), np.array([3e-4, 1e-4, 2e-4])
    st = np.zeros((64, 64, 4), np.float32)
    st[..., 0] = m
    st[..., 1:] = A + B * m[..., None]
    fit = oracle.fit_ycbcr_nlf(st)
    np.testing.assert_allclose(fit.A, A, rtol=2e-4)
    np.testing.assert_allclose(fit.B, B, rtol=2e-5)
    fit2 = oracle.fit_ycbcr_nlf(st, exposure_multiplier=2.0)
    np.testing.assert_allclose(fit2.A, 4 * fit.A, rtol=1e-12)
    np.testing.assert_allclose(fit2.B, 4 * fit.B, rtol=1e-12)


def test_nlf_histogram_bins_and_variance_max():
    # log10 bins with truncation toward zero: (.., 1e-6] -> 0, (1e-6, 1e-5] -> 1, ..., (1e-2, ..) -> 5
    vals = [5e-7, 2e-6, 3e-5, 4e-4, 5e-3, 0.5]
    st = np.zeros((len(vals), 1, 4), np.float32)
    st[:, 0, 0] = 0.5
    st[:, 0, 1:] = np.array(vals)[:, None]
    fit = oracle.fit_ycbcr_nlf(st)
    np.testing.assert_array_equal(fit.histogram, [1, 1, 1, 1, 1, 1])
    assert fit.variance_max == 1.0  # no bin above 1e3 counts
    st = np.zeros((40, 40, 4), np.float32)  # 1600 pixels in bin 3 -> varianceMax 1e-2
    st[..., 0] = np.linspace(0.1, 0.8, 1600).reshape(40, 40)
    st[..., 1:] = 5e-4
    assert oracle.fit_ycbcr_nlf(st).variance_max == pytest.approx(1e-2)


def test_nlf_fit_fallback_branch_keeps_first_pass():
    # half the pixels on the line, half far off it: the filtered pass covers the on-line half only,
    # but a coverage threshold of 1% is still met; make the second pass worse by construction instead:
    # every residual equal in magnitude -> diff^2 == err2 for all, all kept, newErr2 == err2 (accepted).
    # A NaN mean excludes pixels; an image whose valid pixels are < 1% forces the fallback.
    st = np.full((100, 100, 4), np.nan, np.float32)
    rng = np.random.default_rng(28)
    m = rng.uniform(0.1, 0.8, 50)
    st[0, :50, 0] = m
    st[0, :50, 1:] = (1e-5 + 2e-4 * m)[:, None] * rng.uniform(0.5, 1.5, (50, 3))
    fit = oracle.fit_ycbcr_nlf(st)
    assert not fit.accepted
    assert fit.n_first == 50
    # first-pass least squares (by numpy) is what comes back
    for c in range(3):
        Bc, Ac = np.polyfit(m.astype(np.float32).astype(np.float64), st[0, :50, 1 + c].astype(np.float64), 1)
        assert fit.B[c] == pytest.approx(max(Bc, 1e-8), rel=1e-9)
        assert fit.A[c] == pytest.approx(max(Ac, 1e-8), rel=1e-6)


# ---------------- whole pyramid (pyramid_processor.cpp:96-172) ----------------

def test_pyramid_constant_image_is_fixed_point():
    img = np.zeros((64, 48, 4), np.float32)
    img[..., :3] = [0.3, 0.02, -0.04]
    res = oracle.pyramid_denoise(img, default_params(3), nlf_for(3), levels=3, radius=2)
    for lvl in res.denoised:
        np.testing.assert_allclose(lvl[..., :3], np.broadcast_to(img[0, 0, :3], lvl[..., :3].shape), atol=1e-6)


def test_pyramid_f32_tracks_f64_and_denoises():
    # every level samples >= 12 patches, so each top-8 PCA subspace is well defined
    noisy, clean = noisy_ycc(256, 192, seed=29)
    prm = default_params(4, luma=8.0, chroma=2.0)
    r32 = oracle.pyramid_denoise(noisy, prm, nlf_for(4), levels=4, radius=3)
    r64 = oracle.pyramid_denoise(noisy, prm, nlf_for(4), levels=4, radius=3, dtype=np.float64)
    assert np.abs(r32.denoised[0][..., :3] - r64.denoised[0][..., :3]).max() < 1e-4
    err_in = np.sqrt(((noisy[..., 0] - clean[..., 0]) ** 2).mean())
    err_out = np.sqrt(((r64.denoised[0][..., 0] - clean[..., 0]) ** 2).mean())
    assert err_out < 0.85 * err_in


def test_pyramid_calibrate_fits_each_level():
    noisy, _ = noisy_ycc(128, 128, seed=30)
    res = oracle.pyramid_denoise(noisy, default_params(3), np.zeros((3, 2, 3)), levels=3, radius=2, calibrate=True)
    assert np.all(res.nlf >= 1e-8)
    st = oracle.noise_stats(noisy)
    np.testing.assert_allclose(res.nlf[0, 0], oracle.fit_ycbcr_nlf(st).A, rtol=1e-6)
    np.testing.assert_allclose(res.nlf[0, 1], oracle.fit_ycbcr_nlf(st).B, rtol=1e-6)


def test_pyramid_golden_fixture():
    g = np.load("tests/golden/pyramid_64x64_l2_r3.npz")
    res = oracle.pyramid_denoise(g["input"], default_params(2, luma=3.0), g["nlf"], levels=2, radius=3)
    for i in range(2):
        np.testing.assert_allclose(res.denoised[i], g[f"denoised{i}"], atol=2e-6)
        np.testing.assert_allclose(res.bases[i] @ res.bases[i].T, g[f"projector{i}"], atol=1e-9)


# ---------------- denoiseImage (demosaic.metal:664-719) ----------------

def test_denoise_image_limits():
    img = rand_img(12, 14, 40)
    # constant image: fixed point; magnitude channel = 0 without a gradient
    c = np.zeros((9, 9, 4))
    c[..., :3] = [0.3, 0.1, -0.1]
    out = oracle.denoise_image(c, [1e-4] * 3, [1e-3] * 3, [1, 1, 1])
    np.testing.assert_allclose(out[..., :3], c[..., :3], atol=1e-15)
    assert np.all(out[..., 3] == 0)
    # huge thresholds: every sample passes both steps -> plain 5x5 (or 9x9 with gradientBoost) box mean
    for gb, size in ((0.0, 2), (1.0, 4)):
        out = oracle.denoise_image(img, [1e-4] * 3, [1e-3] * 3, [1e9] * 3, gradient_boost=gb)
        pad = np.pad(img[..., :3], ((size, size), (size, size), (0, 0)), mode="edge")
        win = np.lib.stride_tricks.sliding_window_view(pad, (2 * size + 1, 2 * size + 1), axis=(0, 1))
        np.testing.assert_allclose(out[..., :3], win.mean(axis=(-1, -2)), atol=1e-12)
    # tiny thresholds: only the centre passes
    out = oracle.denoise_image(img, [1e-4] * 3, [1e-3] * 3, [1e-9] * 3)
    np.testing.assert_allclose(out[..., :3], img[..., :3], atol=1e-15)


def test_denoise_image_brute_force_with_gradient():
    img = rand_img(8, 9, 41)
    grad = np.random.default_rng(42).normal(0, 0.05, (8, 9, 2))
    va, vb, th = np.array([1e-4, 2e-4, 2e-4]), np.array([1e-2, 5e-3, 5e-3]), np.array([1.5, 1.0, 1.2])
    cB, gB, gT = 1.3, 0.7, 30.0
    out = oracle.denoise_image(img, va, vb, th, cB, gB, gT, grad=grad)
    h, w = 8, 9

    def ss(e0, e1, x):
        t = np.clip((x - e0) / (e1 - e0), 0, 1)
        return t * t * (3 - 2 * t)
    for py in range(h):
        for px in range(w):
            sig = np.sqrt(va + vb * img[py, px, 0])
            dm = 1 / (th * sig)
            g = grad[py, px]
            ang, mag = np.arctan2(g[1], g[0]), np.hypot(*g)
            edge = ss(4, 16, gT * mag / sig[0])
            num, den = np.zeros(3), np.zeros(3)
            for y in range(-4, 5):
                for x in range(-4, 5):
                    qy, qx = min(max(py + y, 0), h - 1), min(max(px + x, 0), w - 1)
                    d = (img[qy, qx, :3] - img[py, px, :3]) * dm
                    gd = (grad[qy, qx] - g) / sig[0]
                    a = x * np.cos(ang) + y * np.sin(ang)
                    dw = 1 + (np.exp(-a * a / 0.25) - 1) * edge
                    gw = 1 - ss(2, 8, np.hypot(*gd))
                    lw = 1.0 if abs(d[0]) < 1 + gB * edge else 0.0
                    cw = 1.0 if np.linalg.norm(d) < cB else 0.0
                    sw = np.array([dw * gw * lw, cw, cw])
                    num += sw * img[qy, qx, :3]
                    den += sw
            np.testing.assert_allclose(out[py, px, :3], num / den, atol=1e-12)
            assert out[py, px, 3] == pytest.approx(mag, rel=1e-12)


def test_pyramid_bilateral_mode_runs_denoise_image():
    noisy, _ = noisy_ycc(128, 96, seed=43)
    res = oracle.pyramid_denoise(noisy, default_params(3), nlf_for(3), levels=3, patch_similarity=False)
    # the coarsest level is denoiseImage of the downsampled input
    lvl2_in = oracle.downsample(oracle.downsample(noisy))
    ref = oracle.denoise_image(lvl2_in, nlf_for(3)[2, 0], nlf_for(3)[2, 1], [1, 1, 1])
    np.testing.assert_allclose(res.denoised[2], ref, atol=1e-6)


# ---------------- differentiation form (FP64 adjoint) ----------------

def test_adjoint_matches_central_finite_differences(oracle_threads):
    """dL/dthr of the FP64 adjoint (eigenvector derivative included) against
    central differences (eps = 1e-3) of the FP64 forward."""
    noisy, clean = noisy_ycc(256, 256, seed=5, a=5e-3, b=2e-3, dtype=np.float64)
    levels, radius = 4, 2
    thr = np.array([[4.0, 2.0, 2.5], [3.0, 2.0, 2.0], [2.5, 1.5, 2.0], [2.0, 1.5, 1.5]])
    cb = np.array([2.0, 1.5, 1.0, 1.0])
    nlf = nlf_for(levels, 5e-3, 2e-3)
    loss, g, den = oracle.pyramid_loss_grad(noisy, clean, thr, cb, nlf, levels=levels, radius=radius)
    eps = 1e-3
    fd = np.zeros_like(g)
    for i in range(levels):
        for c in range(3):
            tp, tm = thr.copy(), thr.copy()
            tp[i, c] += eps
            tm[i, c] -= eps
            lp = oracle.pyramid_loss_grad(noisy, clean, tp, cb, nlf, levels=levels, radius=radius, want_grad=False)[0]
            lm = oracle.pyramid_loss_grad(noisy, clean, tm, cb, nlf, levels=levels, radius=radius, want_grad=False)[0]
            fd[i, c] = (lp - lm) / (2 * eps)
    scale = np.abs(fd).max()
    np.testing.assert_allclose(g, fd, rtol=1e-4, atol=1e-6 * scale)
    # the forward of the differentiable form is the ordinary pyramid when Cb and Cr thresholds agree
    thr2 = thr.copy()
    thr2[:, 2] = thr2[:, 1]
    _, _, den2 = oracle.pyramid_loss_grad(noisy, clean, thr2, cb, nlf, levels=levels, radius=radius, want_grad=False)
    prm = [dict(luma=thr2[i, 0], chroma=thr2[i, 1], chromaBoost=cb[i], gradientBoost=0.0, gradientThreshold=1.0,
                sharpening=1.0) for i in range(levels)]
    ref = oracle.pyramid_denoise(noisy, prm, nlf, levels=levels, radius=radius, dtype=np.float64)
    np.testing.assert_allclose(den2, ref.denoised[0], atol=1e-12)
