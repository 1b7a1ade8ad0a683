"""The registration oracle pinned by what the mathematics fixes (CPU only).

The reference ships no test vectors for ImageRegistration/.  The oracle's
pieces are pinned independently: the warps against scipy's linear
interpolation at the same FP32 sample positions and against exact copies and
shifts; the 9 x 9 Jacobi against numpy's eigh; the least-squares homography by
exact recovery of known homographies (general, affine, translation, minimal
4-point samples); RANSAC by recovering the planted inlier set.  The library's
host-side pieces (sample generator, ScaleHomography) are checked against the
oracle and the closed forms here too, without a device.
"""
import ctypes

import numpy as np
import pytest
from scipy import ndimage

import glasscamerapublic_amd as gcam
from glasscamerapublic_amd import _lib
import oracle

INVALID = 1  # GCAM_ERR_INVALID_ARGUMENT (include/gcam/denoise.h)


def project(H, p):
    ph = np.c_[p.astype(np.float64), np.ones(len(p))] @ np.asarray(H, np.float64).T
    return ph[:, :2] / ph[:, 2:]


def random_homography(rng, scale=1.0):
    H = np.eye(3) + scale * np.array([[0.02, 0.01, 0], [-0.01, 0.015, 0], [2e-5, -1e-5, 0]]) * rng.standard_normal((3, 3))
    H[:2, 2] = rng.uniform(-8, 8, 2)
    return (H / H[2, 2]).astype(np.float32)


def sample_positions(H, w, h, ow=None, oh=None):
    """The sampler's texel-space read positions of registerImage, in FP32 source order."""
    ow, oh = ow or w, oh or h
    H = np.asarray(H, np.float32)
    y, x = np.mgrid[0:oh, 0:ow].astype(np.float32)
    one = np.float32(1)
    u = (H[0, 0] * x + H[0, 1] * y) + H[0, 2] * one
    v = (H[1, 0] * x + H[1, 1] * y) + H[1, 2] * one
    ww = (H[2, 0] * x + H[2, 1] * y) + H[2, 2] * one
    nx, ny = np.float32(1) / np.float32(ow), np.float32(1) / np.float32(oh)
    tx = ((u / ww + np.float32(0.5)) * nx) * np.float32(w) - np.float32(0.5)
    ty = ((v / ww + np.float32(0.5)) * ny) * np.float32(h) - np.float32(0.5)
    return tx, ty


# ---- warps --------------------------------------------------------------------

@pytest.mark.parametrize("seed", [0, 1])
def test_register_image_is_linear_interpolation(seed):
    rng = np.random.default_rng(seed)
    h, w = 37, 53
    img = rng.random((h, w, 4), np.float32)
    H = random_homography(rng)
    got = oracle.register_image(img, H)
    tx, ty = sample_positions(H, w, h)
    for c in range(4):
        want = ndimage.map_coordinates(img[..., c].astype(np.float64), [ty, tx], order=1, mode="nearest")
        np.testing.assert_allclose(got[..., c], want, atol=2e-6)
    one = oracle.register_image(np.ascontiguousarray(img[..., 1]), H)
    np.testing.assert_array_equal(one, got[..., 1])


def test_register_image_identity_and_integer_shift_are_exact():
    rng = np.random.default_rng(2)
    h, w = 32, 64  # powers of two: the normalised coordinates round-trip exactly
    img = rng.random((h, w, 4), np.float32)
    np.testing.assert_array_equal(oracle.register_image(img, np.eye(3)), img)
    H = np.array([[1, 0, 3], [0, 1, -2], [0, 0, 1]], np.float32)
    got = oracle.register_image(img, H)
    yy = np.clip(np.arange(h) - 2, 0, h - 1)
    xx = np.clip(np.arange(w) + 3, 0, w - 1)
    np.testing.assert_array_equal(got, img[yy][:, xx])
    # far outside: the edge texel
    Hfar = np.array([[1, 0, 1e6], [0, 1, -1e6], [0, 0, 1]], np.float32)
    got = oracle.register_image(img, Hfar)
    np.testing.assert_array_equal(got, np.broadcast_to(img[0, w - 1], img.shape))


def test_register_and_fuse_closed_forms():
    rng = np.random.default_rng(4)
    h, w = 24, 40
    fused = rng.random((h, w, 4), np.float32)
    img = rng.random((h, w, 4), np.float32)
    H = random_homography(rng)
    warped = oracle.register_image(img, H)
    np.testing.assert_array_equal(oracle.register_and_fuse(fused, img, H, 1), warped)  # count 1: the new frame
    for count in (2, 3, 7):
        want = (np.float32(count - 1) * fused + warped) / np.float32(count)
        np.testing.assert_array_equal(oracle.register_and_fuse(fused, img, H, count), want)
    # the normalised coordinate uses the output size: an input of twice the size is read at 2x
    big = np.repeat(np.repeat(img, 2, 0), 2, 1)
    out = oracle.register_and_fuse(fused, big, np.eye(3), 1)
    tx, ty = sample_positions(np.eye(3), 2 * w, 2 * h, w, h)
    for c in range(4):
        want = ndimage.map_coordinates(big[..., c].astype(np.float64), [ty, tx], order=1, mode="nearest")
        np.testing.assert_allclose(out[..., c], want, atol=2e-6)


def test_register_bayer_moves_whole_quads():
    rng = np.random.default_rng(5)
    h, w = 20, 28
    raw = rng.integers(0, 65535, (h, w), dtype=np.uint16)
    np.testing.assert_array_equal(oracle.register_bayer(raw, np.eye(3)), raw)
    # +1 quad in x, -2 quads in y: a shift of (2, -4) pixels, colour planes kept
    H = np.array([[1, 0, 1], [0, 1, -2], [0, 0, 1]], np.float32)
    got = oracle.register_bayer(raw, H)
    y, x = np.mgrid[0:h, 0:w]
    want = raw[np.clip(y - 4, 0, h - 1), np.clip(x + 2, 0, w - 1)]
    np.testing.assert_array_equal(got, want)
    # half a quad rounds away from zero: 0.5 -> 1, so x/2 + 0.5 lands on the next quad
    H = np.array([[1, 0, 0.5], [0, 1, 0], [0, 0, 1]], np.float32)
    got = oracle.register_bayer(raw, H)
    want = raw[y, np.clip(2 * np.floor(x // 2 + 1.0).astype(int) + x % 2, 0, w - 1)]
    np.testing.assert_array_equal(got, want)


# ---- homography fit -----------------------------------------------------------

@pytest.mark.parametrize("seed", [0, 1, 2])
def test_jacobi9_matches_eigh(seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((9, 9))
    a = a @ a.T if seed != 2 else a + a.T
    if seed == 1:  # rank 7: the null space the homography fit reads
        u = rng.standard_normal((9, 7))
        a = u @ u.T
    w, v = oracle.jacobi9(a)
    ref = np.linalg.eigh(a)[0][::-1]
    assert np.all(np.diff(w) <= 0)
    np.testing.assert_allclose(w, ref, atol=1e-12 * np.abs(ref).max())
    np.testing.assert_allclose(v @ v.T, np.eye(9), atol=1e-12)
    np.testing.assert_allclose(v @ a @ v.T, np.diag(w), atol=1e-11 * np.abs(ref).max())


@pytest.mark.parametrize("kind", ["projective", "affine", "translation"])
def test_least_squares_recovers_exact_homographies(kind):
    rng = np.random.default_rng(7)
    if kind == "projective":
        H = random_homography(rng, 2.0)
    elif kind == "affine":
        H = random_homography(rng)
        H[2] = (0, 0, 1)
    else:
        H = np.array([[1, 0, 12.5], [0, 1, -7.25], [0, 0, 1]], np.float32)
    p = rng.uniform(0, 1000, (300, 2)).astype(np.float32)
    q = project(H, p).astype(np.float32)
    got = oracle.least_squares_homography(np.c_[p, q])
    np.testing.assert_allclose(got, H, rtol=2e-5, atol=2e-5 * np.abs(H).max())
    np.testing.assert_allclose(project(got, p), q, atol=2e-3)
    # minimal samples: the homography through 4 points
    got4 = oracle.least_squares_homography(np.c_[p[:4], q[:4]])
    np.testing.assert_allclose(project(got4, p[:4]), q[:4], atol=2e-3)


def test_least_squares_degenerate_points():
    p = np.array([[5, 1], [5, 2], [5, 7], [5, 9]], np.float32)  # no spread in x: range_error
    assert oracle.least_squares_homography(np.c_[p, p + 1]) is None


def make_matches(rng, H, n, outlier_frac, noise=0.05, extent=1000.0):
    p = rng.uniform(0, extent, (n, 2)).astype(np.float32)
    q = project(H, p) + noise * rng.standard_normal((n, 2))
    out = rng.random(n) < outlier_frac
    q[out] += rng.uniform(5, 50, (out.sum(), 2)) * rng.choice([-1, 1], (out.sum(), 2))
    return np.c_[p, q.astype(np.float32)].astype(np.float32), np.flatnonzero(~out)


@pytest.mark.parametrize("seed", [0, 1])
def test_ransac_recovers_the_inliers(seed):
    rng = np.random.default_rng(seed)
    H = random_homography(rng)
    m, truth = make_matches(rng, H, 400, 0.3)
    got, inl, best, loss = oracle.find_homography(m, 1.0, 500, seed)
    np.testing.assert_array_equal(inl, truth)
    assert loss == len(m) - len(truth)
    np.testing.assert_allclose(project(got, m[:, :2])[truth], m[truth, 2:], atol=0.25)
    # the refit is the least-squares fit of the inliers
    np.testing.assert_array_equal(got, oracle.least_squares_homography(m[inl]))


def test_ransac_clean_data_takes_the_first_hypothesis():
    rng = np.random.default_rng(9)
    H = random_homography(rng)
    m, truth = make_matches(rng, H, 64, 0.0, noise=0.0)
    _, inl, best, loss = oracle.find_homography(m, 1.0, 50, 3)
    assert (best, loss) == (0, 0) and len(inl) == 64


def test_ransac_samples_library_matches_oracle():
    for seed, n in ((0, 4), (1, 5), (2**63 + 11, 1500), (7, 100000)):
        a = gcam.ransacSamples(seed, n, 0, 3000)
        np.testing.assert_array_equal(a, oracle.ransac_samples(seed, n, 0, 3000))
        assert a.min() >= 0 and a.max() < n
        assert np.all(np.diff(a, axis=1) > 0)  # distinct and ascending
        np.testing.assert_array_equal(gcam.ransacSamples(seed, n, 1000, 5), a[1000:1005])
    # uniform over indices: chi-square of 4 x 20000 draws over 50 indices
    a = gcam.ransacSamples(42, 50, 0, 20000)
    counts = np.bincount(a.reshape(-1), minlength=50)
    expected = a.size / 50
    chi2 = ((counts - expected) ** 2 / expected).sum()
    assert chi2 < 100  # 49 degrees of freedom: p < 1e-4 above this


def test_scale_homography():
    rng = np.random.default_rng(11)
    H = random_homography(rng)
    for s in (2.0, 0.5, 3.0):
        S = np.diag([s, s, 1.0])
        want = S @ H.astype(np.float64) @ np.linalg.inv(S)
        np.testing.assert_allclose(gcam.ScaleHomography(H, s), want, rtol=1e-6)
        np.testing.assert_array_equal(gcam.ScaleHomography(H, s), oracle.scale_homography(H, s))
    # scale 2 is exact: rows 0-1 of column 2 double, row 2 of columns 0-1 halves
    want = H.copy()
    want[:2, 2] *= 2
    want[2, :2] /= 2
    np.testing.assert_array_equal(gcam.ScaleHomography(H, 2.0), want)


def test_registration_argument_checks_need_no_device():
    L = _lib.lib()
    H = (ctypes.c_float * 9)(1, 0, 0, 0, 1, 0, 0, 0, 1)
    fake = ctypes.c_void_p(0x10000)
    cnt = ctypes.c_int32()
    assert L.gcam_find_homography(fake, 3, 1.0, 10, 0, H, None, ctypes.byref(cnt), None) == INVALID
    assert L.gcam_find_homography(fake, 10, 1.0, 0, 0, H, None, ctypes.byref(cnt), None) == INVALID
    assert L.gcam_find_homography(ctypes.c_void_p(0x10004), 10, 1.0, 10, 0, H, None, ctypes.byref(cnt),
                                  None) == INVALID  # misaligned
    assert L.gcam_find_least_squares_homography(fake, 3, H, None) == INVALID
    rgba = _lib.Image(0x10000, 16, 16, 16, _lib.GCAM_PIXEL_RGBA32F)
    r16 = _lib.Image(0x20000, 16, 16, 16, _lib.GCAM_PIXEL_R16U)
    small = _lib.Image(0x30000, 8, 16, 8, _lib.GCAM_PIXEL_RGBA32F)
    st = L.gcam_register_and_fuse_image(ctypes.byref(rgba), ctypes.byref(rgba), ctypes.byref(small), H, 1, None)
    assert st == INVALID
    st = L.gcam_register_and_fuse_image(ctypes.byref(rgba), ctypes.byref(rgba), ctypes.byref(rgba), H, 0, None)
    assert st == INVALID
    nan = (ctypes.c_float * 9)(1, 0, float("nan"), 0, 1, 0, 0, 0, 1)
    st = L.gcam_register_image(ctypes.byref(rgba), ctypes.byref(small), nan, None)
    assert st == INVALID
    st = L.gcam_register_bayer_image(ctypes.byref(rgba), ctypes.byref(r16), H, None)
    assert st == INVALID
    assert b"R16U" in L.gcam_last_error()
