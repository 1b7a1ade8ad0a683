"""Burst registration and fusion on the GPU against the oracle.

The warps evaluate the oracle's FP32 operations in the same order (no
contraction, correctly rounded division), so registerAndFuse, registerImage
and registerBayerImage are compared bit for bit.  The homography fits compute
in FP64 by another route than the oracle (closed-form minimal models, a
parallel-ordered Jacobi), and both round the result to FP32: the least-squares
fit is held to 2e-6 relative to the largest entry, and FindHomography must
pick the same inlier set and agree to 1e-5 of the largest entry.  The test
scenes' inlier noise (0.05 px) and outlier offsets (5 px or more) keep the
winning models' squared errors far from the threshold (1 px^2), so the last
bits of the two minimal models cannot change the inlier set.
"""
import numpy as np
import pytest
import torch

import glasscamerapublic_amd as gcam
import oracle
from tests.test_registration_oracle import make_matches, project, random_homography

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def scene(h, w, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    img = np.zeros((h, w, 4), np.float32)
    for c in range(4):
        fx, fy = rng.uniform(0.01, 0.08, 2)
        img[..., c] = 0.5 + 0.3 * np.sin(fx * x + 1.3 * c) * np.cos(fy * y) + 0.05 * rng.random((h, w))
    return img


@pytest.mark.parametrize("w,h,count", [(256, 192, 1), (250, 130, 2), (66, 18, 4)])
def test_register_and_fuse_is_bit_exact(w, h, count):
    print(f"registerAndFuse {w}x{h} count {count}", flush=True)
    rng = np.random.default_rng(w + count)
    fused, img = scene(h, w, 1), scene(h, w, 2)
    H = random_homography(rng)
    want = oracle.register_and_fuse(fused, img, H, count)
    got = gcam.registerAndFuse(dev(fused), dev(img), None, H, count).cpu().numpy()
    np.testing.assert_array_equal(got, want)
    # in place, as the burst driver calls it
    f = dev(fused)
    gcam.registerAndFuse(f, dev(img), f, H, count)
    np.testing.assert_array_equal(f.cpu().numpy(), want)
    # an input of another size (the sampler's normalised coordinate)
    big = scene(2 * h + 3, 2 * w - 5, 3)
    want = oracle.register_and_fuse(fused, big, H, count)
    np.testing.assert_array_equal(gcam.registerAndFuse(dev(fused), dev(big), None, H, count).cpu().numpy(), want)


@pytest.mark.parametrize("channels", [1, 4])
def test_register_image_is_bit_exact(channels):
    print(f"registerImage channels {channels}", flush=True)
    rng = np.random.default_rng(channels)
    h, w = 173, 301
    img = scene(h, w, 5)
    if channels == 1:
        img = np.ascontiguousarray(img[..., 2])
    for H in (random_homography(rng), np.eye(3, dtype=np.float32),
              np.array([[1, 0, 1e6], [0, 1, -3.5], [0, 0, 1]], np.float32)):
        want = oracle.register_image(img, H)
        np.testing.assert_array_equal(gcam.registerImage(dev(img), H).cpu().numpy(), want)


def test_register_bayer_is_bit_exact():
    print("registerBayerImage", flush=True)
    rng = np.random.default_rng(8)
    h, w = 130, 250
    raw = rng.integers(0, 65536, (h, w)).astype(np.uint16)
    for H in (random_homography(rng), np.array([[1, 0, 0.5], [0, 1, -2.5], [0, 0, 1]], np.float32),
              np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1e-30]], np.float32)):
        want = oracle.register_bayer(raw, H)
        got = gcam.registerBayerImage(dev(raw.view(np.int16)), H).cpu().numpy().view(np.uint16)
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("n", [4, 37, 1500, 20000])
def test_least_squares_homography_matches_oracle(n):
    print(f"FindLeastSquaresHomography n={n}", flush=True)
    rng = np.random.default_rng(n)
    H = random_homography(rng, 2.0)
    p = rng.uniform(0, 2000, (n, 2)).astype(np.float32)
    q = (project(H, p) + 0.1 * rng.standard_normal((n, 2))).astype(np.float32)
    m = np.c_[p, q]
    want = oracle.least_squares_homography(m)
    got = gcam.FindLeastSquaresHomography(m)
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-6 * np.abs(want).max())


def test_least_squares_degenerate_raises():
    p = np.array([[5, 1], [5, 2], [5, 7], [5, 9], [5, 3]], np.float32)
    with pytest.raises(gcam.GcamError, match="homography"):
        gcam.FindLeastSquaresHomography(np.c_[p, p + 1])


@pytest.mark.parametrize("n,outliers,iters,seed", [(1500, 0.3, 2000, 0), (400, 0.6, 2000, 1), (64, 0.0, 50, 2),
                                                   (5000, 0.45, 1000, 3)])
def test_find_homography_matches_oracle(n, outliers, iters, seed):
    print(f"FindHomography n={n} outliers={outliers} iters={iters}", flush=True)
    rng = np.random.default_rng(seed)
    H = random_homography(rng)
    m, truth = make_matches(rng, H, n, outliers, noise=0.05)
    want, want_inl, best, loss = oracle.find_homography(m, 1.0, iters, seed)
    inl = []
    got = gcam.FindHomography(m, 1.0, iters, inl, seed=seed)
    np.testing.assert_array_equal(np.array(inl), want_inl)
    np.testing.assert_array_equal(want_inl, truth)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-5 * np.abs(want).max())
    again = gcam.FindHomography(torch.from_numpy(m).cuda(), 1.0, iters, None, seed=seed)
    np.testing.assert_array_equal(again, got)  # deterministic


def test_find_homography_small_sets():
    # 6 exact matches: every hypothesis fits all of them; fewer than 4 matches is an error
    rng = np.random.default_rng(4)
    H = random_homography(rng)
    p = rng.uniform(0, 100, (6, 2)).astype(np.float32)
    q = project(H, p).astype(np.float32)
    m = np.c_[p, q]
    want, want_inl, _, _ = oracle.find_homography(m, 1.0, 30, 0)
    inl = []
    got = gcam.FindHomography(m, 1.0, 30, inl, seed=0)
    np.testing.assert_array_equal(inl, want_inl)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-5 * np.abs(want).max())
    with pytest.raises(gcam.GcamError):
        gcam.FindHomography(m[:3], 1.0, 30)


def test_fuse_burst_matches_the_oracle_chain():
    """pyramidFusion.cpp:316-346 on 3 synthetic frames: per-channel matches at half
    resolution, averaged homographies, scaled by 2, fused in place."""
    print("fuseBurst", flush=True)
    rng = np.random.default_rng(12)
    h, w = 192, 256
    ref = scene(h, w, 20)
    frames, matches, Hs = [], [], []
    for i in range(3):
        Hh = random_homography(rng, 0.3)  # half-resolution motion
        Hs.append(Hh)
        frames.append(oracle.register_image(ref, oracle.scale_homography(Hh, 2.0)))
        per_channel = []
        for c in range(2):
            m, _ = make_matches(rng, Hh, 300, 0.2, noise=0.03, extent=min(w, h) / 2)
            per_channel.append(m)
        matches.append(per_channel)
    got = gcam.fuseBurst(dev(ref), [dev(f) for f in frames], matches, seed=5).cpu().numpy()
    fused = ref.copy()
    for i in range(3):
        H = np.zeros((3, 3), np.float32)
        for c in range(2):
            Hc = gcam.FindHomography(matches[i][c], 1.0, 2000, seed=5)
            Ho, _, _, _ = oracle.find_homography(matches[i][c], 1.0, 2000, 5)
            np.testing.assert_allclose(Hc, Ho, rtol=0, atol=1e-5 * np.abs(Ho).max())
            H = (Hc + H * np.float32(c)) / np.float32(c + 1)
        fused = oracle.register_and_fuse(fused, frames[i], oracle.scale_homography(H, 2.0), i + 1)
    np.testing.assert_array_equal(got, fused)
