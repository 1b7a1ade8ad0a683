"""Seeded numpy inputs for CPU tests (smooth YCbCr scenes with noise)."""
import numpy as np

SRGB_YCBCR = np.array([[0.2126, 0.7152, 0.0722],
                       [-0.1146, -0.3854, 0.5],
                       [0.5, -0.4542, -0.0458]])


def smooth_scene(h, w, seed, lo=0.02, hi=0.9, period=16.0):
    """Sum of random low-frequency sinusoids per RGB channel, mapped to [lo, hi]."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    rgb = np.zeros((h, w, 3))
    for c in range(3):
        acc = np.zeros((h, w))
        for k in range(4):
            fx, fy = rng.uniform(-1, 1, 2) / period * (k + 1)
            ph = rng.uniform(0, 2 * np.pi)
            acc += np.sin(2 * np.pi * (fx * xx + fy * yy) + ph) / (k + 1)
        acc = (acc - acc.min()) / max(acc.max() - acc.min(), 1e-12)
        rgb[..., c] = lo + (hi - lo) * acc
    return rgb


def noisy_ycc(h, w, seed, a=4e-3, b=2e-3, dtype=np.float32):
    """(noisy YCbCr (H, W, 4), clean YCbCr) with Poisson-Gaussian noise in RGB."""
    rng = np.random.default_rng(seed + 1000)
    rgb = smooth_scene(h, w, seed)
    noisy = a * rng.poisson(rgb / a) + rng.normal(0.0, b, rgb.shape)
    out = np.zeros((h, w, 4), dtype)
    out[..., :3] = noisy @ SRGB_YCBCR.T
    clean = np.zeros((h, w, 4), dtype)
    clean[..., :3] = rgb @ SRGB_YCBCR.T
    return out, clean


def default_params(levels, **kw):
    base = dict(luma=1.0, chroma=1.0, chromaBoost=1.0, gradientBoost=0.0, gradientThreshold=1.0, sharpening=1.0)
    base.update(kw)
    return [dict(base) for _ in range(levels)]


def nlf_for(levels, a=4e-3, b=2e-3):
    """A plausible per-level YCbCr noise model (A, B) for the test noise."""
    out = np.zeros((levels, 2, 3))
    for i in range(levels):
        s = 4.0 ** -i
        out[i, 0] = np.array([b * b, 0.3 * b * b, 0.3 * b * b]) * s + 1e-8
        out[i, 1] = np.array([0.6 * a, 0.25 * a, 0.25 * a]) * s
    return out


def preset_params(levels):
    """The benchmark's denoise parameters as dicts: the iPhone 14 Wide high-ISO preset
    (iPhone14WideCalibration.cpp:75-111 at nlf_alpha = 1) with the chroma multiplier at 0.25."""
    lm, cb = [3.0, 1.5, 1.0, 1.0, 1.0], [8.0, 4.0, 2.0, 2.0, 2.0]
    gb, gt = [2.0, 1.0, 0.0, 0.0, 0.0], [2.0, 2.0, 1.0, 1.0, 1.0]
    return [dict(luma=2.0 * lm[min(i, 4)], chroma=0.25, chromaBoost=cb[min(i, 4)], gradientBoost=gb[min(i, 4)],
                 gradientThreshold=gt[min(i, 4)], sharpening=1.0) for i in range(levels)]
