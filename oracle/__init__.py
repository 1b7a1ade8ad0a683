"""ctypes wrapper of the CPU oracle (``libgcam_oracle.so``).

TEST INFRASTRUCTURE ONLY: the oracle restates the reference's denoise path
(MetalPipeline/demosaic.metal, MetalPipeline/PCA.cpp,
MetalPipeline/pyramid_processor.cpp) on the CPU so that the HIP product can be
checked against it.  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU
baseline leg of ``bench.py`` import this module; the product package never does.

Images are numpy arrays: YCbCr images are (H, W, 4), gradients (H, W, 2),
features (H, W, 8), all C-contiguous float32 or float64.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GCAM_ORACLE_LIB selects another build of the same sources (the Makefile's sanitizer build).
_LIB_PATH = os.environ.get("GCAM_ORACLE_LIB") or os.path.join(_HERE, "libgcam_oracle.so")
_lib = None


def build() -> str:
    """Compile the oracle with its Makefile; returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "libgcam_oracle.so"], check=True)
    return os.path.join(_HERE, "libgcam_oracle.so")


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_fp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_d = ctypes.c_double


class DenoiseParamsC(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in
                ("luma", "chroma", "chromaBoost", "gradientBoost", "gradientThreshold", "sharpening")]


def _declare(L):
    for sfx, R in (("f32", _f), ("f64", _d)):
        getattr(L, f"oracle_downsample_{sfx}").argtypes = [_fp, _i, _i, _i, _fp, _i, _i]
        getattr(L, f"oracle_subtract_noise_{sfx}").argtypes = [_fp, _i, _i, _fp, _fp, _i, _i, _fp, R, R, R, R, _fp]
        getattr(L, f"oracle_collect_patches_{sfx}").argtypes = [_fp, _i, _i, _fp]
        getattr(L, f"oracle_collect_patches_{sfx}").restype = _i
        getattr(L, f"oracle_patch_covariance_{sfx}").argtypes = [_fp, _i, _fp]
        getattr(L, f"oracle_build_pca_space_{sfx}").argtypes = [_fp, _i, _fp, _fp]
        getattr(L, f"oracle_pca_project_{sfx}").argtypes = [_fp, _i, _i, _fp, _fp]
        getattr(L, f"oracle_block_match_{sfx}").argtypes = [_fp, _fp, _fp, _i, _i, _fp, _fp, _fp, R, R, R, R, _i, _fp]
        getattr(L, f"oracle_noise_stats_{sfx}").argtypes = [_fp, _i, _i, _fp]
        getattr(L, f"oracle_denoise_image_{sfx}").argtypes = [_fp, _fp, _i, _i, _fp, _fp, _fp, R, R, R, _fp]
        getattr(L, f"oracle_pyramid_denoise_{sfx}").argtypes = [_fp, _fp, _i, _i, _i, _i, _i, _fp, _fp, R, R, _i,
                                                                 _fp, _fp, _fp, _fp]
        getattr(L, f"oracle_pyramid_denoise_{sfx}").restype = _i
    L.oracle_eigh_jacobi.argtypes = [_fp, _i, _fp, _fp]
    L.oracle_fit_ycbcr_nlf.argtypes = [_fp, _i, _i, _d, _fp, _fp, _fp]
    L.oracle_fit_ycbcr_nlf.restype = _i
    L.oracle_scale_raw.argtypes = [_fp, _i, _i, _i, _fp, _f, _f, _fp]
    L.oracle_raw_sobel.argtypes = [_fp, _i, _i, _fp]
    L.oracle_gaussian_bilinear_weights.argtypes = [_f, _fp, _i]
    L.oracle_gaussian_bilinear_weights.restype = _i
    L.oracle_blur_sobel.argtypes = [_fp, _fp, _i, _i, _fp, _i, _fp, _i, _f, _f, _fp]
    L.oracle_interpolate_green.argtypes = [_fp, _fp, _i, _i, _i, _f, _f, _fp]
    L.oracle_interpolate_red_blue.argtypes = [_fp, _fp, _fp, _i, _i, _i, _fp, _fp, _fp]
    L.oracle_interpolate_red_blue_at_green.argtypes = [_fp, _fp, _i, _i, _i, _fp, _fp, _fp]
    L.oracle_blend_highlights.argtypes = [_fp, _i, _i, _f, _fp]
    L.oracle_transform.argtypes = [_fp, _i, _i, _fp, _fp]
    L.oracle_despeckle.argtypes = [_fp, _i, _i, _fp, _fp, _fp]
    L.oracle_cam_ycbcr.argtypes = [_fp, _fp]
    L.oracle_raw_noise_stats.argtypes = [_fp, _i, _i, _i, _fp, _fp]
    L.oracle_histogram.argtypes = [_fp, _i, _i, _fp]
    L.oracle_guided_filter_ab.argtypes = [_fp, _i, _i, _f, _fp]
    L.oracle_box_filter_gf.argtypes = [_fp, _i, _i, _fp]
    L.oracle_ltm_mask.argtypes = [_fp, _fp, _i, _i, _fp, _i, _i, _fp, _i, _i, _fp, _i, _i, _fp, _f, _f, _f, _f, _fp]
    L.oracle_histogram_statistics.argtypes = [_fp, _fp, _i, _i, _fp]
    L.oracle_fit_raw_nlf.argtypes = [_fp, _fp, _i, _i, _d, _d, _fp, _fp, _fp]
    L.oracle_register_and_fuse.argtypes = [_fp, _fp, _i, _i, _i, _i, _fp, _i, _fp]
    L.oracle_register_image.argtypes = [_fp, _i, _i, _i, _i, _i, _fp, _fp]
    L.oracle_register_bayer.argtypes = [_fp, _i, _i, _fp, _fp]
    L.oracle_jacobi9.argtypes = [_fp, _fp, _fp]
    L.oracle_least_squares_homography.argtypes = [_fp, _i, _fp]
    L.oracle_least_squares_homography.restype = _i
    L.oracle_ransac_samples.argtypes = [ctypes.c_uint64, _i, _i, _i, _fp]
    L.oracle_find_homography.argtypes = [_fp, _i, _f, _i, ctypes.c_uint64, _fp, _fp, _fp, _fp]
    L.oracle_set_threads.argtypes = [_i]
    L.oracle_get_threads.restype = _i


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _sfx(dtype):
    dt = np.dtype(dtype)
    if dt == np.float32:
        return "f32"
    if dt == np.float64:
        return "f64"
    raise TypeError(f"oracle works in float32 or float64, got {dt}")


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def set_threads(n: int) -> None:
    lib().oracle_set_threads(int(n))


def level_sizes(w: int, h: int, levels: int):
    """Pyramid level sizes, width / 2^i (MetalPipeline/pyramid_processor.cpp:38-47)."""
    return [(w >> i, h >> i) for i in range(levels)]


def downsample(img, out_w=None, out_h=None):
    h, w, ch = img.shape
    out_w = w // 2 if out_w is None else out_w
    out_h = h // 2 if out_h is None else out_h
    img = _c(img, img.dtype)
    out = np.zeros((out_h, out_w, ch), img.dtype)
    getattr(lib(), f"oracle_downsample_{_sfx(img.dtype)}")(_p(img), w, h, ch, _p(out), out_w, out_h)
    return out


def subtract_noise(img, img1, den1, grad=None, luma_weight=1.0, sharpening=1.0, nlf=(0.0, 0.0)):
    dt = img.dtype
    h, w, _ = img.shape
    h1, w1, _ = img1.shape
    img, img1, den1 = _c(img, dt), _c(img1, dt), _c(den1, dt)
    grad = None if grad is None else _c(grad, dt)
    out = np.zeros_like(img)
    getattr(lib(), f"oracle_subtract_noise_{_sfx(dt)}")(_p(img), w, h, _p(img1), _p(den1), w1, h1, _p(grad),
                                                        luma_weight, sharpening, nlf[0], nlf[1], _p(out))
    return out


def collect_patches(img):
    dt = img.dtype
    h, w, _ = img.shape
    img = _c(img, dt)
    n = (w * h) // 64
    out = np.zeros((max(n, 1), 25), dt)
    n2 = getattr(lib(), f"oracle_collect_patches_{_sfx(dt)}")(_p(img), w, h, _p(out))
    assert n2 == n
    return out[:n]


def patch_covariance(patches):
    patches = np.ascontiguousarray(patches)
    cov = np.zeros((25, 25), np.float64)
    getattr(lib(), f"oracle_patch_covariance_{_sfx(patches.dtype)}")(_p(patches), patches.shape[0], _p(cov))
    return cov


def eigh_jacobi(mat):
    mat = _c(mat, np.float64)
    n = mat.shape[0]
    vals = np.zeros(n, np.float64)
    vecs = np.zeros((n, n), np.float64)
    lib().oracle_eigh_jacobi(_p(mat), n, _p(vals), _p(vecs))
    return vals, vecs


def build_pca_space(patches):
    """FP64 basis (25, 8) and the 25 eigenvalues, descending (MetalPipeline/PCA.cpp:23-56)."""
    patches = np.ascontiguousarray(patches)
    basis = np.zeros((25, 8), np.float64)
    ev = np.zeros(25, np.float64)
    getattr(lib(), f"oracle_build_pca_space_{_sfx(patches.dtype)}")(_p(patches), patches.shape[0], _p(basis),
                                                                     _p(ev))
    return basis, ev


def pca_project(img, basis):
    dt = img.dtype
    h, w, _ = img.shape
    img = _c(img, dt)
    basis = _c(basis, dt)
    feat = np.zeros((h, w, 8), dt)
    getattr(lib(), f"oracle_pca_project_{_sfx(dt)}")(_p(img), w, h, _p(basis), _p(feat))
    return feat


def block_match(img, feat, var_a, var_b, thr, chroma_boost=1.0, gradient_boost=0.0, gradient_threshold=1.0,
                lens_shading=0.0, radius=10, grad=None):
    dt = img.dtype
    h, w, _ = img.shape
    img, feat = _c(img, dt), _c(feat, dt)
    grad = None if grad is None else _c(grad, dt)
    va, vb, th = (_c(np.asarray(v).reshape(3), dt) for v in (var_a, var_b, thr))
    out = np.zeros_like(img)
    getattr(lib(), f"oracle_block_match_{_sfx(dt)}")(_p(img), _p(grad), _p(feat), w, h, _p(va), _p(vb), _p(th),
                                                     chroma_boost, gradient_boost, gradient_threshold, lens_shading,
                                                     radius, _p(out))
    return out


def noise_stats(img):
    dt = img.dtype
    h, w, _ = img.shape
    img = _c(img, dt)
    out = np.zeros_like(img)
    getattr(lib(), f"oracle_noise_stats_{_sfx(dt)}")(_p(img), w, h, _p(out))
    return out


@dataclass
class NlfFit:
    A: np.ndarray
    B: np.ndarray
    accepted: bool
    histogram: np.ndarray
    n_first: float
    n_second: float
    err2: np.ndarray
    new_err2: np.ndarray
    discarded: int
    variance_max: float


def fit_ycbcr_nlf(stats, exposure_multiplier=1.0) -> NlfFit:
    stats = _c(stats, np.float32)
    h, w, _ = stats.shape
    A = np.zeros(3)
    B = np.zeros(3)
    diag = np.zeros(16)
    acc = lib().oracle_fit_ycbcr_nlf(_p(stats), w, h, exposure_multiplier, _p(A), _p(B), _p(diag))
    return NlfFit(A, B, bool(acc), diag[:6].copy(), diag[6], diag[7], diag[8:11].copy(), diag[11:14].copy(),
                  int(diag[14]), diag[15])


@dataclass
class PyramidResult:
    denoised: list
    subtracted: list
    bases: np.ndarray
    eigvals: np.ndarray
    nlf: np.ndarray


def denoise_image(img, var_a, var_b, thr, chroma_boost=1.0, gradient_boost=0.0, gradient_threshold=1.0,
                  grad=None):
    """denoiseImage (demosaic.metal:664-719): the non-PCA bilateral."""
    dt = img.dtype
    h, w, _ = img.shape
    img = _c(img, dt)
    grad = None if grad is None else _c(grad, dt)
    va, vb, th = (_c(np.asarray(v).reshape(3), dt) for v in (var_a, var_b, thr))
    out = np.zeros_like(img)
    getattr(lib(), f"oracle_denoise_image_{_sfx(dt)}")(_p(img), _p(grad), w, h, _p(va), _p(vb), _p(th),
                                                       chroma_boost, gradient_boost, gradient_threshold, _p(out))
    return out


def pyramid_denoise(image, params, nlf, levels=5, radius=10, gradient=None, exposure_multiplier=1.0,
                    lens_shading=0.0, calibrate=False, dtype=np.float32, patch_similarity=True) -> PyramidResult:
    """Whole multi-level denoise (MetalPipeline/pyramid_processor.cpp:96-172).

    params: sequence of ``levels`` dicts/objects with the DenoiseParameters fields.
    nlf: (levels, 2, 3) array of (A, B) per level; returned (updated if calibrating).
    """
    image = _c(image, dtype)
    h, w, _ = image.shape
    grad = None if gradient is None else _c(gradient, dtype)
    P = (DenoiseParamsC * levels)()
    for i, p in enumerate(params):
        for k in ("luma", "chroma", "chromaBoost", "gradientBoost", "gradientThreshold", "sharpening"):
            setattr(P[i], k, float(p[k] if isinstance(p, dict) else getattr(p, k)))
    nlf_arr = np.ascontiguousarray(np.asarray(nlf, np.float64).reshape(levels, 6).copy())
    sizes = level_sizes(w, h, levels)
    total = sum(a * b for a, b in sizes)
    den = np.zeros(total * 4, dtype)
    sub = np.zeros(total * 4, dtype)
    bases = np.zeros((levels, 25, 8))
    ev = np.zeros((levels, 25))
    rc = getattr(lib(), f"oracle_pyramid_denoise_{_sfx(dtype)}")(
        _p(image), _p(grad), w, h, levels, radius, int(bool(patch_similarity)), ctypes.cast(P, ctypes.c_void_p),
        _p(nlf_arr),
        exposure_multiplier, lens_shading, int(bool(calibrate)), _p(den), _p(sub), _p(bases), _p(ev))
    if rc != 0:
        raise ValueError("pyramid level too small for PCA patch sampling")
    dl, sl, o = [], [], 0
    for (lw, lh) in sizes:
        n = lw * lh * 4
        dl.append(den[o:o + n].reshape(lh, lw, 4))
        sl.append(sub[o:o + n].reshape(lh, lw, 4))
        o += n
    return PyramidResult(dl, sl, bases, ev, nlf_arr.reshape(levels, 2, 3))


def pyramid_loss_grad(image, clean, thr, chroma_boost, nlf, levels=5, radius=4, want_grad=True):
    """Differentiation form, FP64 (oracle_pyramid_loss_grad_f64): (loss, dL/dthr (levels, 3) or None, denoised0).

    thr: (levels, 3) per-channel thresholds; chroma_boost: (levels,); nlf: (levels, 2, 3).
    """
    L = lib()
    if not hasattr(L, "_loss_grad_declared"):
        L.oracle_pyramid_loss_grad_f64.argtypes = [_fp, _fp, _i, _i, _i, _i, _fp, _fp, _fp, _fp, _fp]
        L.oracle_pyramid_loss_grad_f64.restype = _d
        L._loss_grad_declared = True
    image = _c(image, np.float64)
    clean = _c(clean, np.float64)
    h, w, _ = image.shape
    thr = _c(np.asarray(thr).reshape(levels, 3), np.float64)
    cb = _c(np.asarray(chroma_boost).reshape(levels), np.float64)
    nl = _c(np.asarray(nlf).reshape(levels, 6), np.float64)
    g = np.zeros((levels, 3)) if want_grad else None
    den = np.zeros((h, w, 4))
    loss = L.oracle_pyramid_loss_grad_f64(_p(image), _p(clean), w, h, levels, radius, _p(thr), _p(cb), _p(nl),
                                          _p(g), _p(den))
    return loss, g, den


# ---------------------------------------------------------------------------
# Raw front end (FP32): the steps of RawConverter::demosaic that produce the
# denoise's YCbCr image and gradient, and the despeckle of
# RawConverter::denoise (MetalPipeline/raw_converter.cpp:148-213, 70-78).
# Images are (H, W) for one channel, (H, W, C) otherwise; raw is uint16.

BAYER = {"grbg": 0, "gbrg": 1, "rggb": 2, "bggr": 3}


def _f32(a):
    return _c(a, np.float32)


def _vec(v, n):
    a = np.ascontiguousarray(np.asarray(v, np.float32).reshape(-1))
    if a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def scale_raw(raw, bayer=2, scale_mul=(1.0, 1.0, 1.0, 1.0), black_level=0.0, lens_correction=0.0):
    """scaleRawData, MetalPipeline/demosaic.metal:95-117."""
    raw = np.ascontiguousarray(raw, np.uint16)
    h, w = raw.shape
    out = np.zeros((h, w), np.float32)
    sm = _vec(scale_mul, 4)
    lib().oracle_scale_raw(_p(raw), w, h, int(bayer), _p(sm), float(black_level), float(lens_correction), _p(out))
    return out


def raw_sobel(scaled):
    """rawImageSobel, MetalPipeline/demosaic.metal:120-139: (gx, gy, |gx|, |gy|)."""
    scaled = _f32(scaled)
    h, w = scaled.shape
    out = np.empty((h, w, 4), np.float32)
    lib().oracle_raw_sobel(_p(scaled), w, h, _p(out))
    return out


def gaussian_bilinear_weights(radius):
    """gaussianKernelBilinearWeights, MetalPipeline/demosaic_utils.cpp:987-1006: (n, 3) taps."""
    n = lib().oracle_gaussian_bilinear_weights(float(radius), None, 0)
    out = np.empty((n, 3), np.float32)
    lib().oracle_gaussian_bilinear_weights(float(radius), _p(out), n)
    return out


def blur_sobel(scaled, sobel, raw_variance, radius1=1.5, radius2=4.5):
    """sampledConvolutionSobel with the two kernels of gaussianBlurSobelImageKernel
    (MetalPipeline/demosaic.metal:188-210, demosaic_kernels.hpp:66-97)."""
    scaled, sobel = _f32(scaled), _f32(sobel)
    h, w = scaled.shape
    w1, w2 = gaussian_bilinear_weights(radius1), gaussian_bilinear_weights(radius2)
    out = np.empty((h, w, 2), np.float32)
    va, vb = (float(x) for x in raw_variance)
    lib().oracle_blur_sobel(_p(scaled), _p(sobel), w, h, _p(w1), len(w1), _p(w2), len(w2), va, vb, _p(out))
    return out


def raw_variance(raw_nlf_a, raw_nlf_b):
    """getRawVariance, MetalPipeline/demosaic.hpp:186-193: (red, green, blue) x (a, b)."""
    a, b = np.asarray(raw_nlf_a, np.float32), np.asarray(raw_nlf_b, np.float32)
    return np.array([[a[0], b[0]], [(a[1] + a[3]) / 2, (b[1] + b[3]) / 2], [a[2], b[2]]], np.float32)


def demosaic(scaled, gradient, bayer, rvar):
    """demosaicImageKernel, MetalPipeline/demosaic_kernels.hpp:124-181: interpolateGreen,
    interpolateRedBlue, interpolateRedBlueAtGreen (demosaic.metal:214-521).
    rvar = raw_variance(...).  Returns (green, rgb after red/blue, final rgb)."""
    scaled, gradient = _f32(scaled), _f32(gradient)
    h, w = scaled.shape
    rv = np.ascontiguousarray(rvar, np.float32)
    green = np.empty((h, w), np.float32)
    lib().oracle_interpolate_green(_p(scaled), _p(gradient), w, h, int(bayer), float(rv[1, 0]), float(rv[1, 1]),
                                   _p(green))
    tmp = np.empty((h, w, 4), np.float32)
    r, b = np.ascontiguousarray(rv[0]), np.ascontiguousarray(rv[2])
    lib().oracle_interpolate_red_blue(_p(scaled), _p(green), _p(gradient), w, h, int(bayer), _p(r), _p(b), _p(tmp))
    rgb = np.empty((h, w, 4), np.float32)
    lib().oracle_interpolate_red_blue_at_green(_p(tmp), _p(gradient), w, h, int(bayer), _p(r), _p(b), _p(rgb))
    return green, tmp, rgb


def blend_highlights(img, clip=1.0):
    """blendHighlightsImage, MetalPipeline/demosaic.metal:536-564."""
    img = _f32(img)
    h, w = img.shape[:2]
    out = np.empty((h, w, 4), np.float32)
    lib().oracle_blend_highlights(_p(img), w, h, float(clip), _p(out))
    return out


def transform(img, m):
    """transformImage, MetalPipeline/demosaic.metal:649-662."""
    img = _f32(img)
    h, w = img.shape[:2]
    out = np.empty((h, w, 4), np.float32)
    lib().oracle_transform(_p(img), w, h, _p(_vec(m, 9)), _p(out))
    return out


def despeckle(img, var_a, var_b):
    """despeckleLumaMedianChromaImage, MetalPipeline/demosaic.metal:1242-1283."""
    img = _f32(img)
    h, w = img.shape[:2]
    out = np.empty((h, w, 4), np.float32)
    lib().oracle_despeckle(_p(img), w, h, _p(_vec(var_a, 3)), _p(_vec(var_b, 3)), _p(out))
    return out


def cam_ycbcr(rgb_cam=np.eye(3)):
    """cam_ycbcr(rgb_cam, xyz_sRGB), MetalPipeline/demosaic_utils.cpp:703-720."""
    out = np.empty(9, np.float32)
    lib().oracle_cam_ycbcr(_p(_vec(rgb_cam, 9)), _p(out))
    return out.reshape(3, 3)


def raw_noise_stats(scaled, bayer=2):
    """basicRawNoiseStatistics, MetalPipeline/demosaic.metal:1096-1136: (mean, var), (H/2, W/2, 4)."""
    scaled = _f32(scaled)
    h, w = scaled.shape
    mean = np.empty((h // 2, w // 2, 4), np.float32)
    var = np.empty((h // 2, w // 2, 4), np.float32)
    lib().oracle_raw_noise_stats(_p(scaled), w, h, int(bayer), _p(mean), _p(var))
    return mean, var


@dataclass
class RawNlfFit:
    a: np.ndarray
    b: np.ndarray
    n1: float
    n2: float
    accepted: bool
    variance_max: float
    err2: np.ndarray


def fit_raw_nlf(mean, var, raw_w, raw_h, exposure_multiplier=1.0) -> RawNlfFit:
    """MeasureRawNLF's FP64 fit (MetalPipeline/raw_converter.cpp:352-489) on statistics images."""
    mean, var = _f32(mean), _f32(var)
    qh, qw = mean.shape[:2]
    a, b, diag = np.empty(4), np.empty(4), np.empty(8)
    lib().oracle_fit_raw_nlf(_p(mean), _p(var), qw, qh, float(exposure_multiplier), float(raw_w) * float(raw_h),
                             _p(a), _p(b), _p(diag))
    return RawNlfFit(a, b, diag[0], diag[1], bool(diag[2]), diag[3], diag[4:])


def measure_raw_nlf(scaled, bayer=2, exposure_multiplier=1.0) -> RawNlfFit:
    """MeasureRawNLF, MetalPipeline/raw_converter.cpp:339-490."""
    h, w = np.asarray(scaled).shape
    mean, var = raw_noise_stats(scaled, bayer)
    return fit_raw_nlf(mean, var, w, h, exposure_multiplier)


def raw_front_end(raw, bayer=2, scale_mul=(1.0, 1.0, 1.0, 1.0), black_level=0.0, lens_correction=0.0,
                  raw_nlf_a=(1e-5,) * 4, raw_nlf_b=(1e-3,) * 4, rgb_cam=np.eye(3), clip=1.0, despeckle_nlf=None,
                  calibrate=False, exposure_multiplier=1.0):
    """The denoise inputs from a raw frame (MetalPipeline/raw_converter.cpp:179-214 and 70-78):
    returns (ycbcr image after the despeckle, gradient, ycbcr before it); with calibrate the raw
    noise model is fitted from the frame (MeasureRawNLF) and replaces raw_nlf_a / raw_nlf_b."""
    scaled = scale_raw(raw, bayer, scale_mul, black_level, lens_correction)
    if calibrate:
        fit = measure_raw_nlf(scaled, bayer, exposure_multiplier)
        raw_nlf_a, raw_nlf_b = fit.a.astype(np.float32), fit.b.astype(np.float32)
    rv = raw_variance(raw_nlf_a, raw_nlf_b)
    grad = blur_sobel(scaled, raw_sobel(scaled), rv[1])
    _, _, rgb = demosaic(scaled, grad, bayer, rv)
    ycc = transform(blend_highlights(rgb, clip), cam_ycbcr(rgb_cam))
    if despeckle_nlf is None:
        return ycc, grad, ycc
    return despeckle(ycc, despeckle_nlf[0], despeckle_nlf[1]), grad, ycc


def histogram(img):
    """histogramImage + histogramStatistics, MetalPipeline/demosaic.metal:1357-1453, on a zeroed
    histogram_data: returns dict(histogram[256], bands[8], black_level, white_level, shadows,
    highlights, mean, median)."""
    img = _f32(img)
    h, w = img.shape[:2]
    hist = np.zeros(256, np.uint32)
    bands = np.zeros(8, np.uint32)
    stats = np.zeros(6, np.float32)
    lib().oracle_histogram(_p(img), w, h, _p(hist))
    lib().oracle_histogram_statistics(_p(hist), _p(bands), w, h, _p(stats))
    keys = ("black_level", "white_level", "shadows", "highlights", "mean", "median")
    return dict(histogram=hist, bands=bands, **{k: float(v) for k, v in zip(keys, stats)})


LTM_DEFAULTS = dict(eps=0.01, shadows=0.8, highlights=1.05, detail=(1.1, 1.2, 1.3))  # demosaic.hpp:95-100


def guided_filter_ab(guide, eps):
    """GuidedFilterABImage, MetalPipeline/demosaic.metal:1295-1323: (H, W, 2) of (a, b)."""
    guide = _f32(guide)
    h, w = guide.shape[:2]
    out = np.empty((h, w, 2), np.float32)
    lib().oracle_guided_filter_ab(_p(guide), w, h, float(eps), _p(out))
    return out


def box_filter_gf(img2):
    """BoxFilterGFImage, MetalPipeline/demosaic.metal:1325-1355."""
    img2 = _f32(img2)
    h, w = img2.shape[:2]
    out = np.empty((h, w, 2), np.float32)
    lib().oracle_box_filter_gf(_p(img2), w, h, _p(out))
    return out


def ltm_mask(image, gradient, guides, params=None, nlf=(1e-5, 1e-3), hist_shadows=0.8, hist_highlights=1.0):
    """LocalToneMapping::createMask (MetalPipeline/raw_converter.hpp:61-75, demosaic_kernels.hpp:537-600):
    guided-filter (a, b) of the three guides (denoised levels 4, 2, 0), their box means, and the
    mask.  Returns (mask (H, W), [abMean lf, mf, hf])."""
    p = dict(LTM_DEFAULTS, **(params or {}))
    image, gradient = _f32(image), _f32(gradient)
    h, w = image.shape[:2]
    means = [box_filter_gf(guided_filter_ab(g, p["eps"])) for g in guides]
    par = np.array([p["eps"], p["shadows"], p["highlights"], *p["detail"]], np.float32)
    mask = np.empty((h, w), np.float32)
    (lh, lw), (mh, mw), (hh, hw) = (m.shape[:2] for m in means)
    lib().oracle_ltm_mask(_p(image), _p(gradient), w, h, _p(means[0]), lw, lh, _p(means[1]), mw, mh,
                          _p(means[2]), hw, hh, _p(par), float(nlf[0]), float(nlf[1]), float(hist_shadows),
                          float(hist_highlights), _p(mask))
    return mask, means


# ---- burst registration and fusion (ImageRegistration/) ----

def _h9(H):
    return np.ascontiguousarray(np.asarray(H, np.float32).reshape(9))


def register_and_fuse(fused, image, H, count):
    """registerAndFuse, ImageRegistration/SURF.metal:493-516: ((count - 1) fused + image(H p)) / count
    over fused's size ((H, W, 4) float32)."""
    fused, image = _f32(fused), _f32(image)
    oh, ow = fused.shape[:2]
    ih, iw = image.shape[:2]
    out = np.empty_like(fused)
    lib().oracle_register_and_fuse(_p(fused), _p(image), ow, oh, iw, ih, _p(_h9(H)), int(count), _p(out))
    return out


def register_image(image, H):
    """registerImage, ImageRegistration/SURF.metal:518-538 ((H, W) or (H, W, 4) float32)."""
    image = _f32(image)
    h, w = image.shape[:2]
    ch = 1 if image.ndim == 2 else image.shape[2]
    out = np.empty_like(image)
    lib().oracle_register_image(_p(image), ch, w, h, w, h, _p(_h9(H)), _p(out))
    return out


def register_bayer(raw, H):
    """registerBayerImage, ImageRegistration/SURF.metal:540-559 ((H, W) uint16)."""
    raw = np.ascontiguousarray(raw, np.uint16)
    h, w = raw.shape
    out = np.empty_like(raw)
    lib().oracle_register_bayer(_p(raw), w, h, _p(_h9(H)), _p(out))
    return out


def jacobi9(a):
    """The 9 x 9 symmetric eigensolve of LeastSquaresHomography.cpp:216-349: (eigenvalues
    descending, eigenvectors as rows)."""
    a = np.ascontiguousarray(a, np.float64)
    w = np.empty(9)
    v = np.empty((9, 9))
    lib().oracle_jacobi9(_p(a), _p(w), _p(v))
    return w, v


def least_squares_homography(matches):
    """FindLeastSquaresHomography, LeastSquaresHomography.cpp:356-421, on (n, 4) matches
    (x1, y1, x2, y2); None where the reference throws std::range_error."""
    m = _f32(matches)
    H = np.empty(9, np.float32)
    if lib().oracle_least_squares_homography(_p(m), m.shape[0], _p(H)) != 0:
        return None
    return H.reshape(3, 3)


def ransac_samples(seed, count, first, n):
    out = np.empty((n, 4), np.int32)
    lib().oracle_ransac_samples(int(seed), int(count), int(first), int(n), _p(out))
    return out


def find_homography(matches, threshold=1.0, max_iterations=2000, seed=0):
    """FindHomography, ImageRegistration/Homography.cpp:72-116: (H, inlier indices, best
    hypothesis, its outlier count)."""
    m = _f32(matches)
    n = m.shape[0]
    H = np.empty(9, np.float32)
    inl = np.empty(n, np.int32)
    cnt = np.zeros(1, np.int32)
    best = np.zeros(2, np.int32)
    lib().oracle_find_homography(_p(m), n, float(threshold), int(max_iterations), int(seed), _p(H), _p(inl),
                                 _p(cnt), _p(best))
    return H.reshape(3, 3), inl[:cnt[0]].copy(), int(best[0]), int(best[1])


def scale_homography(H, scale):
    """ScaleHomography, ImageRegistration/Homography.hpp:29-37: S H S^-1, S = diag(scale, scale, 1)."""
    H = np.asarray(H, np.float32).reshape(3, 3)
    s = np.array([scale, scale, 1.0], np.float32)
    inv = np.array([1.0 / np.float32(scale), 1.0 / np.float32(scale), 1.0], np.float32)
    return ((s[:, None] * H) * inv[None, :]).astype(np.float32)
