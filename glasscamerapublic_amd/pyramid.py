"""Host-side mirror of the reference's C++ operator surface for the denoise path.

``PyramidProcessor`` keeps the names, argument meaning and ownership rules of
``PyramidProcessor<levels>`` (MetalPipeline/pyramid_processor.hpp:22-76):
``denoise(denoiseParameters, image, gradientImage, nlfParameters,
exposure_multiplier, lensShadingCorrection, calibrateFromImage)`` returns the
level-0 denoised image owned by the processor (valid until the next call), and
``denoisedImagePyramid`` exposes every level, as ``RawConverter::denoise``
reads them (MetalPipeline/raw_converter.cpp:88,107-109).

Images are torch tensors on a HIP device: YCbCr images (H, W, 4) float32
(gls::pixel_float4), gradients (H, W, 2) float32.  torch only provides device
memory and streams; all arithmetic runs in the gfx950 kernels of
``libgcam_denoise.so`` through the C ABI.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, fields
from typing import Optional, Sequence

import torch

from . import _lib
from ._lib import (GCAM_PIXEL_PCA8_F32, GCAM_PIXEL_RG32F, GCAM_PIXEL_RGBA32F, ConfigC, DenoiseParamsC, Image,
                   NlfC, SynthParamsC, check, lib)


@dataclass
class DenoiseParameters:
    """DenoiseParameters (MetalPipeline/demosaic.hpp:34-41)."""
    luma: float = 1.0
    chroma: float = 1.0
    chromaBoost: float = 1.0
    gradientBoost: float = 0.0
    gradientThreshold: float = 1.0
    sharpening: float = 1.0


# YCbCrNLF (MetalPipeline/demosaic.hpp:54): (A, B), var = A + B * Y per channel.
YCbCrNLF = tuple


def _stream(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream if hasattr(s, "cuda_stream") else int(s)


def _fmt_of(t: torch.Tensor) -> int:
    ch = t.shape[-1]
    return {4: GCAM_PIXEL_RGBA32F, 2: GCAM_PIXEL_RG32F, 8: GCAM_PIXEL_PCA8_F32}[ch]


def as_image(t: Optional[torch.Tensor], fmt: Optional[int] = None) -> Optional[Image]:
    """gcam_image view of an (H, W, C) float32 device tensor with packed pixels."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("gcam images live in device memory (a HIP tensor)")
    if t.dtype != torch.float32 or t.dim() != 3:
        raise ValueError("expected an (H, W, C) float32 tensor")
    ch = t.shape[2]
    if t.stride(2) != 1 or t.stride(1) != ch or t.stride(0) % ch != 0:
        raise ValueError("pixels must be packed (C contiguous channels, rows of whole pixels)")
    return Image(t.data_ptr(), t.shape[1], t.shape[0], t.stride(0) // ch, _fmt_of(t) if fmt is None else fmt)


class _DeviceArray:
    """Exposes a handle-owned device buffer to torch without a copy."""

    def __init__(self, ptr: int, shape, device: torch.device):
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": "<f4", "data": (ptr, False),
                                         "version": 2, "strides": None}
        self.device = device


def _wrap(img: Image, channels: int, device: torch.device) -> torch.Tensor:
    arr = _DeviceArray(img.data, (img.height, img.width, channels), device)
    return torch.as_tensor(arr, device=device)


class PyramidProcessor:
    """PyramidProcessor<levels> (MetalPipeline/pyramid_processor.hpp:22-76) on gfx950.

    search_radius: half-width of the block-matching window; the reference
    hard-codes 10 (21 x 21, demosaic.metal:853).
    precision: "fp32" (plain form), "fp16" or "bf16" (contractions in 16 bits,
    FP32 accumulation).
    usePatchSimiliarity: the reference's switch (pyramid_processor.hpp:27);
    False selects the non-PCA denoiseImage bilateral.
    """

    def __init__(self, width: int, height: int, levels: int = 5, search_radius: int = 10,
                 precision: str = "fp32", device: Optional[torch.device] = None, bm_rows_per_thread: int = 0,
                 usePatchSimiliarity: bool = True):
        self.width, self.height, self.levels = int(width), int(height), int(levels)
        self.search_radius = int(search_radius)
        self.precision = precision
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        cfg = ConfigC(self.search_radius, _lib.GCAM_PRECISION[precision], int(bm_rows_per_thread),
                      0 if usePatchSimiliarity else 1)
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(lib().gcam_pyramid_create(self.width, self.height, self.levels, ctypes.byref(cfg),
                                            ctypes.byref(handle)))
        self._h = handle
        # Counts the calls that overwrite the handle's per-call state (denoise, denoise_fwd); a reverse
        # pass records the count of its forward and refuses to run after another forward.
        self.generation = 0

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            torch.cuda.synchronize(self.device)
            lib().gcam_pyramid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def level_size(self, i: int):
        return self.width >> i, self.height >> i

    def denoise(self, denoiseParameters: Sequence[DenoiseParameters], image: torch.Tensor,
                gradientImage: Optional[torch.Tensor], nlfParameters: list, exposure_multiplier: float = 1.0,
                lensShadingCorrection: float = 0.0, calibrateFromImage: bool = False, stream=None) -> torch.Tensor:
        """PyramidProcessor::denoise (MetalPipeline/pyramid_processor.cpp:96-172).

        nlfParameters is a list of ``levels`` (A, B) pairs; when
        calibrateFromImage is set it is overwritten with the fitted model, as the
        reference writes ``(*nlfParameters)[i]``.
        """
        if len(denoiseParameters) != self.levels or len(nlfParameters) != self.levels:
            raise ValueError(f"need {self.levels} denoise parameters and noise models")
        P = (DenoiseParamsC * self.levels)()
        for i, dp in enumerate(denoiseParameters):
            for f in fields(DenoiseParameters):
                setattr(P[i], f.name, float(getattr(dp, f.name)))
        N = (NlfC * self.levels)()
        for i, (A, B) in enumerate(nlfParameters):
            for c in range(3):
                N[i].a[c] = float(A[c])
                N[i].b[c] = float(B[c])
        img = as_image(image, GCAM_PIXEL_RGBA32F)
        grad = as_image(gradientImage, GCAM_PIXEL_RG32F)
        out = Image()
        self.generation += 1
        with torch.cuda.device(self.device):
            check(lib().gcam_pyramid_denoise(self._h, P, ctypes.byref(img),
                                             ctypes.byref(grad) if grad is not None else None, N,
                                             float(exposure_multiplier), float(lensShadingCorrection),
                                             int(bool(calibrateFromImage)), ctypes.c_void_p(_stream(stream)),
                                             ctypes.byref(out)))
        if calibrateFromImage:
            for i in range(self.levels):
                nlfParameters[i] = (tuple(N[i].a), tuple(N[i].b))
        return _wrap(out, 4, self.device)

    def _level(self, level: int, which: int, channels: int) -> torch.Tensor:
        img = Image()
        check(lib().gcam_pyramid_level(self._h, level, which, ctypes.byref(img)))
        return _wrap(img, channels, self.device)

    @property
    def denoisedImagePyramid(self):
        return [self._level(i, _lib.GCAM_LEVEL_DENOISED, 4) for i in range(self.levels)]

    @property
    def subtractedImagePyramid(self):
        return [self._level(i, _lib.GCAM_LEVEL_SUBTRACTED, 4) for i in range(self.levels - 1)]

    @property
    def imagePyramid(self):
        return [self._level(i, _lib.GCAM_LEVEL_NOISY, 4) for i in range(1, self.levels)]

    def _params_nlf(self, denoiseParameters, nlfParameters):
        P = (DenoiseParamsC * self.levels)()
        for i, dp in enumerate(denoiseParameters):
            for f in fields(DenoiseParameters):
                setattr(P[i], f.name, float(getattr(dp, f.name)))
        N = (NlfC * self.levels)()
        for i, (A, B) in enumerate(nlfParameters):
            for c in range(3):
                N[i].a[c] = float(A[c])
                N[i].b[c] = float(B[c])
        return P, N

    def denoise_fwd(self, denoiseParameters, thresholds, image: torch.Tensor, nlfParameters,
                    stream=None) -> torch.Tensor:
        """Differentiable forward: denoise with per-level (Y, Cb, Cr) thresholds (levels x 3).

        Keeps what denoise_bwd needs; `image` must stay alive until then.
        """
        P, N = self._params_nlf(denoiseParameters, nlfParameters)
        th = (ctypes.c_float * (3 * self.levels))(*[float(v) for v in torch.as_tensor(thresholds).reshape(-1)])
        out = Image()
        self.generation += 1
        with torch.cuda.device(self.device):
            check(lib().gcam_pyramid_denoise_fwd(self._h, P, th, ctypes.byref(as_image(image, GCAM_PIXEL_RGBA32F)),
                                                 N, ctypes.c_void_p(_stream(stream)), ctypes.byref(out)))
        self._fwd_image = image
        return _wrap(out, 4, self.device)

    def denoise_bwd(self, grad_output: torch.Tensor, stream=None, generation=None) -> torch.Tensor:
        """dL/dthresholds (levels x 3, float32 CPU tensor) of the last denoise_fwd, given dL/d(output).

        generation: the processor's ``generation`` right after the forward being differentiated; if
        another denoise or denoise_fwd ran on this processor since, the saved forward state is gone and
        this raises instead of differentiating the later call."""
        if generation is not None and generation != self.generation:
            raise RuntimeError(f"denoise_bwd: the processor ran {self.generation - generation} more forward "
                               "call(s) since the forward being differentiated; its saved state is overwritten")
        g = (ctypes.c_float * (3 * self.levels))()
        grad_output = grad_output.contiguous()
        with torch.cuda.device(self.device):
            check(lib().gcam_pyramid_denoise_bwd(self._h, ctypes.byref(as_image(grad_output, GCAM_PIXEL_RGBA32F)),
                                                 g, ctypes.c_void_p(_stream(stream))))
        return torch.tensor(list(g), dtype=torch.float32).reshape(self.levels, 3)

    def set_profiling(self, enable: bool = True) -> None:
        """Record HIP events around every stage of each denoise call."""
        check(lib().gcam_pyramid_set_profiling(self._h, int(bool(enable))))

    def stage_times(self) -> dict:
        """{stage: [ms per level]} of the last call (synchronizes on its events)."""
        ms = (ctypes.c_float * (self.levels * len(_lib.STAGES)))()
        check(lib().gcam_pyramid_stage_times(self._h, ms))
        n = len(_lib.STAGES)
        return {name: [ms[i * n + k] for i in range(self.levels)] for k, name in enumerate(_lib.STAGES)}

    def pcaSpace(self, level: int):
        """(basis (25, 8) float32, eigenvalues (25,) float64) of a level, from the last call."""
        b = (ctypes.c_float * 200)()
        e = (ctypes.c_double * 25)()
        check(lib().gcam_pyramid_basis(self._h, level, b, e))
        return torch.tensor(list(b)).reshape(25, 8), torch.tensor(list(e), dtype=torch.float64)

    def MeasureYCbCrNLF(self, inputImage: torch.Tensor, exposure_multiplier: float = 1.0, stream=None):
        """PyramidProcessor::MeasureYCbCrNLF (MetalPipeline/pyramid_processor.cpp:174-324)."""
        return measure_ycbcr_nlf(inputImage, exposure_multiplier, stream=stream)


# ---- per-operation functions (MetalPipeline/demosaic_kernels.hpp wrappers) ----

def downsample_image(image: torch.Tensor, out: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
    """resampleImageKernel: downsampleImageXYZ / downsampleImageXY (demosaic.metal:886-920)."""
    h, w, ch = image.shape
    if out is None:
        out = torch.empty((h // 2, w // 2, ch), dtype=torch.float32, device=image.device)
    check(lib().gcam_downsample_image(ctypes.byref(as_image(image)), ctypes.byref(as_image(out)),
                                      ctypes.c_void_p(_stream(stream))))
    return out


def subtract_noise_image(image, image1, denoised1, gradient=None, luma_weight=1.0, sharpening=1.0,
                         nlf=(0.0, 0.0), stream=None) -> torch.Tensor:
    """subtractNoiseImage (demosaic.metal:922-959)."""
    out = torch.empty_like(image)
    g = as_image(gradient)
    check(lib().gcam_subtract_noise_image(ctypes.byref(as_image(image)), ctypes.byref(as_image(image1)),
                                          ctypes.byref(as_image(denoised1)),
                                          ctypes.byref(g) if g is not None else None, float(luma_weight),
                                          float(sharpening), float(nlf[0]), float(nlf[1]),
                                          ctypes.byref(as_image(out)), ctypes.c_void_p(_stream(stream))))
    return out


def collect_patches(image: torch.Tensor, stream=None) -> torch.Tensor:
    """collectPatches (demosaic.metal:755-772): (W*H/64, 25) float32 patches."""
    h, w, _ = image.shape
    n = (w * h) // 64
    out = torch.empty((max(n, 1), 25), dtype=torch.float32, device=image.device)
    cnt = ctypes.c_int64()
    check(lib().gcam_collect_patches(ctypes.byref(as_image(image)), ctypes.c_void_p(out.data_ptr()), out.shape[0],
                                     ctypes.byref(cnt), ctypes.c_void_p(_stream(stream))))
    return out[:n]


def build_pca_space(patches: torch.Tensor, stream=None):
    """build_pca_space<25,8> (PCA.cpp:23-56) on the device: (basis (25, 8), eigenvalues (25,))."""
    patches = patches.contiguous()
    basis = torch.empty((25, 8), dtype=torch.float32, device=patches.device)
    ev = torch.empty(25, dtype=torch.float64, device=patches.device)
    check(lib().gcam_build_pca_space(ctypes.c_void_p(patches.data_ptr()), patches.shape[0],
                                     ctypes.c_void_p(basis.data_ptr()), ctypes.c_void_p(ev.data_ptr()),
                                     ctypes.c_void_p(_stream(stream))))
    return basis, ev


def pca_projection(image: torch.Tensor, basis: torch.Tensor, stream=None) -> torch.Tensor:
    """pcaProjection (demosaic.metal:774-796): (H, W, 8) float32 features."""
    h, w, _ = image.shape
    feat = torch.empty((h, w, 8), dtype=torch.float32, device=image.device)
    basis = basis.to(device=image.device, dtype=torch.float32).contiguous()
    check(lib().gcam_pca_projection(ctypes.byref(as_image(image)), ctypes.c_void_p(basis.data_ptr()),
                                    ctypes.byref(as_image(feat, GCAM_PIXEL_PCA8_F32)),
                                    ctypes.c_void_p(_stream(stream))))
    return feat


def block_matching_denoise_image(image, features, var_a, var_b, thresholdMultipliers, chromaBoost=1.0,
                                 gradientBoost=0.0, gradientThreshold=1.0, lensShadingCorrection=0.0,
                                 search_radius=10, gradient=None, stream=None) -> torch.Tensor:
    """blockMatchingDenoiseImage (demosaic.metal:806-884)."""
    out = torch.empty_like(image)
    va = (ctypes.c_float * 3)(*[float(v) for v in var_a])
    vb = (ctypes.c_float * 3)(*[float(v) for v in var_b])
    th = (ctypes.c_float * 3)(*[float(v) for v in thresholdMultipliers])
    g = as_image(gradient)
    check(lib().gcam_block_matching_denoise_image(
        ctypes.byref(as_image(image)), ctypes.byref(g) if g is not None else None,
        ctypes.byref(as_image(features, GCAM_PIXEL_PCA8_F32)), va, vb, th, float(chromaBoost), float(gradientBoost),
        float(gradientThreshold), float(lensShadingCorrection), int(search_radius), ctypes.byref(as_image(out)),
        ctypes.c_void_p(_stream(stream))))
    return out


def block_matching_denoise_image_lp(image, basis, var_a, var_b, thresholdMultipliers, chromaBoost=1.0,
                                    gradientBoost=0.0, gradientThreshold=1.0, lensShadingCorrection=0.0,
                                    search_radius=10, gradient=None, precision="fp16", stream=None) -> torch.Tensor:
    """Precision form of pcaProjection + blockMatchingDenoiseImage (features projected in-kernel)."""
    out = torch.empty_like(image)
    basis = basis.to(device=image.device, dtype=torch.float32).contiguous()
    va = (ctypes.c_float * 3)(*[float(v) for v in var_a])
    vb = (ctypes.c_float * 3)(*[float(v) for v in var_b])
    th = (ctypes.c_float * 3)(*[float(v) for v in thresholdMultipliers])
    g = as_image(gradient)
    check(lib().gcam_block_matching_denoise_image_lp(
        ctypes.byref(as_image(image)), ctypes.byref(g) if g is not None else None, ctypes.c_void_p(basis.data_ptr()),
        va, vb, th, float(chromaBoost), float(gradientBoost), float(gradientThreshold), float(lensShadingCorrection),
        int(search_radius), _lib.GCAM_PRECISION[precision], ctypes.byref(as_image(out)),
        ctypes.c_void_p(_stream(stream))))
    return out


def denoise_image(image, var_a, var_b, thresholdMultipliers, chromaBoost=1.0, gradientBoost=0.0,
                  gradientThreshold=1.0, gradient=None, stream=None) -> torch.Tensor:
    """denoiseImage (demosaic.metal:664-719): the non-PCA bilateral."""
    out = torch.empty_like(image)
    va = (ctypes.c_float * 3)(*[float(v) for v in var_a])
    vb = (ctypes.c_float * 3)(*[float(v) for v in var_b])
    th = (ctypes.c_float * 3)(*[float(v) for v in thresholdMultipliers])
    g = as_image(gradient)
    check(lib().gcam_denoise_image(ctypes.byref(as_image(image)), ctypes.byref(g) if g is not None else None, va, vb,
                                   th, float(chromaBoost), float(gradientBoost), float(gradientThreshold),
                                   ctypes.byref(as_image(out)), ctypes.c_void_p(_stream(stream))))
    return out


def basic_noise_statistics(image: torch.Tensor, stream=None) -> torch.Tensor:
    """basicNoiseStatistics (demosaic.metal:1138-1159): (meanY, varY, varCb, varCr)."""
    out = torch.empty_like(image)
    check(lib().gcam_basic_noise_statistics(ctypes.byref(as_image(image)), ctypes.byref(as_image(out)),
                                            ctypes.c_void_p(_stream(stream))))
    return out


def measure_ycbcr_nlf(image: torch.Tensor, exposure_multiplier: float = 1.0, stream=None):
    """MeasureYCbCrNLF (pyramid_processor.cpp:174-324), FP64 fit on the device: (A, B)."""
    scratch = torch.empty_like(image)
    n = NlfC()
    check(lib().gcam_measure_ycbcr_nlf(ctypes.byref(as_image(image)), ctypes.byref(as_image(scratch)),
                                       float(exposure_multiplier), ctypes.byref(n), ctypes.c_void_p(_stream(stream))))
    return tuple(n.a), tuple(n.b)


def synth_frame(width: int, height: int, shot_gain: float, read_sigma: float, scene_seed: int = 1234,
                noise_seed: int = 5678, lo: float = 0.02, hi: float = 0.9, period: float = 64.0, octaves: int = 3,
                device=None, want_raw: bool = False, want_clean: bool = False, stream=None):
    """Seeded RGGB 12-bit Poisson-Gaussian frame, bilinear-demosaiced to YCbCr on the device.

    Returns the (H, W, 4) YCbCr tensor, plus the raw (H, W) int16-stored DN
    and/or the clean YCbCr scene when asked.
    """
    device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    ycc = torch.empty((height, width, 4), dtype=torch.float32, device=device)
    raw = torch.empty((height, width), dtype=torch.int16, device=device) if want_raw else None
    clean = torch.empty((height, width, 4), dtype=torch.float32, device=device) if want_clean else None
    p = SynthParamsC(scene_seed, noise_seed, shot_gain, read_sigma, lo, hi, period, octaves)
    cimg = as_image(clean)
    check(lib().gcam_synth_frame(width, height, ctypes.byref(p),
                                 ctypes.c_void_p(raw.data_ptr()) if raw is not None else None,
                                 ctypes.byref(as_image(ycc)), ctypes.byref(cimg) if cimg is not None else None,
                                 ctypes.c_void_p(_stream(stream))))
    out = [ycc]
    if want_raw:
        out.append(raw)
    if want_clean:
        out.append(clean)
    return out[0] if len(out) == 1 else tuple(out)
