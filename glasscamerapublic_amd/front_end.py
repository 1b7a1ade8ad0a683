"""Raw front end: from a Bayer frame to the inputs of ``PyramidProcessor.denoise``.

Python mirror of the kernel wrappers RawConverter::demosaic drives before the
denoise (MetalPipeline/demosaic_kernels.hpp, raw_converter.cpp:148-214) and
of the despeckle RawConverter::denoise runs first (raw_converter.cpp:70-78),
over the HIP kernels of ``libgcam_denoise.so`` (C ABI: ``include/gcam/front_end.h``).

Images are device tensors: raw (H, W) int16/uint16 holding 16-bit values,
single-channel images (H, W) float32, multi-channel (H, W, C) float32.
``RawConverter`` chains the front end, the pyramid denoise and the return to
camera RGB, like RawConverter::demosaic with noise reduction and without the
post-processing.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from ._lib import (GCAM_PIXEL_R16U, GCAM_PIXEL_R32F, GCAM_PIXEL_RG32F, GCAM_PIXEL_RGBA32F, FrontEndParamsC, Image,
                   NlfC, check, lib)
from .pyramid import DenoiseParameters, PyramidProcessor, _stream

# BayerPattern (MetalPipeline/demosaic.hpp; demosaic.metal:24-29)
BAYER = {"grbg": 0, "gbrg": 1, "rggb": 2, "bggr": 3}

_FMT_BY_CH = {1: GCAM_PIXEL_R32F, 2: GCAM_PIXEL_RG32F, 4: GCAM_PIXEL_RGBA32F}


def _bayer(p) -> int:
    v = BAYER[p] if isinstance(p, str) else int(p)
    if v not in (0, 1, 2, 3):
        raise ValueError(f"bad Bayer pattern {p!r}")
    return v


def _img(t: torch.Tensor) -> Image:
    """gcam_image view of a packed device tensor: raw (H, W) 16-bit, (H, W) float32 or (H, W, C) float32."""
    if not t.is_cuda:
        raise ValueError("gcam images live in device memory (a HIP tensor)")
    if t.dtype in (torch.int16, torch.uint16):
        if t.dim() != 2 or t.stride(1) != 1:
            raise ValueError("raw frames are (H, W) with contiguous rows")
        return Image(t.data_ptr(), t.shape[1], t.shape[0], t.stride(0), GCAM_PIXEL_R16U)
    if t.dtype != torch.float32:
        raise ValueError("expected float32 images")
    if t.dim() == 2:
        if t.stride(1) != 1:
            raise ValueError("rows must be contiguous")
        return Image(t.data_ptr(), t.shape[1], t.shape[0], t.stride(0), GCAM_PIXEL_R32F)
    ch = t.shape[2]
    if t.dim() != 3 or ch not in _FMT_BY_CH or t.stride(2) != 1 or t.stride(1) != ch or t.stride(0) % ch:
        raise ValueError("expected packed (H, W, 1|2|4) float32 pixels")
    return Image(t.data_ptr(), t.shape[1], t.shape[0], t.stride(0) // ch, _FMT_BY_CH[ch])


def _floats(v, n) -> ctypes.Array:
    a = [float(x) for x in np.asarray(v, dtype=np.float64).reshape(-1)]
    if len(a) != n:
        raise ValueError(f"expected {n} values, got {len(a)}")
    return (ctypes.c_float * n)(*a)


def _like(t: torch.Tensor, *shape) -> torch.Tensor:
    return torch.empty(shape, dtype=torch.float32, device=t.device)


def scaleRawData(rawImage: torch.Tensor, bayerPattern=2, scaleMul=(1.0, 1.0, 1.0, 1.0), blackLevel: float = 0.0,
                 lensShadingCorrection: float = 0.0, out: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
    """scaleRawDataKernel (demosaic_kernels.hpp:30-49); blackLevel normalised (black_level / 0xffff)."""
    h, w = rawImage.shape
    out = _like(rawImage, h, w) if out is None else out
    check(lib().gcam_scale_raw_data(ctypes.byref(_img(rawImage)), ctypes.byref(_img(out)), _bayer(bayerPattern),
                                    _floats(scaleMul, 4), float(blackLevel), float(lensShadingCorrection),
                                    ctypes.c_void_p(_stream(stream))))
    return out


def rawImageSobel(scaledRawImage: torch.Tensor, out: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
    """rawImageSobelKernel (demosaic_kernels.hpp:51-64): (gx, gy, |gx|, |gy|)."""
    h, w = scaledRawImage.shape
    out = _like(scaledRawImage, h, w, 4) if out is None else out
    check(lib().gcam_raw_image_sobel(ctypes.byref(_img(scaledRawImage)), ctypes.byref(_img(out)),
                                     ctypes.c_void_p(_stream(stream))))
    return out


def gaussianKernelBilinearWeights(radius: float) -> np.ndarray:
    """gaussianKernelBilinearWeights (demosaic_utils.cpp:987-1006): (n, 3) taps {weight, x, y}."""
    n = lib().gcam_gaussian_kernel_bilinear_weights(float(radius), None, 0)
    if n < 0:
        raise ValueError(f"bad radius {radius}")
    buf = (ctypes.c_float * (3 * n))()
    lib().gcam_gaussian_kernel_bilinear_weights(float(radius), buf, n)
    return np.array(buf, dtype=np.float32).reshape(n, 3)


def gaussianBlurSobelImage(rawImage: torch.Tensor, sobelImage: Optional[torch.Tensor], rawNoiseModel,
                           radius1: float = 1.5, radius2: float = 4.5, out: Optional[torch.Tensor] = None,
                           stream=None) -> torch.Tensor:
    """gaussianBlurSobelImageKernel (demosaic_kernels.hpp:66-97).  rawNoiseModel = (a, b) of the
    green raw channel; sobelImage None computes the Sobel filter inside the blur."""
    h, w = rawImage.shape
    out = _like(rawImage, h, w, 2) if out is None else out
    sob = ctypes.byref(_img(sobelImage)) if sobelImage is not None else None
    check(lib().gcam_gaussian_blur_sobel_image(ctypes.byref(_img(rawImage)), sob, float(radius1), float(radius2),
                                               _floats(rawNoiseModel, 2), ctypes.byref(_img(out)),
                                               ctypes.c_void_p(_stream(stream))))
    return out


def getRawVariance(rawNlf) -> np.ndarray:
    """getRawVariance (demosaic.hpp:186-193): rawNlf = (a[4], b[4]) -> [[red a, b], [green a, b], [blue a, b]]."""
    out = (ctypes.c_float * 6)()
    lib().gcam_raw_variance(_floats(rawNlf[0], 4), _floats(rawNlf[1], 4), out)
    return np.array(out, dtype=np.float32).reshape(3, 2)


def demosaicImage(rawImage: torch.Tensor, gradientImage: torch.Tensor, bayerPattern, rawVariance,
                  stream=None) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """demosaicImageKernel (demosaic_kernels.hpp:124-181): returns (green, red/blue pass, linear RGB)."""
    h, w = rawImage.shape
    green, tmp, rgb = _like(rawImage, h, w), _like(rawImage, h, w, 4), _like(rawImage, h, w, 4)
    check(lib().gcam_demosaic_image(ctypes.byref(_img(rawImage)), ctypes.byref(_img(gradientImage)),
                                    ctypes.byref(_img(green)), ctypes.byref(_img(tmp)), ctypes.byref(_img(rgb)),
                                    _bayer(bayerPattern), _floats(rawVariance, 6), ctypes.c_void_p(_stream(stream))))
    return green, tmp, rgb


def blendHighlightsImage(inputImage: torch.Tensor, clip: float = 1.0, out: Optional[torch.Tensor] = None,
                         stream=None) -> torch.Tensor:
    """blendHighlightsImageKernel (demosaic_kernels.hpp:255-269)."""
    out = torch.empty_like(inputImage) if out is None else out
    check(lib().gcam_blend_highlights_image(ctypes.byref(_img(inputImage)), float(clip), ctypes.byref(_img(out)),
                                            ctypes.c_void_p(_stream(stream))))
    return out


def transformImage(inputImage: torch.Tensor, transform, out: Optional[torch.Tensor] = None,
                   stream=None) -> torch.Tensor:
    """transformImageKernel (demosaic_kernels.hpp:271-295): out.xyz = transform @ in.xyz."""
    out = torch.empty_like(inputImage) if out is None else out
    check(lib().gcam_transform_image(ctypes.byref(_img(inputImage)), _floats(transform, 9), ctypes.byref(_img(out)),
                                     ctypes.c_void_p(_stream(stream))))
    return out


def despeckleImage(inputImage: torch.Tensor, var_a, var_b, out: Optional[torch.Tensor] = None,
                   stream=None) -> torch.Tensor:
    """despeckleImageKernel (demosaic_kernels.hpp:297-315)."""
    out = torch.empty_like(inputImage) if out is None else out
    check(lib().gcam_despeckle_image(ctypes.byref(_img(inputImage)), _floats(var_a, 3), _floats(var_b, 3),
                                     ctypes.byref(_img(out)), ctypes.c_void_p(_stream(stream))))
    return out


def rawNoiseStatistics(scaledRawImage: torch.Tensor, bayerPattern=2,
                       stream=None) -> Tuple[torch.Tensor, torch.Tensor]:
    """basicRawNoiseStatistics (demosaic.metal:1110-1136): per-quad (mean, variance) images, (H/2, W/2, 4)."""
    h, w = scaledRawImage.shape
    mean, var = _like(scaledRawImage, h // 2, w // 2, 4), _like(scaledRawImage, h // 2, w // 2, 4)
    check(lib().gcam_raw_noise_statistics(ctypes.byref(_img(scaledRawImage)), _bayer(bayerPattern),
                                          ctypes.byref(_img(mean)), ctypes.byref(_img(var)),
                                          ctypes.c_void_p(_stream(stream))))
    return mean, var


def measureRawNLF(scaledRawImage: torch.Tensor, bayerPattern=2, exposure_multiplier: float = 1.0,
                  stream=None) -> Tuple[Tuple[float, ...], Tuple[float, ...]]:
    """MeasureRawNLF (raw_converter.cpp:339-490): the RawNLF (a[4], b[4]); synchronizes."""
    a, b = (ctypes.c_float * 4)(), (ctypes.c_float * 4)()
    check(lib().gcam_measure_raw_nlf(ctypes.byref(_img(scaledRawImage)), _bayer(bayerPattern),
                                     float(exposure_multiplier), a, b, ctypes.c_void_p(_stream(stream))))
    return tuple(a), tuple(b)


HISTOGRAM_WORDS = 256 + 8 + 6  # histogramImageKernel::histogram_data: 256 + 8 uint32, 6 float


def _histogram_dict(words: torch.Tensor) -> dict:
    host = words.cpu().numpy()
    f = host[264:].view(np.float32)
    keys = ("black_level", "white_level", "shadows", "highlights", "mean", "median")
    return dict(histogram=host[:256].view(np.uint32).copy(), bands=host[256:264].view(np.uint32).copy(),
                **{k: float(v) for k, v in zip(keys, f)})


def _histogram_device(image: torch.Tensor, stream=None) -> torch.Tensor:
    """histogram_data of `image` in device memory (reset, histogramImage, histogramStatistics)."""
    words = torch.empty(HISTOGRAM_WORDS, dtype=torch.int32, device=image.device)
    ptr = ctypes.c_void_p(words.data_ptr())
    s = ctypes.c_void_p(_stream(stream))
    check(lib().gcam_histogram_reset(ptr, s))
    check(lib().gcam_histogram_image(ctypes.byref(_img(image)), ptr, s))
    check(lib().gcam_histogram_statistics(ptr, image.shape[1], image.shape[0], s))
    return words


def histogramImage(image: torch.Tensor, stream=None) -> dict:
    """histogramImageKernel (demosaic_kernels.hpp:485-535): reset, histogramImage and
    histogramStatistics on the device; returns histogram_data as a dict (synchronizes)."""
    return _histogram_dict(_histogram_device(image, stream))


@dataclass
class LTMParameters:
    """LTMParameters (demosaic.hpp:95-100)."""
    eps: float = 0.01
    shadows: float = 0.8
    highlights: float = 1.05
    detail: Sequence[float] = (1.1, 1.2, 1.3)

    def to_c(self) -> "_lib.LtmParamsC":
        return _lib.LtmParamsC(float(self.eps), float(self.shadows), float(self.highlights), _floats(self.detail, 3))


def guidedFilterABImage(guideImage: torch.Tensor, eps: float, out: Optional[torch.Tensor] = None,
                        stream=None) -> torch.Tensor:
    """GuidedFilterABImage (demosaic.metal:1295-1323): (H, W, 2) of (a, b)."""
    h, w = guideImage.shape[:2]
    out = _like(guideImage, h, w, 2) if out is None else out
    check(lib().gcam_guided_filter_ab_image(ctypes.byref(_img(guideImage)), float(eps), ctypes.byref(_img(out)),
                                            ctypes.c_void_p(_stream(stream))))
    return out


def boxFilterGFImage(inputImage: torch.Tensor, out: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
    """BoxFilterGFImage (demosaic.metal:1325-1355): 5x5 box mean of an (H, W, 2) image."""
    out = torch.empty_like(inputImage) if out is None else out
    check(lib().gcam_box_filter_gf_image(ctypes.byref(_img(inputImage)), ctypes.byref(_img(out)),
                                         ctypes.c_void_p(_stream(stream))))
    return out


class LocalToneMapping:
    """LocalToneMapping (raw_converter.hpp:25-78): createMask from the denoised image, its
    gradient, the guides denoisedImagePyramid[4], [2], [0], the level-0 luma noise model and the
    histogram of denoisedImagePyramid[3]; getMask() returns the (H, W) multiplier mask."""

    def __init__(self, width: int, height: int, device=None):
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        sizes = [(height // 16, width // 16), (height // 4, width // 4), (height, width)]
        self.ab = [torch.empty((h, w, 2), dtype=torch.float32, device=dev) for h, w in sizes]
        self.ab_mean = [torch.empty_like(t) for t in self.ab]
        self.mask = torch.empty((height, width), dtype=torch.float32, device=dev)

    def createMask(self, image: torch.Tensor, gradientImage: torch.Tensor, guideImage: Sequence[torch.Tensor],
                   nlf: Tuple[float, float], ltmParameters: LTMParameters, histogram: torch.Tensor,
                   stream=None) -> torch.Tensor:
        for g, ab, m in zip(guideImage, self.ab, self.ab_mean):
            if tuple(g.shape[:2]) != tuple(ab.shape[:2]):
                raise ValueError(f"guide {tuple(g.shape[:2])} does not match its band {tuple(ab.shape[:2])}")
            guidedFilterABImage(g, ltmParameters.eps, out=ab, stream=stream)
            boxFilterGFImage(ab, out=m, stream=stream)
        pc = ltmParameters.to_c()
        check(lib().gcam_local_tone_mapping_mask_image(
            ctypes.byref(_img(image)), ctypes.byref(_img(gradientImage)), ctypes.byref(_img(self.ab_mean[0])),
            ctypes.byref(_img(self.ab_mean[1])), ctypes.byref(_img(self.ab_mean[2])), ctypes.byref(pc),
            float(nlf[0]), float(nlf[1]), ctypes.c_void_p(histogram.data_ptr()), ctypes.byref(_img(self.mask)),
            ctypes.c_void_p(_stream(stream))))
        return self.mask

    def getMask(self) -> torch.Tensor:
        return self.mask


def cam_ycbcr(rgb_cam=np.eye(3)) -> np.ndarray:
    """cam_ycbcr(rgb_cam, xyz_sRGB) (demosaic_utils.cpp:703-720)."""
    out = (ctypes.c_float * 9)()
    lib().gcam_cam_ycbcr(_floats(rgb_cam, 9), out)
    return np.array(out, dtype=np.float32).reshape(3, 3)


def inverse3(m) -> np.ndarray:
    """inverse() of a 3x3 matrix (raw_converter.cpp:169)."""
    out = (ctypes.c_float * 9)()
    check(lib().gcam_matrix3_inverse(_floats(m, 9), out))
    return np.array(out, dtype=np.float32).reshape(3, 3)


@dataclass
class FrontEndParameters:
    """The fields of DemosaicParameters / NoiseModel (demosaic.hpp:52-60, 102-124) the front end reads."""
    bayerPattern: int = 2
    black_level: float = 0.0            # normalised: DemosaicParameters.black_level / 0xffff
    lensShadingCorrection: float = 0.0
    scale_mul: Sequence[float] = (1.0, 1.0, 1.0, 1.0)
    rawNlf: Tuple[Sequence[float], Sequence[float]] = ((1e-5,) * 4, (1e-3,) * 4)
    rgb_cam: Sequence[Sequence[float]] = field(default_factory=lambda: np.eye(3).tolist())
    highlightClip: float = 1.0
    despeckleNlf: Optional[Tuple[Sequence[float], Sequence[float]]] = None  # pyramidNlf[0]; None: no despeckle
    calibrateRawNlf: bool = False       # fit rawNlf from the frame (MeasureRawNLF); the run then synchronizes
    exposure_multiplier: float = 1.0

    def to_c(self) -> FrontEndParamsC:
        c = FrontEndParamsC()
        c.bayer_pattern = _bayer(self.bayerPattern)
        c.black_level = float(self.black_level)
        c.lens_shading = float(self.lensShadingCorrection)
        c.scale_mul = _floats(self.scale_mul, 4)
        c.raw_nlf_a = _floats(self.rawNlf[0], 4)
        c.raw_nlf_b = _floats(self.rawNlf[1], 4)
        c.rgb_cam = _floats(self.rgb_cam, 9)
        c.highlight_clip = float(self.highlightClip)
        c.despeckle = 1 if self.despeckleNlf is not None else 0
        if self.despeckleNlf is not None:
            c.despeckle_nlf = NlfC(_floats(self.despeckleNlf[0], 3), _floats(self.despeckleNlf[1], 3))
        c.calibrate_raw_nlf = 1 if self.calibrateRawNlf else 0
        c.exposure_multiplier = float(self.exposure_multiplier)
        return c


class RawFrontEnd:
    """The front end of RawConverter::demosaic for one frame size: raw -> (YCbCr, gradient)."""

    def __init__(self, width: int, height: int, device=None):
        self.width, self.height = int(width), int(height)
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(lib().gcam_front_end_create(self.width, self.height, ctypes.byref(h)))
        self._h = h

    def run(self, rawImage: torch.Tensor, params: FrontEndParameters, ycbcr: Optional[torch.Tensor] = None,
            gradient: Optional[torch.Tensor] = None, stream=None) -> Tuple[torch.Tensor, torch.Tensor]:
        if self._h is None:
            raise RuntimeError("RawFrontEnd is closed")
        shape = (self.height, self.width)
        if tuple(rawImage.shape) != shape:
            raise ValueError(f"raw frame is {tuple(rawImage.shape)}, the front end was built for {shape}")
        ycbcr = torch.empty((*shape, 4), dtype=torch.float32, device=self.device) if ycbcr is None else ycbcr
        gradient = torch.empty((*shape, 2), dtype=torch.float32, device=self.device) if gradient is None else gradient
        pc = params.to_c()
        check(lib().gcam_front_end_run(self._h, ctypes.byref(_img(rawImage)), ctypes.byref(pc),
                                       ctypes.byref(_img(ycbcr)), ctypes.byref(_img(gradient)),
                                       ctypes.c_void_p(_stream(stream))))
        return ycbcr, gradient

    @property
    def rawNlf(self) -> Tuple[Tuple[float, ...], Tuple[float, ...]]:
        """The raw noise model of the last run (fitted when it calibrated)."""
        a, b = (ctypes.c_float * 4)(), (ctypes.c_float * 4)()
        check(lib().gcam_front_end_raw_nlf(self._h, a, b))
        return tuple(a), tuple(b)

    def close(self) -> None:
        if self._h is not None:
            lib().gcam_front_end_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RawConverter:
    """RawConverter::demosaic with noise reduction, without post-processing
    (raw_converter.cpp:148-223): front end, despeckle, pyramid denoise with the
    front end's gradient, and the return to camera RGB (inverse of cam_ycbcr)."""

    def __init__(self, width: int, height: int, levels: int = 5, search_radius: int = 4, precision: str = "fp16",
                 device=None):
        self.front = RawFrontEnd(width, height, device)
        self.pyramid = PyramidProcessor(width, height, levels=levels, search_radius=search_radius,
                                        precision=precision, device=device)
        self.levels = levels
        self.ltm: Optional[LocalToneMapping] = None
        self.histogramData: Optional[dict] = None

    def demosaic(self, rawImage: torch.Tensor, params: FrontEndParameters,
                 denoiseParameters: Sequence[DenoiseParameters], pyramidNlf, exposure_multiplier: float = 1.0,
                 calibrateFromImage: bool = False, ltmParameters: Optional[LTMParameters] = None,
                 stream=None) -> torch.Tensor:
        """Returns linear camera RGB (H, W, 4); pyramidNlf is updated in place when calibrating.
        histogramData holds the statistics of denoised level 3 and, with ltmParameters, ltm.getMask()
        the local tone-mapping mask."""
        if params.despeckleNlf is None:
            params = FrontEndParameters(**{**params.__dict__, "despeckleNlf": pyramidNlf[0]})
        ycc, grad = self.front.run(rawImage, params, stream=stream)
        den = self.pyramid.denoise(denoiseParameters, ycc, grad, pyramidNlf, exposure_multiplier,
                                   params.lensShadingCorrection, calibrateFromImage, stream=stream)
        # RawConverter::denoise's histogram of the denoised level 3 (raw_converter.cpp:87-89)
        hist_level = min(3, self.levels - 1)
        hist = _histogram_device(self.pyramid.denoisedImagePyramid[hist_level], stream)
        if ltmParameters is not None:  # the local tone-mapping mask (raw_converter.cpp:104-112)
            if self.levels != 5:
                raise ValueError("the local tone-mapping mask uses pyramid levels 4, 2 and 0 (5 levels)")
            if self.ltm is None:
                self.ltm = LocalToneMapping(self.front.width, self.front.height, self.front.device)
            d = self.pyramid.denoisedImagePyramid
            self.ltm.createMask(den, grad, [d[4], d[2], d[0]], (pyramidNlf[0][0][0], pyramidNlf[0][1][0]),
                                ltmParameters, hist, stream=stream)
        self.histogramData = _histogram_dict(hist)
        back = inverse3(cam_ycbcr(np.asarray(params.rgb_cam, np.float32).reshape(3, 3)))
        return transformImage(den, back, stream=stream)

    def close(self) -> None:
        self.front.close()
        self.pyramid.close()
        self.ltm = None
