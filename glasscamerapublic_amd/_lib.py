"""ctypes binding of the C ABI declared in ``include/gcam/denoise.h`` and
``include/gcam/front_end.h``.

The product path is the HIP library ``libgcam_denoise.so`` built in-tree by
``glasscamerapublic_amd.build``.  There is no fallback: if the library is
missing, ``lib()`` raises and every operation fails loudly.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgcam_denoise.so")

GCAM_OK = 0
GCAM_ERR_INVALID_ARGUMENT = 1
GCAM_PIXEL_RGBA32F = 0
GCAM_PIXEL_RG32F = 1
GCAM_PIXEL_R32F = 2
GCAM_PIXEL_PCA8_F32 = 3
GCAM_PIXEL_PCA8_F16 = 4
GCAM_PIXEL_R16U = 5

GCAM_PRECISION = {"fp32": 0, "fp16": 1, "bf16": 2}
GCAM_LEVEL_DENOISED, GCAM_LEVEL_SUBTRACTED, GCAM_LEVEL_NOISY, GCAM_LEVEL_FEATURES = 0, 1, 2, 3
STAGES = ("downsample", "nlf", "subtract", "pca", "block_match", "moments")

# Every symbol the header declares (checked by the CPU test suite).
EXPORTED = (
    "gcam_abi_version", "gcam_status_string", "gcam_last_error",
    "gcam_pyramid_create", "gcam_pyramid_destroy", "gcam_pyramid_denoise", "gcam_pyramid_level",
    "gcam_pyramid_basis", "gcam_pyramid_set_profiling", "gcam_pyramid_stage_times",
    "gcam_pyramid_denoise_fwd", "gcam_pyramid_denoise_bwd", "gcam_downsample_image", "gcam_subtract_noise_image", "gcam_collect_patches",
    "gcam_build_pca_space", "gcam_build_pca_space_noalloc", "gcam_pca_projection", "gcam_block_matching_denoise_image",
    "gcam_block_matching_denoise_image_lp", "gcam_denoise_image",
    "gcam_basic_noise_statistics", "gcam_measure_ycbcr_nlf", "gcam_synth_frame",
    "gcam_burst_create", "gcam_burst_destroy", "gcam_burst_denoise", "gcam_burst_output", "gcam_burst_tune",
    "gcam_burst_get_tuning", "gcam_burst_set_tuning",
    "gcam_scale_raw_data", "gcam_raw_image_sobel", "gcam_gaussian_kernel_bilinear_weights",
    "gcam_gaussian_blur_sobel_image", "gcam_raw_variance", "gcam_demosaic_image", "gcam_blend_highlights_image",
    "gcam_transform_image", "gcam_despeckle_image", "gcam_cam_ycbcr", "gcam_matrix3_inverse",
    "gcam_front_end_create", "gcam_front_end_destroy", "gcam_front_end_run",
    "gcam_raw_noise_statistics", "gcam_measure_raw_nlf", "gcam_front_end_raw_nlf",
    "gcam_histogram_reset", "gcam_histogram_image", "gcam_histogram_statistics",
    "gcam_guided_filter_ab_image", "gcam_box_filter_gf_image", "gcam_local_tone_mapping_mask_image",
    "gcam_register_and_fuse_image", "gcam_register_image", "gcam_register_bayer_image",
    "gcam_find_least_squares_homography", "gcam_find_homography", "gcam_ransac_samples", "gcam_scale_homography",
)
GCAM_MAX_LEVELS = 8
GCAM_MAX_BURST = 64


class GcamError(RuntimeError):
    """A gcam_status other than GCAM_OK, with gcam_last_error()'s message."""

    def __init__(self, status: int, message: str):
        super().__init__(f"gcam status {status}: {message}")
        self.status = status


class Image(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("stride", ctypes.c_int32), ("format", ctypes.c_int32)]


class DenoiseParamsC(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in
                ("luma", "chroma", "chromaBoost", "gradientBoost", "gradientThreshold", "sharpening")]


class NlfC(ctypes.Structure):
    _fields_ = [("a", ctypes.c_float * 3), ("b", ctypes.c_float * 3)]


class ConfigC(ctypes.Structure):
    _fields_ = [("search_radius", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("bm_rows_per_thread", ctypes.c_int32), ("bilateral", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 4)]


class SynthParamsC(ctypes.Structure):
    _fields_ = [("scene_seed", ctypes.c_uint64), ("noise_seed", ctypes.c_uint64), ("shot_gain", ctypes.c_float),
                ("read_sigma", ctypes.c_float), ("lo", ctypes.c_float), ("hi", ctypes.c_float),
                ("period", ctypes.c_float), ("octaves", ctypes.c_int32)]


class BurstTuningC(ctypes.Structure):
    _fields_ = [("rows_per_thread", ctypes.c_int32 * 8), ("grid", ctypes.c_int32 * 8),
                ("schedule", ctypes.c_int32 * 8), ("tile_width", ctypes.c_int32 * 8), ("level_ms", ctypes.c_float * 8),
                ("tuned", ctypes.c_int32)]


class LtmParamsC(ctypes.Structure):
    _fields_ = [("eps", ctypes.c_float), ("shadows", ctypes.c_float), ("highlights", ctypes.c_float),
                ("detail", ctypes.c_float * 3)]


class FrontEndParamsC(ctypes.Structure):
    _fields_ = [("bayer_pattern", ctypes.c_int32), ("black_level", ctypes.c_float), ("lens_shading", ctypes.c_float),
                ("scale_mul", ctypes.c_float * 4), ("raw_nlf_a", ctypes.c_float * 4),
                ("raw_nlf_b", ctypes.c_float * 4), ("rgb_cam", ctypes.c_float * 9),
                ("highlight_clip", ctypes.c_float), ("despeckle", ctypes.c_int32), ("despeckle_nlf", NlfC),
                ("calibrate_raw_nlf", ctypes.c_int32), ("exposure_multiplier", ctypes.c_float)]


_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int32
_F = ctypes.c_float
_IP = ctypes.POINTER(Image)


def _declare(L: ctypes.CDLL) -> None:
    L.gcam_abi_version.restype = _I
    L.gcam_status_string.argtypes = [_I]
    L.gcam_status_string.restype = ctypes.c_char_p
    L.gcam_last_error.restype = ctypes.c_char_p
    L.gcam_pyramid_create.argtypes = [_I, _I, _I, ctypes.POINTER(ConfigC), ctypes.POINTER(_P)]
    L.gcam_pyramid_destroy.argtypes = [_P]
    L.gcam_pyramid_destroy.restype = None
    L.gcam_pyramid_denoise.argtypes = [_P, ctypes.POINTER(DenoiseParamsC), _IP, _IP, ctypes.POINTER(NlfC), _F, _F, _I,
                                       _P, _IP]
    L.gcam_pyramid_level.argtypes = [_P, _I, _I, _IP]
    L.gcam_pyramid_basis.argtypes = [_P, _I, _P, _P]
    L.gcam_pyramid_denoise_fwd.argtypes = [_P, ctypes.POINTER(DenoiseParamsC), _P, _IP, ctypes.POINTER(NlfC), _P,
                                           _IP]
    L.gcam_pyramid_denoise_bwd.argtypes = [_P, _IP, _P, _P]
    L.gcam_pyramid_set_profiling.argtypes = [_P, _I]
    L.gcam_pyramid_stage_times.argtypes = [_P, _P]
    L.gcam_downsample_image.argtypes = [_IP, _IP, _P]
    L.gcam_subtract_noise_image.argtypes = [_IP, _IP, _IP, _IP, _F, _F, _F, _F, _IP, _P]
    L.gcam_collect_patches.argtypes = [_IP, _P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), _P]
    L.gcam_build_pca_space.argtypes = [_P, ctypes.c_int64, _P, _P, _P]
    L.gcam_build_pca_space_noalloc.argtypes = [_P, ctypes.c_int64, _P, _P, _P]
    L.gcam_pca_projection.argtypes = [_IP, _P, _IP, _P]
    L.gcam_block_matching_denoise_image.argtypes = [_IP, _IP, _IP, _P, _P, _P, _F, _F, _F, _F, _I, _IP, _P]
    L.gcam_block_matching_denoise_image_lp.argtypes = [_IP, _IP, _P, _P, _P, _P, _F, _F, _F, _F, _I, _I, _IP, _P]
    L.gcam_denoise_image.argtypes = [_IP, _IP, _P, _P, _P, _F, _F, _F, _IP, _P]
    L.gcam_basic_noise_statistics.argtypes = [_IP, _IP, _P]
    L.gcam_measure_ycbcr_nlf.argtypes = [_IP, _IP, _F, ctypes.POINTER(NlfC), _P]
    L.gcam_synth_frame.argtypes = [_I, _I, ctypes.POINTER(SynthParamsC), _P, _IP, _IP, _P]
    L.gcam_burst_create.argtypes = [_I, ctypes.POINTER(_I), ctypes.POINTER(_I), _I, ctypes.POINTER(ConfigC),
                                    ctypes.POINTER(_P)]
    L.gcam_burst_destroy.argtypes = [_P]
    L.gcam_burst_destroy.restype = None
    L.gcam_burst_denoise.argtypes = [_P, ctypes.POINTER(DenoiseParamsC), _IP, _IP, ctypes.POINTER(NlfC), _F, _P]
    L.gcam_burst_tune.argtypes = [_P, ctypes.POINTER(DenoiseParamsC), _IP, _IP, ctypes.POINTER(NlfC), _F, _P]
    L.gcam_burst_output.argtypes = [_P, _I, _I, _IP]
    L.gcam_burst_get_tuning.argtypes = [_P, ctypes.POINTER(BurstTuningC)]
    L.gcam_burst_set_tuning.argtypes = [_P, ctypes.POINTER(BurstTuningC)]
    _FP = ctypes.POINTER(ctypes.c_float)
    L.gcam_scale_raw_data.argtypes = [_IP, _IP, _I, _FP, _F, _F, _P]
    L.gcam_raw_image_sobel.argtypes = [_IP, _IP, _P]
    L.gcam_gaussian_kernel_bilinear_weights.argtypes = [_F, _FP, _I]
    L.gcam_gaussian_blur_sobel_image.argtypes = [_IP, _IP, _F, _F, _FP, _IP, _P]
    L.gcam_raw_variance.argtypes = [_FP, _FP, _FP]
    L.gcam_raw_variance.restype = None
    L.gcam_demosaic_image.argtypes = [_IP, _IP, _IP, _IP, _IP, _I, _FP, _P]
    L.gcam_blend_highlights_image.argtypes = [_IP, _F, _IP, _P]
    L.gcam_transform_image.argtypes = [_IP, _FP, _IP, _P]
    L.gcam_despeckle_image.argtypes = [_IP, _FP, _FP, _IP, _P]
    L.gcam_cam_ycbcr.argtypes = [_FP, _FP]
    L.gcam_cam_ycbcr.restype = None
    L.gcam_matrix3_inverse.argtypes = [_FP, _FP]
    L.gcam_front_end_create.argtypes = [_I, _I, ctypes.POINTER(_P)]
    L.gcam_front_end_destroy.argtypes = [_P]
    L.gcam_front_end_destroy.restype = None
    L.gcam_front_end_run.argtypes = [_P, _IP, ctypes.POINTER(FrontEndParamsC), _IP, _IP, _P]
    L.gcam_raw_noise_statistics.argtypes = [_IP, _I, _IP, _IP, _P]
    L.gcam_measure_raw_nlf.argtypes = [_IP, _I, _F, _FP, _FP, _P]
    L.gcam_front_end_raw_nlf.argtypes = [_P, _FP, _FP]
    L.gcam_histogram_reset.argtypes = [_P, _P]
    L.gcam_histogram_image.argtypes = [_IP, _P, _P]
    L.gcam_histogram_statistics.argtypes = [_P, _I, _I, _P]
    L.gcam_guided_filter_ab_image.argtypes = [_IP, _F, _IP, _P]
    L.gcam_box_filter_gf_image.argtypes = [_IP, _IP, _P]
    L.gcam_local_tone_mapping_mask_image.argtypes = [_IP, _IP, _IP, _IP, _IP, ctypes.POINTER(LtmParamsC), _F, _F,
                                                     _P, _IP, _P]
    L.gcam_register_and_fuse_image.argtypes = [_IP, _IP, _IP, _FP, _I, _P]
    L.gcam_register_image.argtypes = [_IP, _IP, _FP, _P]
    L.gcam_register_bayer_image.argtypes = [_IP, _IP, _FP, _P]
    L.gcam_find_least_squares_homography.argtypes = [_P, _I, _FP, _P]
    L.gcam_find_homography.argtypes = [_P, _I, _F, _I, ctypes.c_uint64, _FP, ctypes.POINTER(ctypes.c_int32),
                                       ctypes.POINTER(ctypes.c_int32), _P]
    L.gcam_ransac_samples.argtypes = [ctypes.c_uint64, _I, _I, _I, ctypes.POINTER(ctypes.c_int32)]
    L.gcam_scale_homography.argtypes = [_FP, _F, _FP]
    L.gcam_scale_homography.restype = None
    for name in EXPORTED:
        fn = getattr(L, name)
        if fn.restype is ctypes.c_int:  # default restype: gcam_status
            fn.restype = _I


def lib() -> ctypes.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m glasscamerapublic_amd.build` "
                "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != GCAM_OK:
        msg = lib().gcam_last_error()
        raise GcamError(status, msg.decode() if msg else "")
