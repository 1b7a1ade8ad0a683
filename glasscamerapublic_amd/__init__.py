"""MI355X-native pyramid PCA block-matching denoise (GlassCamera's
PyramidProcessor<levels>::denoise path) over hand-written gfx950 HIP kernels.

The compute lives in ``libgcam_denoise.so`` (C ABI: ``include/gcam/denoise.h``
and, for the raw front end that produces the denoise's inputs,
``include/gcam/front_end.h``); this package is the Python mirror of the
reference's operator surface.
"""
from ._lib import GcamError, lib  # noqa: F401
from .pyramid import (DenoiseParameters, PyramidProcessor, basic_noise_statistics,  # noqa: F401
                      block_matching_denoise_image, block_matching_denoise_image_lp, build_pca_space,
                      collect_patches, denoise_image, downsample_image, measure_ycbcr_nlf, pca_projection,
                      subtract_noise_image, synth_frame)
from .burst import BurstProcessor  # noqa: F401
from .front_end import (FrontEndParameters, RawConverter, RawFrontEnd, blendHighlightsImage, cam_ycbcr,  # noqa: F401
                        demosaicImage, despeckleImage, gaussianBlurSobelImage, gaussianKernelBilinearWeights,
                        getRawVariance, histogramImage, inverse3, LocalToneMapping, LTMParameters,
                        guidedFilterABImage, boxFilterGFImage, measureRawNLF, rawImageSobel, rawNoiseStatistics, scaleRawData,
                        transformImage)
from .registration import (FindHomography, FindLeastSquaresHomography, ScaleHomography, fuseBurst,  # noqa: F401
                           ransacSamples, registerAndFuse, registerBayerImage, registerImage)

__all__ = [
    "GcamError", "lib", "DenoiseParameters", "PyramidProcessor", "BurstProcessor", "basic_noise_statistics",
    "block_matching_denoise_image", "block_matching_denoise_image_lp", "build_pca_space", "collect_patches",
    "denoise_image", "downsample_image", "measure_ycbcr_nlf", "pca_projection", "subtract_noise_image",
    "synth_frame", "FrontEndParameters", "RawConverter", "RawFrontEnd", "blendHighlightsImage", "cam_ycbcr",
    "demosaicImage", "despeckleImage", "gaussianBlurSobelImage", "gaussianKernelBilinearWeights", "getRawVariance",
    "histogramImage", "inverse3", "LocalToneMapping", "LTMParameters", "guidedFilterABImage", "boxFilterGFImage", "measureRawNLF", "rawImageSobel", "rawNoiseStatistics", "scaleRawData", "transformImage",
    "FindHomography", "FindLeastSquaresHomography", "ScaleHomography", "fuseBurst", "ransacSamples",
    "registerAndFuse", "registerBayerImage", "registerImage",
]
