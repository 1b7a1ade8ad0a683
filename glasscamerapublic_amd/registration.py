"""Burst registration and fusion: homography warps and the robust homography fit.

Python mirror of the registration kernels' wrappers
(MetalPipeline/demosaic_kernels.hpp:690-749: RegisterAndFuseKernel,
RegisterImageKernel, RegisterBayerImageKernel), of FindHomography /
FindLeastSquaresHomography / ScaleHomography (ImageRegistration/Homography.hpp,
Homography.cpp:72-116, LeastSquaresHomography.cpp:356-421) and of the burst
loop of PyramidFusion/pyramidFusion.cpp:316-346, over the HIP kernels of
``libgcam_denoise.so`` (C ABI: ``include/gcam/registration.h``).

Homographies are 3 x 3 float32 arrays mapping output (reference) coordinates
to input coordinates.  Matches are (n, 4) arrays of (x1, y1, x2, y2), a fitted
homography mapping (x1, y1) to (x2, y2).  Images are device tensors as in
``front_end``: (H, W, 4) or (H, W) float32, raw frames (H, W) 16-bit.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np
import torch

from ._lib import check, lib
from .front_end import _img
from .pyramid import _stream


def _h9(homography) -> ctypes.Array:
    h = np.asarray(homography, dtype=np.float32).reshape(-1)
    if h.size != 9:
        raise ValueError("a homography is 3 x 3")
    return (ctypes.c_float * 9)(*[float(v) for v in h])


def _matches(matchpoints, device=None) -> torch.Tensor:
    """(n, 4) float32 device tensor of the matches (16-byte aligned rows)."""
    if isinstance(matchpoints, torch.Tensor):
        m = matchpoints.to(dtype=torch.float32)
        if not m.is_cuda:
            m = m.to(device or "cuda")
    else:
        arr = np.asarray(matchpoints, dtype=np.float32)
        m = torch.from_numpy(np.ascontiguousarray(arr)).to(device or "cuda")
    m = m.reshape(-1, 4).contiguous()
    return m


def registerAndFuse(inputImage0: torch.Tensor, inputImage1: torch.Tensor, outputImage: Optional[torch.Tensor],
                    homography, count: int, stream=None) -> torch.Tensor:
    """RegisterAndFuseKernel::operator() (demosaic_kernels.hpp:701-709):
    output = ((count - 1) inputImage0 + inputImage1(H p)) / count over inputImage0's size;
    outputImage may be inputImage0 (the burst driver fuses in place)."""
    out = torch.empty_like(inputImage0) if outputImage is None else outputImage
    check(lib().gcam_register_and_fuse_image(ctypes.byref(_img(inputImage0)), ctypes.byref(_img(inputImage1)),
                                             ctypes.byref(_img(out)), _h9(homography), int(count),
                                             ctypes.c_void_p(_stream(stream))))
    return out


def registerImage(inputImage: torch.Tensor, homography, outputImage: Optional[torch.Tensor] = None,
                  stream=None) -> torch.Tensor:
    """RegisterImageKernel<T>::operator() (demosaic_kernels.hpp:722-729): output = input(H p), bilinear."""
    out = torch.empty_like(inputImage) if outputImage is None else outputImage
    check(lib().gcam_register_image(ctypes.byref(_img(inputImage)), ctypes.byref(_img(out)), _h9(homography),
                                    ctypes.c_void_p(_stream(stream))))
    return out


def registerBayerImage(inputImage: torch.Tensor, homography, outputImage: Optional[torch.Tensor] = None,
                       stream=None) -> torch.Tensor:
    """RegisterBayerImageKernel::operator() (demosaic_kernels.hpp:741-748): nearest-neighbour warp
    of each colour plane of a raw frame by a homography of the half-resolution planes."""
    out = torch.empty_like(inputImage) if outputImage is None else outputImage
    check(lib().gcam_register_bayer_image(ctypes.byref(_img(inputImage)), ctypes.byref(_img(out)),
                                          _h9(homography), ctypes.c_void_p(_stream(stream))))
    return out


def FindLeastSquaresHomography(matchpoints, stream=None) -> np.ndarray:
    """FindLeastSquaresHomography (LeastSquaresHomography.cpp:356-421) of all the matches, on the
    device; raises GcamError where the reference throws std::range_error."""
    m = _matches(matchpoints)
    H = (ctypes.c_float * 9)()
    check(lib().gcam_find_least_squares_homography(ctypes.c_void_p(m.data_ptr()), m.shape[0], H,
                                                   ctypes.c_void_p(_stream(stream))))
    return np.array(H[:], dtype=np.float32).reshape(3, 3)


def FindHomography(matchpoints, threshold: float, max_iterations: int,
                   inlier_indices: Optional[List[int]] = None, seed: int = 0, stream=None) -> np.ndarray:
    """FindHomography (Homography.cpp:72-116): RANSAC over max_iterations minimal samples, then the
    least-squares refit of the winner's inliers; fills inlier_indices (ascending) when given."""
    m = _matches(matchpoints)
    n = m.shape[0]
    H = (ctypes.c_float * 9)()
    idx = (ctypes.c_int32 * n)()
    cnt = ctypes.c_int32(0)
    check(lib().gcam_find_homography(ctypes.c_void_p(m.data_ptr()), n, float(threshold), int(max_iterations),
                                     ctypes.c_uint64(int(seed) & (2**64 - 1)), H, idx, ctypes.byref(cnt),
                                     ctypes.c_void_p(_stream(stream))))
    if inlier_indices is not None:
        inlier_indices[:] = list(idx[:cnt.value])
    return np.array(H[:], dtype=np.float32).reshape(3, 3)


def ransacSamples(seed: int, count: int, first: int, n: int) -> np.ndarray:
    """The (n, 4) minimal samples FindHomography draws for hypotheses first .. first + n - 1."""
    out = (ctypes.c_int32 * (4 * n))()
    check(lib().gcam_ransac_samples(ctypes.c_uint64(int(seed) & (2**64 - 1)), int(count), int(first), int(n), out))
    return np.array(out[:], dtype=np.int32).reshape(n, 4)


def ScaleHomography(homography, scale: float) -> np.ndarray:
    """ScaleHomography (Homography.hpp:29-37): S H S^-1 with S = diag(scale, scale, 1)."""
    out = (ctypes.c_float * 9)()
    lib().gcam_scale_homography(_h9(homography), float(scale), out)
    return np.array(out[:], dtype=np.float32).reshape(3, 3)


def fuseBurst(referenceImage: torch.Tensor, images: Sequence[torch.Tensor], matches: Sequence[Sequence],
              threshold: float = 1.0, max_iterations: int = 2000, scale: float = 2.0, count_base: int = 1,
              seed: int = 0, fusedImage: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
    """The burst loop of pyramidFusion.cpp:316-346.  matches[i] holds, per feature channel, the
    matches of the reference against images[i] at 1 / scale resolution; their homographies are
    averaged (the running average of :330-338), scaled to full resolution and image i is fused
    into the copy of the reference with count i + count_base.  The reference passes i + 1, so
    its first registered frame replaces the reference frame; count_base=2 keeps the reference
    in the average."""
    if len(images) != len(matches):
        raise ValueError("one match set per image")
    fused = referenceImage.clone() if fusedImage is None else fusedImage.copy_(referenceImage)
    for i, (image, per_channel) in enumerate(zip(images, matches)):
        H = np.zeros((3, 3), np.float32)
        for c, m in enumerate(per_channel):
            Hc = FindHomography(m, threshold, max_iterations, seed=seed, stream=stream)
            H = (Hc + H * np.float32(c)) / np.float32(c + 1)
        registerAndFuse(fused, image, fused, ScaleHomography(H, scale), i + count_base, stream=stream)
    return fused
