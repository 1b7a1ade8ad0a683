"""Batched form: a burst of frames of uneven sizes denoised in one graph launch.

``BurstProcessor`` holds one pyramid per frame (sizes fixed at construction)
and runs ``PyramidProcessor::denoise`` (MetalPipeline/pyramid_processor.cpp:96-172)
on all of them through ``gcam_burst_denoise``: every stage of a level is one
kernel launch over all frames, the block matching one persistent launch over
every frame's tiles, and the whole call a captured hipGraph.  The block
matching launch shape is autotuned on the first call (or by ``tune``).
"""
from __future__ import annotations

import ctypes
from dataclasses import fields
from typing import Optional, Sequence

import torch

from . import _lib
from ._lib import (GCAM_PIXEL_RG32F, GCAM_PIXEL_RGBA32F, BurstTuningC, ConfigC, DenoiseParamsC, Image, NlfC,
                   check, lib)
from .pyramid import DenoiseParameters, _stream, _wrap, as_image


class BurstProcessor:
    """Many PyramidProcessor<levels> runs in one launch (precision form: "fp16" or "bf16")."""

    def __init__(self, sizes: Sequence[tuple], levels: int = 5, search_radius: int = 4, precision: str = "fp16",
                 device: Optional[torch.device] = None):
        self.sizes = [(int(w), int(h)) for (w, h) in sizes]
        self.n, self.levels = len(self.sizes), int(levels)
        self.search_radius, self.precision = int(search_radius), precision
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        W = (ctypes.c_int32 * self.n)(*[w for w, _ in self.sizes])
        H = (ctypes.c_int32 * self.n)(*[h for _, h in self.sizes])
        cfg = ConfigC(self.search_radius, _lib.GCAM_PRECISION[precision], 0, 0)
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(lib().gcam_burst_create(self.n, W, H, self.levels, ctypes.byref(cfg), ctypes.byref(handle)))
        self._h = handle

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            torch.cuda.synchronize(self.device)
            lib().gcam_burst_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _args(self, params, frames, nlf, gradients):
        """params / nlf: one list of `levels` entries per frame, or one list shared by all frames."""
        if len(frames) != self.n:
            raise ValueError(f"need {self.n} frames")
        L = self.levels
        if len(params) == L and isinstance(params[0], DenoiseParameters):
            params = [params] * self.n
        if len(params) != self.n:
            raise ValueError(f"params: one list of {L} per frame, or one shared list")
        # a shared model is a list of (A, B) pairs, A = 3 numbers
        if len(nlf) == L and isinstance(nlf[0][0][0], (int, float)):
            nlf = [nlf] * self.n
        if len(nlf) != self.n:
            raise ValueError(f"nlf: one list of {L} (A, B) pairs per frame, or one shared list")
        P = (DenoiseParamsC * (self.n * L))()
        N = (NlfC * (self.n * L))()
        for f in range(self.n):
            if len(params[f]) != L or len(nlf[f]) != L:
                raise ValueError(f"frame {f}: need {L} parameters and noise models")
            for i in range(L):
                for fld in fields(DenoiseParameters):
                    setattr(P[f * L + i], fld.name, float(getattr(params[f][i], fld.name)))
                A, B = nlf[f][i]
                for c in range(3):
                    N[f * L + i].a[c] = float(A[c])
                    N[f * L + i].b[c] = float(B[c])
        F = (Image * self.n)(*[as_image(t, GCAM_PIXEL_RGBA32F) for t in frames])
        G = None
        if gradients is not None:
            G = (Image * self.n)(*[as_image(g, GCAM_PIXEL_RG32F) if g is not None else Image() for g in gradients])
        return P, F, G, N

    def denoise(self, params, frames: Sequence[torch.Tensor], nlf, gradients=None,
                lensShadingCorrection: float = 0.0, stream=None):
        """Denoise the burst; returns the level-0 outputs (handle-owned views, valid until the next call)."""
        P, F, G, N = self._args(params, frames, nlf, gradients)
        with torch.cuda.device(self.device):
            check(lib().gcam_burst_denoise(self._h, P, F, G, N, float(lensShadingCorrection),
                                           ctypes.c_void_p(_stream(stream))))
        return [self.output(f, 0) for f in range(self.n)]

    def tune(self, params, frames, nlf, gradients=None, lensShadingCorrection: float = 0.0, stream=None) -> dict:
        """Run the autotuner on these inputs (synchronizes); returns the tuning."""
        P, F, G, N = self._args(params, frames, nlf, gradients)
        with torch.cuda.device(self.device):
            check(lib().gcam_burst_tune(self._h, P, F, G, N, float(lensShadingCorrection),
                                        ctypes.c_void_p(_stream(stream))))
        return self.tuning

    def output(self, frame: int, level: int = 0) -> torch.Tensor:
        img = Image()
        check(lib().gcam_burst_output(self._h, int(frame), int(level), ctypes.byref(img)))
        return _wrap(img, 4, self.device)

    @property
    def tuning(self) -> dict:
        t = BurstTuningC()
        check(lib().gcam_burst_get_tuning(self._h, ctypes.byref(t)))
        L = self.levels
        return {"rows_per_thread": list(t.rows_per_thread)[:L], "grid": list(t.grid)[:L],
                "schedule": list(t.schedule)[:L], "tile_width": list(t.tile_width)[:L],
                "level_ms": list(t.level_ms)[:L], "tuned": bool(t.tuned)}

    def set_tuning(self, rows_per_thread, grid, schedule, tile_width=None) -> None:
        """Fix the block-matching launch per level: rows per thread (1, 2, 4),
        persistent grid (0: one workgroup per tile), tile order (0, 1) and tile
        width (32, or 64 with two rows per thread; default 32)."""
        t = BurstTuningC()
        for i in range(self.levels):
            t.rows_per_thread[i] = int(rows_per_thread[i])
            t.grid[i] = int(grid[i])
            t.schedule[i] = int(schedule[i])
            t.tile_width[i] = 32 if tile_width is None else int(tile_width[i])
        check(lib().gcam_burst_set_tuning(self._h, ctypes.byref(t)))
