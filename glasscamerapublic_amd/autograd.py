"""torch.autograd wrapper of the differentiable denoise (gcam_pyramid_denoise_fwd/_bwd).

``denoise_thresholds(thresholds, images, processors, params, nlf)`` denoises a
batch of crops with shared per-level, per-channel thresholds (levels x 3:
Y, Cb, Cr) and returns the stacked (N, H, W, 3) outputs; autograd carries
gradients back to ``thresholds`` through the hand-written HIP adjoint (match
weights recomputed, the PCA basis differentiated too).  The images are inputs
of the forward only: no gradient flows to them.  Each processor keeps the state
of its last forward only, so backward raises if a processor ran another
forward between this forward and its backward.
"""
from __future__ import annotations

from typing import Sequence

import torch

from .pyramid import DenoiseParameters, PyramidProcessor


class _DenoiseThresholds(torch.autograd.Function):
    @staticmethod
    def forward(ctx, thresholds, images, processors, params, nlf):
        outs = []
        for img, proc, nl in zip(images, processors, nlf):
            out = proc.denoise_fwd(params, thresholds.detach().cpu(), img, nl)
            outs.append(out[..., :3].clone())
        ctx.processors = processors
        ctx.generations = [proc.generation for proc in processors]
        ctx.thr_device = thresholds.device
        return torch.stack(outs)

    @staticmethod
    def backward(ctx, grad_out):
        total = None
        for k, proc in enumerate(ctx.processors):
            g = torch.zeros(grad_out.shape[1:3] + (4,), dtype=torch.float32, device=grad_out.device)
            g[..., :3] = grad_out[k]
            gk = proc.denoise_bwd(g, generation=ctx.generations[k])
            total = gk if total is None else total + gk
        return total.to(ctx.thr_device), None, None, None, None


def denoise_thresholds(thresholds: torch.Tensor, images: Sequence[torch.Tensor],
                       processors: Sequence[PyramidProcessor], params: Sequence[DenoiseParameters],
                       nlf: Sequence[list]) -> torch.Tensor:
    """Batch denoise with differentiable thresholds; one processor per image (each keeps its own forward state)."""
    if not (len(images) == len(processors) == len(nlf)):
        raise ValueError("one processor and one noise model per image")
    return _DenoiseThresholds.apply(thresholds, list(images), list(processors), list(params), list(nlf))
