"""Differentiation form on the GPU: the hand-written HIP adjoint
(gcam_pyramid_denoise_bwd) against the FP64 CPU adjoint of the oracle and
against central finite differences (eps = 1e-3) of the GPU forward.

Bar: within 1e-3 of the largest gradient component (per workload), i.e.
relative to the gradient's scale; components that are themselves ~0 (the
coarsest level's chroma) are held to the same absolute bound.
"""
import numpy as np
import pytest
import torch

import glasscamerapublic_amd as gcam
import oracle
from glasscamerapublic_amd.autograd import denoise_thresholds
from tests.helpers import nlf_for

pytestmark = pytest.mark.gpu


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


THR5 = np.array([[4.0, 2.0, 2.5], [3.0, 2.0, 2.0], [2.5, 1.5, 2.0], [2.0, 1.5, 1.5], [2.0, 1.5, 1.5]], np.float32)
CB5 = np.array([2.0, 1.5, 1.0, 1.0, 1.0], np.float32)


def params_for(levels):
    return [gcam.DenoiseParameters(luma=float(THR5[i, 0]), chroma=float(THR5[i, 1]), chromaBoost=float(CB5[i]))
            for i in range(levels)]


def gpu_loss_grad(proc, ycc, clean, thr, levels, nlf):
    out = proc.denoise_fwd(params_for(levels), torch.from_numpy(thr), ycc, nlf)
    diff = (out[..., :3] - clean[..., :3]).double()
    loss = 0.5 * float((diff * diff).sum())
    g = torch.zeros_like(out)
    g[..., :3] = out[..., :3] - clean[..., :3]
    return loss, proc.denoise_bwd(g).numpy().astype(np.float64)


@pytest.mark.parametrize("w,h,levels,radius", [(256, 256, 4, 2), (512, 512, 5, 4)])
def test_adjoint_matches_fp64_oracle(w, h, levels, radius, oracle_threads):
    ycc, clean = gcam.synth_frame(w, h, 5e-3, 2e-3, scene_seed=71, noise_seed=72, want_clean=True)
    nlf = [(tuple(x[0]), tuple(x[1])) for x in nlf_for(levels, 5e-3, 2e-3)]
    thr = THR5[:levels].copy()
    proc = gcam.PyramidProcessor(w, h, levels=levels, search_radius=radius)
    loss, g = gpu_loss_grad(proc, ycc, clean, thr, levels, nlf)
    ref_loss, ref_g, _ = oracle.pyramid_loss_grad(host(ycc).astype(np.float64), host(clean).astype(np.float64),
                                                  thr.astype(np.float64), CB5[:levels].astype(np.float64),
                                                  np.array(nlf), levels=levels, radius=radius)
    print("loss gpu", loss, "fp64", ref_loss)
    print(np.c_[g.ravel(), ref_g.ravel()])
    assert loss == pytest.approx(ref_loss, rel=1e-5)
    scale = np.abs(ref_g).max()
    np.testing.assert_allclose(g, ref_g, rtol=0, atol=1e-3 * scale)


def test_adjoint_matches_gpu_finite_differences():
    """512x512 crop, 5 levels, 9x9 window: the differentiation workload's crop."""
    w, h, levels, radius = 512, 512, 5, 4
    ycc, clean = gcam.synth_frame(w, h, 5e-3, 2e-3, scene_seed=81, noise_seed=82, want_clean=True)
    nlf = [(tuple(x[0]), tuple(x[1])) for x in nlf_for(levels, 5e-3, 2e-3)]
    thr = THR5.copy()
    proc = gcam.PyramidProcessor(w, h, levels=levels, search_radius=radius)
    _, g = gpu_loss_grad(proc, ycc, clean, thr, levels, nlf)
    eps = 1e-3
    fd = np.zeros_like(g)
    for i in range(levels):
        for c in range(3):
            tp, tm = thr.copy(), thr.copy()
            tp[i, c] += eps
            tm[i, c] -= eps
            lp, _ = gpu_loss_grad(proc, ycc, clean, tp, levels, nlf)
            lm, _ = gpu_loss_grad(proc, ycc, clean, tm, levels, nlf)
            fd[i, c] = (lp - lm) / (2 * eps)
    print(np.c_[g.ravel(), fd.ravel()])
    scale = np.abs(fd).max()
    np.testing.assert_allclose(g, fd, rtol=0, atol=1e-3 * scale)


def test_forward_of_differentiable_form_equals_denoise():
    w, h, levels = 256, 256, 4
    ycc = gcam.synth_frame(w, h, 5e-3, 2e-3, scene_seed=91, noise_seed=92)
    nlf = [(tuple(x[0]), tuple(x[1])) for x in nlf_for(levels)]
    prm = params_for(levels)
    thr = np.array([[p.luma, p.chroma, p.chroma] for p in prm], np.float32)
    proc = gcam.PyramidProcessor(w, h, levels=levels, search_radius=3)
    a = proc.denoise(prm, ycc, None, list(nlf)).clone()
    b = proc.denoise_fwd(prm, torch.from_numpy(thr), ycc, nlf).clone()
    assert torch.equal(a, b)


def test_autograd_batch_of_crops():
    """torch.autograd over a batch of crops with shared thresholds equals the sum of per-crop adjoints."""
    w, h, levels, radius, n = 512, 512, 5, 4, 3
    frames = [gcam.synth_frame(w, h, 5e-3, 2e-3, scene_seed=100 + k, noise_seed=200 + k, want_clean=True)
              for k in range(n)]
    nlf = [[(tuple(x[0]), tuple(x[1])) for x in nlf_for(levels, 5e-3, 2e-3)] for _ in range(n)]
    procs = [gcam.PyramidProcessor(w, h, levels=levels, search_radius=radius) for _ in range(n)]
    thr = torch.tensor(THR5, requires_grad=True)
    out = denoise_thresholds(thr, [f[0] for f in frames], procs, params_for(levels), nlf)
    clean = torch.stack([f[1][..., :3] for f in frames])
    loss = 0.5 * ((out - clean) ** 2).sum()
    loss.backward()
    expect = np.zeros((levels, 3))
    for k in range(n):
        expect += gpu_loss_grad(procs[k], frames[k][0], frames[k][1], THR5.copy(), levels, nlf[k])[1]
    np.testing.assert_allclose(thr.grad.numpy(), expect, rtol=1e-5, atol=1e-6 * np.abs(expect).max())


def test_backward_after_another_forward_raises():
    """Each processor keeps only its last forward: a second forward (a validation pass, say) between a
    forward and its backward must make the backward raise, not differentiate the later call."""
    w, h, levels = 256, 256, 3
    frames = [gcam.synth_frame(w, h, 5e-3, 2e-3, scene_seed=400 + k, noise_seed=500 + k) for k in range(2)]
    nlf = [[(tuple(x[0]), tuple(x[1])) for x in nlf_for(levels, 5e-3, 2e-3)] for _ in range(2)]
    procs = [gcam.PyramidProcessor(w, h, levels=levels, search_radius=2) for _ in range(2)]
    thr = torch.tensor(THR5[:levels], requires_grad=True)
    out = denoise_thresholds(thr, frames, procs, params_for(levels), nlf)
    with torch.no_grad():
        denoise_thresholds(thr, frames[::-1], procs, params_for(levels), nlf)  # overwrites both processors
    with pytest.raises(RuntimeError, match="more forward"):
        out.sum().backward()
    out = denoise_thresholds(thr, frames, procs, params_for(levels), nlf)
    out.sum().backward()  # a backward right after its forward still works
    assert torch.isfinite(thr.grad).all()


def test_unsupported_settings_are_rejected():
    proc = gcam.PyramidProcessor(256, 256, levels=3, search_radius=2)
    ycc = gcam.synth_frame(256, 256, 5e-3, 2e-3)
    prm = [gcam.DenoiseParameters(gradientBoost=1.0)] * 3
    nlf = [(tuple(x[0]), tuple(x[1])) for x in nlf_for(3)]
    with pytest.raises(gcam.GcamError, match="gradientBoost"):
        proc.denoise_fwd(prm, torch.ones(3, 3), ycc, nlf)
    with pytest.raises(gcam.GcamError, match="no gcam_pyramid_denoise_fwd"):
        proc.denoise_bwd(torch.zeros((256, 256, 4), device="cuda"))


def test_differentiation_workload_16_crops(oracle_threads):
    """The differentiation workload: 16 crops of 512x512, 5 levels, 9x9 window,
    a = 5e-3, b = 2e-3, L = 1/2 |out - clean|^2 summed over the batch; the 15
    threshold gradients from torch.autograd through the HIP adjoint, against
    central differences (eps = 1e-3) of the GPU batch loss and against the sum
    of the FP64 CPU adjoints.  Components above 5 % of the largest are also held
    to 1e-3 relative against the FP64 gradient."""
    w, h, levels, radius, n = 512, 512, 5, 4, 16
    frames = [gcam.synth_frame(w, h, 5e-3, 2e-3, scene_seed=300 + k, noise_seed=400 + k, want_clean=True)
              for k in range(n)]
    nlf = [(tuple(x[0]), tuple(x[1])) for x in nlf_for(levels, 5e-3, 2e-3)]
    procs = [gcam.PyramidProcessor(w, h, levels=levels, search_radius=radius) for _ in range(n)]
    clean = torch.stack([f[1][..., :3] for f in frames])

    def batch_loss(t):
        out = denoise_thresholds(t, [f[0] for f in frames], procs, params_for(levels), [nlf] * n)
        return 0.5 * ((out - clean).double() ** 2).sum()

    thr = torch.tensor(THR5, requires_grad=True)
    batch_loss(thr).backward()
    g = thr.grad.numpy().astype(np.float64)

    eps = 1e-3
    fd = np.zeros((levels, 3))
    with torch.no_grad():
        for i in range(levels):
            for c in range(3):
                tp, tm = THR5.copy(), THR5.copy()
                tp[i, c] += eps
                tm[i, c] -= eps
                fd[i, c] = (float(batch_loss(torch.from_numpy(tp))) - float(batch_loss(torch.from_numpy(tm)))) / (2 * eps)

    ref = np.zeros((levels, 3))
    for k in range(n):
        ref += oracle.pyramid_loss_grad(host(frames[k][0]).astype(np.float64), host(frames[k][1]).astype(np.float64),
                                        THR5.astype(np.float64), CB5.astype(np.float64), np.array(nlf),
                                        levels=levels, radius=radius)[1]
    print(np.c_[g.ravel(), fd.ravel(), ref.ravel()])
    scale = np.abs(ref).max()
    np.testing.assert_allclose(g, ref, rtol=0, atol=1e-3 * scale)
    np.testing.assert_allclose(g, fd, rtol=0, atol=1e-3 * scale)
    big = np.abs(ref) > 0.05 * scale
    rel = np.abs(g - ref)[big] / np.abs(ref)[big]
    print("relative error of the significant components vs FP64:", rel.max())
    assert rel.max() < 1e-3
