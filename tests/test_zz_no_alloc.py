"""gcam_build_pca_space_noalloc against gcam_build_pca_space, through the C ABI (GPU).

The allocation-free entry point runs the same kernels with the moment partials
in the library's fixed device array, so its basis and eigenvalues must equal
the allocating entry point's bit for bit.  Two calls in a row, the second on
fewer patches (the array keeps the first call's partials, which the second must
not see), then the size limits.
"""
import ctypes

import pytest
import torch

from glasscamerapublic_amd import _lib

pytestmark = pytest.mark.gpu


def patches(n, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    mix = torch.randn((25, 25), generator=g) / 5 + torch.eye(25)
    x = torch.randn((n, 25), generator=g) @ mix * torch.linspace(2.0, 0.1, 25)
    return x.to(torch.float32).contiguous().to("cuda")


def run(entry, p, n):
    basis = torch.full((25, 8), float("nan"), dtype=torch.float32, device="cuda")
    ev = torch.full((25,), float("nan"), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    st = entry(ctypes.c_void_p(p.data_ptr()), ctypes.c_int64(n), ctypes.c_void_p(basis.data_ptr()),
               ctypes.c_void_p(ev.data_ptr()), ctypes.c_void_p(stream))
    torch.cuda.synchronize()
    return st, basis.cpu(), ev.cpu()


def test_noalloc_equals_allocating_entry_point():
    L = _lib.lib()
    # each call has fewer patches than the one before
    for n, seed in ((200_000, 1), (3_001, 2), (700, 4)):
        p = patches(n, seed)
        st_a, basis_a, ev_a = run(L.gcam_build_pca_space, p, n)
        st_b, basis_b, ev_b = run(L.gcam_build_pca_space_noalloc, p, n)
        assert st_a == 0 and st_b == 0, _lib.lib().gcam_last_error()
        assert torch.isfinite(basis_b).all() and torch.isfinite(ev_b).all()
        assert torch.equal(basis_a, basis_b), f"n={n}: basis differs"
        assert torch.equal(ev_a, ev_b), f"n={n}: eigenvalues differ"


def test_noalloc_rejects_sizes_out_of_range():
    L = _lib.lib()
    p = patches(16, 3)
    for n in (1, (1 << 24) + 1):
        st, _, _ = run(L.gcam_build_pca_space_noalloc, p, n)
        assert st == _lib.GCAM_ERR_INVALID_ARGUMENT
        assert (b"at most" if n > 2 else b"at least") in L.gcam_last_error()
