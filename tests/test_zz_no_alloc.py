"""The allocation-free A^T B entry point, mbpt_gemm_tn_f64_noalloc, called
through the C ABI against mbpt_gemm_tn_f64 (the same kernels with a caller's
work buffer) on two calls in a row, the second smaller, so the fixed scratch
array's leftovers from the first call must not leak into the second.  The
chunking is the same, so the results agree to the bit; both are also checked
against an FP64 product.  An oversized product is refused before any launch.
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _tn_with_work(lib, A, B):
    n, a = A.shape
    b = B.shape[1]
    C = torch.full((a, b), float('nan'), dtype=torch.float64, device='cuda')
    elems = int(lib.mbpt_gemm_tn_f64_work(n, a, b))
    work = torch.empty(max(elems, 1), dtype=torch.float64, device='cuda')
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.mbpt_gemm_tn_f64(n, a, b, _ptr(A), a, _ptr(B), b, _ptr(C), b, _ptr(work),
                              work.numel(), stream)
    assert rc == 0, lib.mbpt_last_error()
    return C


def _tn_noalloc(lib, A, B):
    n, a = A.shape
    b = B.shape[1]
    C = torch.full((a, b), float('nan'), dtype=torch.float64, device='cuda')
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.mbpt_gemm_tn_f64_noalloc(n, a, b, _ptr(A), a, _ptr(B), b, _ptr(C), b, stream)
    assert rc == 0, lib.mbpt_last_error()
    return C


def test_noalloc_matches_the_work_buffer_entry_on_two_calls():
    from mbptcode_amd import _lib
    lib = _lib.lib()
    # the headline Davidson's widest product (70000 pairs, 161 trial vectors
    # against 30 new ones), then a smaller one after it
    for n, a, b, seed in ((70000, 161, 30, 0), (5000, 40, 20, 1)):
        print(f'noalloc tn: n={n} a={a} b={b}', flush=True)
        g = torch.Generator(device='cuda').manual_seed(seed)
        A = torch.randn((n, a), generator=g, dtype=torch.float64, device='cuda')
        B = torch.randn((n, b), generator=g, dtype=torch.float64, device='cuda')
        ref = _tn_with_work(lib, A, B)
        got = _tn_noalloc(lib, A, B)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), float((got - ref).abs().max())
        exact = (A.cpu().numpy().T @ B.cpu().numpy())
        assert np.abs(got.cpu().numpy() - exact).max() <= 1e-12 * np.abs(exact).max()


def test_noalloc_refuses_a_product_beyond_its_scratch():
    from mbptcode_amd import _lib
    lib = _lib.lib()
    cap = int(lib.mbpt_gemm_tn_f64_noalloc_max())
    assert cap == 1024 * 64 * 64
    # a, b = 2100: 33 x 33 > 1024 output tiles, so one chunk of a * b > cap
    n, a, b = 16, 2100, 2100
    assert int(lib.mbpt_gemm_tn_f64_work(n, a, b)) > cap
    A = torch.zeros((n, a), dtype=torch.float64, device='cuda')
    C = torch.zeros((a, b), dtype=torch.float64, device='cuda')
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.mbpt_gemm_tn_f64_noalloc(n, a, b, _ptr(A), a, _ptr(A), a, _ptr(C), b, stream)
    assert rc == _lib.ERR_VALUE
    assert b'fixed scratch' in lib.mbpt_last_error()
