"""Device FP64 products of the Davidson subspace.

`tn(A, B)` = A^T B for tall A (n, a) and B (n, b): the reductions over the
pair index of the Rayleigh-Ritz step and of the trial-space
orthogonalisation, run by libmbpt_hip.so's split-contraction MFMA kernel
(subspace.hip).  A and B may be column slices of wider row-major buffers.
"""
import ctypes

import torch

from . import _lib


class _Work:
    buf = {}

    @classmethod
    def get(cls, dev, elems):
        b = cls.buf.get(dev)
        if b is None or b.numel() < elems:
            b = torch.empty(max(elems, 1 << 16), dtype=torch.float64, device=dev)
            cls.buf[dev] = b
        return b


def _rowmajor(t):
    """(pointer, row stride) of a 2-D FP64 device tensor whose rows are
    unit-stride (a column slice of a contiguous buffer qualifies)."""
    if t.dim() != 2 or t.dtype != torch.float64 or not t.is_cuda:
        raise ValueError('expected a 2-D float64 device tensor')
    if t.shape[1] > 1 and t.stride(1) != 1:
        raise ValueError('rows must be unit-stride')
    return ctypes.c_void_p(t.data_ptr()), int(t.stride(0))


def tn(A, B, out=None):
    """A^T B, (a, b) FP64 on A's device (A (n, a), B (n, b))."""
    n, a = A.shape
    n2, b = B.shape
    if n != n2:
        raise ValueError(f'row counts differ: {n} vs {n2}')
    if out is None:
        out = torch.empty((a, b), dtype=torch.float64, device=A.device)
    if a == 0 or b == 0:
        return out
    if a > 1 and A.stride(1) != 1:
        A = A.contiguous()
    if b > 1 and B.stride(1) != 1:
        B = B.contiguous()
    lib = _lib.lib()
    pa, lda = _rowmajor(A)
    pb, ldb = _rowmajor(B)
    pc, ldc = _rowmajor(out)
    elems = int(lib.mbpt_gemm_tn_f64_work(n, a, b))
    work = _Work.get(A.device, elems)
    _lib.check(lib.mbpt_gemm_tn_f64(n, a, b, pa, max(lda, a), pb, max(ldb, b), pc,
                                    max(ldc, b), _lib.dptr(work), work.numel(),
                                    _lib.stream_ptr()), 'mbpt_gemm_tn_f64')
    return out
