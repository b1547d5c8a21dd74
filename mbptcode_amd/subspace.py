"""Device FP64 products of the Davidson subspace.

`tn(A, B)` = A^T B for tall A (n, a) and B (n, b): the reductions over the
pair index of the Rayleigh-Ritz step and of the trial-space
orthogonalisation, run by libmbpt_hip.so's split-contraction MFMA kernel
(subspace.hip).  A and B may be column slices of wider row-major buffers.

The single-instance Davidson keeps V and AV as column blocks of one
row-major buffer and names an operand by up to two column ranges of it
(`Ranges`): `tn_ranges` (A^T B), `nn_ranges` (C = beta C + alpha A Y),
`ritz_residual` (X = V Y, R = AV Y - X diag(theta)) and `jd_correction`
(the Jacobi-Davidson correction), all in dvd.hip.
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


class Ranges:
    """Up to two column ranges of one row-major FP64 device buffer `buf`
    (n, ld): ranges [(start, count), ...] read as one operand of
    sum(count) logical columns."""
    __slots__ = ('buf', 'ranges')

    def __init__(self, buf, *ranges):
        if buf.dim() != 2 or buf.dtype != torch.float64:
            raise ValueError('expected a 2-D float64 buffer')
        if not buf.is_contiguous():
            raise ValueError('the buffer must be contiguous (row-major)')
        rs = [(int(a), int(c)) for a, c in ranges if int(c) > 0]
        if len(rs) > 2:
            raise ValueError('at most two column ranges')
        for a, c in rs:
            if a < 0 or a + c > buf.shape[1]:
                raise ValueError(f'columns [{a}, {a + c}) outside a buffer of {buf.shape[1]}')
        self.buf, self.ranges = buf, rs

    @property
    def ncol(self):
        return sum(c for _, c in self.ranges)

    def columns(self):
        """The logical columns as one (n, ncol) tensor (a copy when two ranges)."""
        parts = [self.buf[:, a:a + c] for a, c in self.ranges]
        if not parts:
            return self.buf[:, :0]
        return parts[0] if len(parts) == 1 else torch.cat(parts, dim=1)

    def abi(self):
        """(pointer, ld, n1, off1, off2) for the C ABI."""
        rs = self.ranges + [(0, 0)] * (2 - len(self.ranges))
        (o1, n1), (o2, _) = rs
        return (ctypes.c_void_p(self.buf.data_ptr()), int(self.buf.shape[1]), n1, o1, o2)


def tn_ranges(A, B, out=None):
    """A^T B for two `Ranges` of buffers with equal row counts, (a, b) FP64."""
    n = A.buf.shape[0]
    if B.buf.shape[0] != n:
        raise ValueError('row counts differ')
    a, b = A.ncol, B.ncol
    if out is None:
        out = torch.empty((a, b), dtype=torch.float64, device=A.buf.device)
    if a == 0 or b == 0:
        return out
    pc, ldc = _rowmajor(out)
    lib = _lib.lib()
    elems = int(lib.mbpt_gemm_tn_seg_f64_work(n, a, b))
    work = _Work.get(A.buf.device, elems)
    _lib.check(lib.mbpt_gemm_tn_seg_f64(n, a, b, *A.abi(), *B.abi(), pc, max(ldc, b),
                                        _lib.dptr(work), work.numel(), _lib.stream_ptr()),
               'mbpt_gemm_tn_seg_f64')
    return out


def nn_ranges(A, Y, C, c_off, alpha=1.0, beta=0.0):
    """C[:, c_off : c_off + c] = beta C + alpha A Y for `Ranges` A (n, k), Y
    (k, c) a row-major device tensor (c <= 64), C a contiguous (n, ldc)
    buffer."""
    n, k = A.buf.shape[0], A.ncol
    c = Y.shape[1]
    if Y.shape[0] != k:
        raise ValueError(f'Y has {Y.shape[0]} rows, A {k} columns')
    if C.shape[0] != n or not C.is_contiguous():
        raise ValueError('C must be a contiguous buffer with the rows of A')
    py, ldy = _rowmajor(Y)
    _lib.check(_lib.lib().mbpt_gemm_nn_seg_f64(n, k, c, *A.abi(), py, max(ldy, c),
                                               _lib.dptr(C), int(C.shape[1]), int(c_off),
                                               float(alpha), float(beta), _lib.stream_ptr()),
               'mbpt_gemm_nn_seg_f64')
    return C


def ritz_residual(buf, m, v_off, av_off, Y, theta, X, R):
    """X = V Y and R = AV Y - X diag(theta) for the m columns of V at v_off and
    of AV at av_off in `buf` (n, ld); Y (m, k), theta (k), X and R (n, k)
    contiguous, k <= 64."""
    n = buf.shape[0]
    k = Y.shape[1]
    py, ldy = _rowmajor(Y)
    _lib.check(_lib.lib().mbpt_ritz_residual_f64(
        n, m, k, _lib.dptr(buf), int(buf.shape[1]), int(v_off), _lib.dptr(buf),
        int(buf.shape[1]), int(av_off), py, max(ldy, k), _lib.dptr(theta), _lib.dptr(X),
        int(X.shape[1]), _lib.dptr(R), int(R.shape[1]), _lib.stream_ptr()),
        'mbpt_ritz_residual_f64')
    return X, R


def block_scale(stage, m, p, drop, buf, t_off):
    """Unit scaling of the projected block buf[:, t_off : t_off + p] from the
    orthonormalisation's projections `stage` ((3m + 2p + 2, p) contiguous):
    its last two rows receive the columns' norms and keep flags, and the
    block's columns are scaled to unit norm or zeroed (dvd.hip)."""
    n = buf.shape[0]
    if stage.shape != (3 * m + 2 * p + 2, p) or not stage.is_contiguous():
        raise ValueError('stage must be a contiguous (3m + 2p + 2, p) buffer')
    if not (buf.is_contiguous() and 0 <= t_off and t_off + p <= buf.shape[1]):
        raise ValueError('the block must be p columns of a contiguous (n, ld) buffer')
    _lib.check(_lib.lib().mbpt_block_scale_f64(
        n, m, p, float(drop), _lib.dptr(stage),
        ctypes.c_void_p(buf.data_ptr() + 8 * t_off), int(buf.shape[1]), _lib.stream_ptr()),
        'mbpt_block_scale_f64')


def ritz_max_m():
    """The largest Rayleigh matrix `ritz_eigh` takes."""
    return int(_lib.lib().mbpt_ritz_max_m())


def ritz_eigh(H, G, k, Y=None):
    """The lowest k eigenpairs of the symmetric part of H, on the device.
    H and G are (m, m) FP64 views with unit column stride (blocks of the
    Davidson's [H | G] buffer); returns Y (m, k) with the eigenvectors as
    columns and wq (3k,): the eigenvalues ascending, y^T G y and y^T H y of
    each column (ritz.hip)."""
    m = H.shape[0]
    if H.shape != (m, m) or G.shape != (m, m) or H.stride(1) != 1 or G.stride(1) != 1:
        raise ValueError('H and G must be (m, m) views with unit column stride')
    lib = _lib.lib()
    if Y is None:
        Y = torch.empty((m, k), dtype=torch.float64, device=H.device)
    wq = torch.empty(3 * k, dtype=torch.float64, device=H.device)
    work = _Work.get(H.device, int(lib.mbpt_ritz_eigh_f64_work(m, k)))
    _lib.check(lib.mbpt_ritz_eigh_f64(
        m, k, ctypes.c_void_p(H.data_ptr()), int(H.stride(0)), ctypes.c_void_p(G.data_ptr()),
        int(G.stride(0)), _lib.dptr(Y), int(Y.stride(0)), _lib.dptr(wq), _lib.dptr(work),
        work.numel(), _lib.stream_ptr()), 'mbpt_ritz_eigh_f64')
    return Y, wq


def jd_correction(U, R, diag, theta, T, t_off=0, floor=1e-8):
    """T[:, t_off : t_off + k] = K^-1 R - alpha K^-1 U per column (the
    Jacobi-Davidson correction, Solvers/davidson.py:43-77); U, R (n, k)
    contiguous, T a contiguous (n, ldt) buffer."""
    n, k = R.shape
    if not (T.is_contiguous() and T.shape[0] == n and 0 <= t_off and t_off + k <= T.shape[1]):
        raise ValueError('T must be a contiguous (n, ldt) buffer with room for k columns at t_off')
    if U.shape != R.shape or not (U.is_contiguous() and R.is_contiguous()):
        raise ValueError('U and R must be contiguous and of one shape')
    lib = _lib.lib()
    work = _Work.get(R.device, int(lib.mbpt_jd_correction_f64_work(n, k)))
    Tv = T[:, t_off:]
    _lib.check(lib.mbpt_jd_correction_f64(
        n, k, _lib.dptr(U), int(U.shape[1]), _lib.dptr(R), int(R.shape[1]), _lib.dptr(diag),
        _lib.dptr(theta), float(floor), ctypes.c_void_p(Tv.data_ptr()), int(T.shape[1]),
        _lib.dptr(work), work.numel(), _lib.stream_ptr()), 'mbpt_jd_correction_f64')
    return T
