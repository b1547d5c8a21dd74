"""Space-time polarizability, static screening and the dRPA energy on the GPU.

Drop-ins for src/SingleReference/LinearResponse/space_time.py:

* `split_branches` (:448), `polarizability_projected_tau` (:478),
  `polarizability_projected_sweep` (:533), `chi0_imaginary_frequency` (:607),
  `static_screening_matrix` (:1003), `dyson_screening` (:956),
  `dielectric_log_determinant` (:1073), `rpa_correlation_energy_space_time`
  (:1100);
* plus the grid-gauge forms the synthetic workloads use
  (`chi0_grid`, `screened_interaction_grid`, `rpa_correlation_energy_grid`),
  where the interaction is given on the points (V) instead of as a factor D.

The N^3 step -- every tau point's Go o Gv -- runs in libmbpt_hip.so (chi0.hip,
FP64 or FP32 MFMA, tau loop fused).  The steps the reference keeps in FP64
and that are N^3 only once (the D projection, the Dyson solve, log-dets) run
in FP64 on the device through torch (rocBLAS / rocSOLVER): W through a
Cholesky factorisation of the positive-definite I - D^T chi0 D, never an
explicit inverse.
"""
import ctypes
import time
import warnings

import numpy as np
import torch

from . import _lib
from .constants import ISDF_TILE_GB
from .grids import static_screening_grid
from .orbital_spaces import particle_hole_window


class NonPositiveDielectric(RuntimeWarning):
    """det(1 - chi0) is not positive at an imaginary frequency."""


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError('the space-time chi0 kernels run on a ROCm GPU; none is visible')
    return torch.device('cuda', torch.cuda.current_device())


def _dev(a, dev):
    if isinstance(a, torch.Tensor):
        return a.to(device=dev, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)


def split_branches(X, eps, nocc, mu=None):
    """(X_o, X_v, e_o, e_v, mu, occ, virt): the collocation split at the gap,
    energies shifted by mu (default mid-gap) (space_time.py:448-475)."""
    eps = np.asarray(eps, float)
    if mu is None:
        mu = particle_hole_window(eps, nocc)[2]
    occ, virt = np.arange(nocc), np.arange(nocc, len(eps))
    if isinstance(X, torch.Tensor):
        X_o, X_v = X[:, :nocc].contiguous(), X[:, nocc:].contiguous()
    else:
        X = np.asarray(X, float)
        X_o, X_v = np.ascontiguousarray(X[:, occ]), np.ascontiguousarray(X[:, virt])
    return X_o, X_v, eps[occ] - mu, eps[virt] - mu, mu, occ, virt


class Chi0Builder:
    """Device-resident collocation for repeated chi0 builds (one handle)."""

    def __init__(self, X, eps, nocc, mu=None, precision='fp32', device=None):
        dev = _device() if device is None else torch.device(device)
        eps = np.asarray(eps, float)
        if mu is None:
            mu = particle_hole_window(eps, nocc)[2]
        Xd = _dev(X, dev)
        self.M, norb = Xd.shape
        if norb != len(eps):
            raise ValueError(f'X has {norb} orbitals, eps has {len(eps)}')
        self.nocc, self.nvir = nocc, norb - nocc
        self.device, self.precision, self.mu = dev, precision, float(mu)
        self._lib = _lib.lib()
        h = ctypes.c_void_p()
        _, eps_p = _lib.host_doubles(eps)
        self._eps_keep = _
        with torch.cuda.device(dev):
            _lib.check(self._lib.mbpt_chi0_create(
                ctypes.byref(h), self.M, nocc, self.nvir,
                _lib.precision_code(precision), _lib.dptr(Xd), eps_p, self.mu,
                _lib.stream_ptr()), 'mbpt_chi0_create')
            torch.cuda.current_stream().synchronize()
        self._h = h

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            try:
                torch.cuda.synchronize(self.device)
            except Exception:
                pass
            self._lib.mbpt_chi0_destroy(h)
            self._h = None

    def profile(self, enable=True):
        _lib.check(self._lib.mbpt_chi0_profile(self._h, int(enable)), 'mbpt_chi0_profile')

    def profile_read(self):
        """{'ms', 'launches', 'flops'} of the fused chi0 launches since `profile()`."""
        ms, fl = ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_long()
        _lib.check(self._lib.mbpt_chi0_profile_read(self._h, ctypes.byref(ms),
                                                     ctypes.byref(n), ctypes.byref(fl)),
                   'mbpt_chi0_profile_read')
        return {'ms': float(ms.value), 'launches': int(n.value), 'flops': float(fl.value)}

    def sigma_sums(self, tau, Zt, states):
        """(2, ntau, nstates) FP64 device: the space-time self-energy's branch
        sums (`mbpt_chi0_sigma`), + sum X_s (Zt o Go) X_s and
        - sum X_s (Zt o Gv) X_s per time point, Zt (ntau, M, M) FP32 device."""
        if self.precision != 'fp32':
            raise NotImplementedError('the self-energy sums run on FP32 handles')
        tau = np.atleast_1d(np.asarray(tau, float))
        states = np.atleast_1d(np.asarray(states, dtype=np.int32))
        if Zt.dtype != torch.float32 or tuple(Zt.shape) != (len(tau), self.M, self.M):
            raise ValueError(f'Zt must be FP32 ({len(tau)}, {self.M}, {self.M})')
        if not Zt.is_contiguous():
            raise ValueError('Zt must be contiguous')
        out = torch.empty((2, len(tau), len(states)), dtype=torch.float64,
                          device=self.device)
        for t0 in range(0, len(tau), _lib.CHI0_MAX_TAU):
            t1 = min(len(tau), t0 + _lib.CHI0_MAX_TAU)
            tt, tp = _lib.host_doubles(tau[t0:t1])
            part = out if (t0 == 0 and t1 == len(tau)) else torch.empty(
                (2, t1 - t0, len(states)), dtype=torch.float64, device=self.device)
            st = np.ascontiguousarray(states)
            _lib.check(self._lib.mbpt_chi0_sigma(
                self._h, t1 - t0, tp, _lib.dptr(Zt[t0:t1]), len(st),
                st.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _lib.dptr(part),
                _lib.stream_ptr()), 'mbpt_chi0_sigma')
            if part is not out:
                out[:, t0:t1] = part
        return out

    def sigma_profile_read(self):
        """{'ms', 'launches', 'flops'} of the self-energy sums since `profile()`."""
        ms, fl = ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_long()
        _lib.check(self._lib.mbpt_chi0_sigma_profile_read(
            self._h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(fl)),
            'mbpt_chi0_sigma_profile_read')
        return {'ms': float(ms.value), 'launches': int(n.value), 'flops': float(fl.value)}

    def terms(self, tau):
        """Build and keep P(t) = -2 Go(t) o Gv(t) for every time point
        (`mbpt_chi0_terms`): each Green's-function product once, for any
        number of later `transform` calls."""
        tau = np.atleast_1d(np.asarray(tau, float))
        if not 1 <= len(tau) <= _lib.CHI0_MAX_TAU:
            raise ValueError(f'{len(tau)} time points: 1 to {_lib.CHI0_MAX_TAU} are held')
        tt, tp = _lib.host_doubles(tau)
        _lib.check(self._lib.mbpt_chi0_terms(self._h, len(tt), tp, _lib.stream_ptr()),
                   'mbpt_chi0_terms')
        self._terms_nt = len(tt)
        return self

    def transform(self, weights, out=None, accumulate=False):
        """(nfreq, M, M) FP64: out[f] = [out[f]] + sum_t weights[f, t] P(t)
        over the kept terms (`mbpt_chi0_transform`), weights (nfreq, ntau)."""
        W = np.atleast_2d(np.asarray(weights, float))
        nt = getattr(self, '_terms_nt', 0)
        if W.shape[1] != nt:
            raise ValueError(f'weights have {W.shape[1]} time points, the terms {nt}')
        if out is None:
            out = torch.empty((W.shape[0], self.M, self.M), dtype=torch.float64,
                              device=self.device)
            accumulate = False
        if tuple(out.shape) != (W.shape[0], self.M, self.M):
            raise ValueError(f'out must be ({W.shape[0]}, {self.M}, {self.M})')
        ww, wp = _lib.host_doubles(W)
        _lib.check(self._lib.mbpt_chi0_transform(self._h, W.shape[0], wp, _lib.dptr(out),
                                                  int(accumulate), _lib.stream_ptr()),
                   'mbpt_chi0_transform')
        return out

    def transform_profile_read(self):
        """{'ms', 'calls'} of `transform` since `profile()`."""
        ms, n = ctypes.c_double(), ctypes.c_long()
        _lib.check(self._lib.mbpt_chi0_transform_profile_read(self._h, ctypes.byref(ms),
                                                               ctypes.byref(n)),
                   'mbpt_chi0_transform_profile_read')
        return {'ms': float(ms.value), 'calls': int(n.value)}

    def build(self, tau, weights, out=None, accumulate=False):
        """out (M, M) FP64 = [out] + sum_t weights[t] (-2) Go(t) o Gv(t)."""
        tau = np.atleast_1d(np.asarray(tau, float))
        weights = np.atleast_1d(np.asarray(weights, float))
        if tau.shape != weights.shape:
            raise ValueError('tau and weights differ in length')
        if out is None:
            out = torch.empty((self.M, self.M), dtype=torch.float64, device=self.device)
            accumulate = False
        for t0 in range(0, len(tau), _lib.CHI0_MAX_TAU):
            tt, tp = _lib.host_doubles(tau[t0:t0 + _lib.CHI0_MAX_TAU])
            ww, wp = _lib.host_doubles(weights[t0:t0 + _lib.CHI0_MAX_TAU])
            _lib.check(self._lib.mbpt_chi0_build(self._h, len(tt), tp, wp,
                                                  _lib.dptr(out),
                                                  int(accumulate or t0 > 0),
                                                  _lib.stream_ptr()),
                       'mbpt_chi0_build')
        return out


#: Largest sum over the time points of |weight| for which chi0 at a frequency
#: is formed from FP32 terms.  Each FP32 term carries a rounding of about
#: 1e-7 of its size and the weights multiply it: the regularized 20-point
#: grids keep the sum below 600 (E_c within 0.03 uHa of FP64 on the 4000-point
#: instance), while the unregularized 16-point fit of the 1200-point instance
#: reaches 1e9 with cancelling signs, where FP32 terms leave nothing of chi0.
#: Frequencies above the limit are formed from FP64 terms.
FP32_WEIGHT_SUM_LIMIT = 1e4


def fp32_weights_ok(weights):
    """Per row of `weights` (nfreq, ntau): whether FP32 terms suffice."""
    W = np.atleast_2d(np.asarray(weights, float))
    return np.abs(W).sum(axis=1) <= FP32_WEIGHT_SUM_LIMIT


def chi0_grid(X, eps, nocc, tau, weights, mu=None, precision='fp32', builder=None):
    """sum_t weights[t] (-2) Go(t) o Gv(t) on the grid, (M, M) FP64 device.
    precision='fp32' builds in FP64 when the weights would amplify FP32
    rounding beyond `FP32_WEIGHT_SUM_LIMIT`."""
    if builder is None:
        if precision == 'fp32' and not fp32_weights_ok(weights)[0]:
            precision = 'fp64'
        builder = Chi0Builder(X, eps, nocc, mu=mu, precision=precision)
    return builder.build(tau, weights)


class FrequencySweep:
    """chi0 at the frequencies of a weight matrix (nfreq, ntau) from terms
    built once per time point (`Chi0Builder.terms`), the one-pass form of the
    reference's tau-to-frequency accumulation (space_time.py:634-651,
    :854-905).  precision='fp32': FP32 terms for the rows `fp32_weights_ok`
    passes, FP64 terms for the others; 'fp64': FP64 terms throughout."""

    def __init__(self, X, eps, nocc, tau, weights, mu=None, precision='fp32',
                 device=None, profile=False):
        self.W = np.atleast_2d(np.asarray(weights, float))
        tau = np.atleast_1d(np.asarray(tau, float))
        if self.W.shape[1] != len(tau):
            raise ValueError(f'weights have {self.W.shape[1]} time points, tau {len(tau)}')
        if precision not in ('fp32', 'fp64'):
            raise ValueError(f"precision {precision!r}: 'fp32' or 'fp64'")
        self.use64 = (np.ones(len(self.W), bool) if precision == 'fp64'
                      else ~fp32_weights_ok(self.W))
        self.builders = {}
        for key, need in (('fp32', ~self.use64), ('fp64', self.use64)):
            if need.any():
                b = Chi0Builder(X, eps, nocc, mu=mu, precision=key, device=device)
                if profile:
                    b.profile(True)
                self.builders[key] = b.terms(tau)
        first = next(iter(self.builders.values()))
        self.M, self.device = first.M, first.device

    @property
    def nfreq(self):
        return len(self.W)

    def build(self, rows, out=None):
        """(len(rows), M, M) FP64 device: chi0 at the frequencies `rows`."""
        rows = np.atleast_1d(np.asarray(rows, int))
        if out is None:
            out = torch.empty((len(rows), self.M, self.M), dtype=torch.float64,
                              device=self.device)
        hi = self.use64[rows]
        for key, mask in (('fp32', ~hi), ('fp64', hi)):
            idx = np.flatnonzero(mask)
            if idx.size == len(rows):
                self.builders[key].transform(self.W[rows], out=out[:len(rows)])
            elif idx.size:
                for j in idx:
                    self.builders[key].transform(self.W[rows[j]][None], out=out[j:j + 1])
        return out

    def profile(self, enable=True):
        for b in self.builders.values():
            b.profile(enable)

    def profile_read(self):
        """{'terms_ms', 'terms_flops', 'transform_ms'} summed over the builders."""
        out = {'terms_ms': 0.0, 'terms_flops': 0.0, 'transform_ms': 0.0}
        for b in self.builders.values():
            p, t = b.profile_read(), b.transform_profile_read()
            out['terms_ms'] += p['ms']
            out['terms_flops'] += p['flops']
            out['transform_ms'] += t['ms']
        return out


def _project(P, D):
    """D^T P D in FP64 (D None: the grid gauge, P itself)."""
    if D is None:
        return P
    return D.T @ (P @ D)


def polarizability_projected_tau(X_o, X_v, e_o, e_v, D, tau,
                                 tile_memory_gb=ISDF_TILE_GB, out=None, work=None,
                                 tiles=None, precision='fp64'):
    """proj(tau) = -2 D^T (Go o Gv) D at one time point (space_time.py:478-530).

    X_o, X_v, e_o, e_v as `split_branches` returns them (energies already
    shifted).  D (M, naux), or None for the grid gauge.  `out` (naux, naux)
    is overwritten when given; `work` is accepted and never read (the kernel
    tiles in registers and LDS); `tiles` must be None (the whole sum).
    Returns a NumPy array for NumPy inputs, else a device tensor.
    """
    if tiles is not None:
        raise NotImplementedError('partial tile sums are a multi-rank feature')
    numpy_in = not isinstance(X_o, torch.Tensor)
    dev = _device()
    Xo, Xv = _dev(X_o, dev), _dev(X_v, dev)
    X = torch.cat([Xo, Xv], dim=1)
    eps = np.concatenate([np.asarray(e_o, float), np.asarray(e_v, float)])
    P = chi0_grid(X, eps, Xo.shape[1], [float(tau)], [1.0], mu=0.0,
                  precision=precision)
    Dd = None if D is None else _dev(D, dev)
    res = _project(P, Dd)
    if out is not None:
        if isinstance(out, torch.Tensor):
            out.copy_(res)
        else:
            out[...] = res.cpu().numpy()
        return out
    return res.cpu().numpy() if numpy_in else res


def polarizability_projected_sweep(X, D, eps, nocc, tau_points, mu=None,
                                   tau_indices=None, tile_memory_gb=ISDF_TILE_GB,
                                   precision='fp64'):
    """proj(tau) on every time point, (ntau, naux, naux) (space_time.py:533-560);
    slots outside `tau_indices` stay zero."""
    dev = _device()
    b = Chi0Builder(X, eps, nocc, mu=mu, precision=precision)
    Dd = None if D is None else _dev(D, dev)
    tau_points = np.asarray(tau_points, float)
    naux = b.M if Dd is None else Dd.shape[1]
    proj = torch.zeros((len(tau_points), naux, naux), dtype=torch.float64, device=dev)
    which = range(len(tau_points)) if tau_indices is None else np.atleast_1d(tau_indices)
    for k in which:
        proj[k] = _project(b.build([tau_points[k]], [1.0]), Dd)
    return proj


def chi0_imaginary_frequency(X, D, eps, nocc, grid, mu=None, tau_indices=None,
                             tile_memory_gb=ISDF_TILE_GB, precision='fp64'):
    """chi0(i w_k) = sum_tau cosft_wt[k, tau] proj(tau), (nfreq, naux, naux)
    (space_time.py:607-631): every proj(tau) built once, then one transform
    pass forms all frequencies (`FrequencySweep`)."""
    dev = _device()
    Dd = None if D is None else _dev(D, dev)
    which = np.arange(grid.ntau) if tau_indices is None else np.atleast_1d(tau_indices)
    tau = np.asarray(grid.tau_points)[which]
    W = np.asarray(grid.cosft_wt, float)[:, which]
    sweep = FrequencySweep(X, eps, nocc, tau, W, mu=mu, precision=precision, device=dev)
    if Dd is None:
        return sweep.build(np.arange(grid.nfreq))
    chi = torch.empty((grid.nfreq, Dd.shape[1], Dd.shape[1]), dtype=torch.float64,
                      device=dev)
    tmp = torch.empty((1, sweep.M, sweep.M), dtype=torch.float64, device=dev)
    for k in range(grid.nfreq):
        chi[k] = _project(sweep.build([k], out=tmp)[0], Dd)
    return chi


def dyson_screening(chi0, transform=None, correlation=False):
    """W = [I - chi0]^-1 (space_time.py:956-968), through the Cholesky factor of
    the positive-definite I - chi0 (chi0 is negative semidefinite on the
    imaginary axis)."""
    c = chi0 if isinstance(chi0, torch.Tensor) else torch.as_tensor(np.asarray(chi0, float))
    if transform is not None:
        T = transform if isinstance(transform, torch.Tensor) else torch.as_tensor(
            np.asarray(transform, float), device=c.device)
        c = T.T @ c @ T
    n = c.shape[-1]
    A = torch.eye(n, dtype=c.dtype, device=c.device) - 0.5 * (c + c.T)
    L, info = torch.linalg.cholesky_ex(A)
    if int(info) != 0:
        raise np.linalg.LinAlgError('I - chi0 is not positive definite')
    W = torch.cholesky_inverse(L)
    if correlation:
        W.diagonal().sub_(1.0)
    return W.cpu().numpy() if not isinstance(chi0, torch.Tensor) else W


def static_screening_matrix(X, D, eps, nocc, grid=None, mu=None, comm=None,
                            tile_memory_gb=ISDF_TILE_GB, branches=None,
                            precision='fp64'):
    """W_aux = [1 - chi0(i w = 0)]^-1 in D's auxiliary basis
    (space_time.py:1003-1050).  Returns NumPy for NumPy inputs."""
    if comm is not None and comm.Get_size() > 1:
        raise NotImplementedError('one GPU: no tau split over ranks')
    grid = static_screening_grid(eps, nocc) if grid is None else grid
    chi0 = chi0_imaginary_frequency(X, D, eps, nocc, grid, mu=mu,
                                    precision=precision)[0]
    W = dyson_screening(chi0)
    numpy_in = not isinstance(X, torch.Tensor)
    return W.cpu().numpy() if numpy_in else W


def screened_interaction_grid(X, V, eps, nocc, grid=None, mu=None,
                              precision='fp32', chi0=None):
    """(W, chi0) on the grid: W = (I - V chi0(0))^-1 V computed as
    L (I - L^T chi0 L)^-1 L^T with L = chol(V) -- two FP64 Cholesky
    factorisations, a triangular solve and a symmetric product on the device;
    chi0(0) from the fused kernel at `precision`.  precision 'fp32': the
    products L^T chi0 L and Y^T Y (library GEMMs) in FP32 as well, widened
    to FP64 before I - K is factorized and for W; the factorizations and the
    solve stay FP64."""
    dev = _device()
    grid = static_screening_grid(eps, nocc) if grid is None else grid
    Vd = _dev(V, dev)
    if chi0 is None:
        chi0 = chi0_grid(X, eps, nocc, grid.tau_points, grid.cosft_wt[0], mu=mu,
                         precision=precision)
    L = torch.linalg.cholesky(Vd)
    if precision == 'fp32':
        L32 = L.float()
        K = (L32.T @ (chi0.float() @ L32)).double()
    else:
        K = L.T @ (chi0 @ L)
    n = K.shape[0]
    A = torch.eye(n, dtype=K.dtype, device=dev) - 0.5 * (K + K.T)
    C, info = torch.linalg.cholesky_ex(A)
    if int(info) != 0:
        raise np.linalg.LinAlgError('I - L^T chi0 L is not positive definite')
    # W = L A^-1 L^T = (C^-1 L^T)^T (C^-1 L^T)
    Y = solve_lower_blocked(C, L.T.contiguous())
    if precision == 'fp32':
        Y32 = Y.float()
        W = (Y32.T @ Y32).double()
    else:
        W = Y.T @ Y
    return 0.5 * (W + W.T), chi0


#: device memory for one batched Cholesky factorization: the stacked FP64
#: matrices and their factors (16 matrices of M = 6000 take 9.2 GB)
FACTOR_BATCH_BYTES = 16 * 2 ** 30


def factor_batch_size(n, count, budget=FACTOR_BATCH_BYTES, per_matrix=2):
    """How many n x n FP64 matrices (each with `per_matrix` buffers: the
    matrix, its factor, ...) one batched Cholesky takes, at most `count`.
    rocSOLVER's factorization of one matrix of a few thousand is
    latency-bound (22 ms at n = 6000); sixteen at once take 3.9 ms each."""
    return int(max(1, min(count, budget // (per_matrix * 8 * n * n))))


def solve_lower_blocked(C, B, block=1024):
    """C^-1 B for a lower-triangular C (n, n) and B (n, m): block forward
    substitution with each diagonal block inverted by one triangular solve
    against the identity, everything else GEMMs (a triangular solve with
    thousands of right-hand sides runs at a third of the GEMM rate, and
    rocBLAS's trsm workspace under torch is bounded)."""
    n = C.shape[0]
    out = torch.empty_like(B)
    for k in range(0, n, block):
        kb = min(block, n - k)
        eye = torch.eye(kb, dtype=C.dtype, device=C.device)
        dinv = torch.linalg.solve_triangular(C[k:k + kb, k:k + kb].contiguous(), eye,
                                             upper=False)
        rhs = B[k:k + kb]
        if k:
            rhs = torch.addmm(rhs, C[k:k + kb, :k], out[:k], alpha=-1.0)
        out[k:k + kb] = dinv @ rhs
    return out


def dielectric_log_determinant(one_minus_c, where, flags=None):
    """log|det(1 - c)| through a Cholesky factor (space_time.py:1073-1097);
    a matrix that is not positive definite falls to slogdet, warns
    `NonPositiveDielectric` when the determinant is not positive and records
    `where` in flags['nonpositive_det']."""
    A = one_minus_c
    L, info = torch.linalg.cholesky_ex(0.5 * (A + A.T))
    if int(info) == 0:
        return float(2.0 * torch.log(torch.diagonal(L)).sum())
    sign, logdet = torch.linalg.slogdet(A)
    if not float(sign) > 0:
        warnings.warn(f'det(1 - chi0) is not positive at {where} (sign {float(sign)})',
                      NonPositiveDielectric, stacklevel=2)
        if flags is not None:
            flags.setdefault('nonpositive_det', []).append(where)
    return float(logdet)


def rpa_correlation_energy_space_time(X, D, eps, nocc, grid, mu=None,
                                      tile_gb=ISDF_TILE_GB, screening=None,
                                      counter_term=None, comm=None,
                                      want_tape=False, flags=None,
                                      precision='fp64'):
    """E_c = (1/2 pi) sum_k w_k [log det(I - chi0_k) + Tr chi0_k]
    (space_time.py:1100-1183), chi0_k in D's auxiliary basis.  The
    rescaled-interaction and counter-term variants of the solvation routes are
    not part of this path."""
    if screening is not None or counter_term is not None or want_tape:
        raise NotImplementedError('screening / counter_term / want_tape are not built')
    if comm is not None and comm.Get_size() > 1:
        raise NotImplementedError('one GPU: no frequency split over ranks')
    dev = _device()
    sweep = FrequencySweep(X, eps, nocc, grid.tau_points, grid.cosft_wt, mu=mu,
                           precision=precision, device=dev)
    Dd = None if D is None else _dev(D, dev)
    mine = {}
    e_c = 0.0
    tmp = torch.empty((1, sweep.M, sweep.M), dtype=torch.float64, device=dev)
    for k in range(grid.nfreq):
        c = _project(sweep.build([k], out=tmp)[0], Dd)
        n = c.shape[0]
        logdet = dielectric_log_determinant(torch.eye(n, dtype=c.dtype, device=dev) - c,
                                            k, mine)
        e_c += grid.omega_weights[k] * (logdet + float(torch.trace(c)))
    if flags is not None and mine.get('nonpositive_det'):
        flags.setdefault('nonpositive_det', []).extend(mine['nonpositive_det'])
    return e_c / (2.0 * np.pi)


def rpa_correlation_energy_grid(X, V, eps, nocc, grid, mu=None, precision='fp32',
                                flags=None, timings=None):
    """E_c with the interaction on the grid: log det(I - L^T chi0 L) with
    L = chol(V), Tr(L^T chi0 L) = sum(chi0 o V); every proj(tau) built once
    and transformed to the frequencies a batch at a time (`FrequencySweep`),
    Cholesky log-dets in FP64.  precision 'fp32': the terms (where their
    weights allow) and the products L^T chi0 L in FP32 (library GEMMs, the
    result widened to FP64 before I - K is formed and factorized); 'fp64':
    everything in FP64.

    The matrices of up to `factor_batch_size` frequencies are factorized in
    one batched Cholesky; the log-dets and traces stay on the device until
    the end (no host synchronisation per frequency); a frequency whose
    I - L^T chi0 L is not positive definite is redone through
    `dielectric_log_determinant` (slogdet, warning, flags).  One stream:
    concurrent factorizations through torch's rocBLAS / rocSOLVER (several
    streams, one or several host threads) corrupted results at random.
    `timings`, a dict, receives the device time of the term builds and of
    the transforms (HIP events, waited for), their flops and the count of
    frequencies formed from FP64 terms."""
    dev = _device()
    if timings is not None:
        torch.cuda.synchronize()
        t_all = time.perf_counter()
    sweep = FrequencySweep(X, eps, nocc, grid.tau_points, grid.cosft_wt, mu=mu,
                           precision=precision, device=dev, profile=timings is not None)
    Vd = _dev(V, dev)
    L = torch.linalg.cholesky(Vd)
    n = L.shape[0]
    nf = grid.nfreq
    eye = torch.eye(n, dtype=torch.float64, device=dev)
    logdet = torch.zeros(nf, dtype=torch.float64, device=dev)
    trace = torch.zeros(nf, dtype=torch.float64, device=dev)
    info = torch.zeros(nf, dtype=torch.int32, device=dev)
    g = factor_batch_size(n, nf, per_matrix=3)
    chis = torch.empty((g, n, n), dtype=torch.float64, device=dev)
    A = torch.empty((g, n, n), dtype=torch.float64, device=dev)
    if precision == 'fp32':
        L32 = L.float()
        L32t = L32.T.contiguous()
        c32, Y32, K32 = (torch.empty((n, n), dtype=torch.float32, device=dev)
                         for _ in range(3))
    for k0 in range(0, nf, g):
        ks = list(range(k0, min(nf, k0 + g)))
        sweep.build(ks, out=chis[:len(ks)])
        for j, k in enumerate(ks):
            if precision == 'fp32':
                c32.copy_(chis[j])
                torch.matmul(c32, L32, out=Y32)
                torch.matmul(L32t, Y32, out=K32)
                K = K32.double()
            else:
                K = L.T @ (chis[j] @ L)
            trace[k] = torch.trace(K)
            torch.sub(eye, 0.5 * (K + K.T), out=A[j])
        C, inf = torch.linalg.cholesky_ex(A[:len(ks)])
        logdet[k0:k0 + len(ks)] = 2.0 * torch.log(torch.diagonal(C, dim1=-2, dim2=-1)).sum(-1)
        info[k0:k0 + len(ks)] = inf
        del C
    del A
    logdet_h, trace_h = logdet.cpu().numpy(), trace.cpu().numpy()
    mine = {}
    for k in np.flatnonzero(info.cpu().numpy()):
        chi = sweep.build([k], out=chis[:1])[0]
        K = L.T @ (chi @ L)
        logdet_h[k] = dielectric_log_determinant(eye - K, int(k), mine)
    e_c = float(np.dot(np.asarray(grid.omega_weights, float), logdet_h + trace_h))
    if flags is not None and mine.get('nonpositive_det'):
        flags.setdefault('nonpositive_det', []).extend(mine['nonpositive_det'])
    if timings is not None:
        torch.cuda.synchronize()
        timings.update(sweep.profile_read())
        timings.update({'seconds': time.perf_counter() - t_all, 'nfreq': nf,
                        'ntau': len(grid.tau_points),
                        'fp64_frequencies': int(sweep.use64.sum())})
    return e_c / (2.0 * np.pi)
