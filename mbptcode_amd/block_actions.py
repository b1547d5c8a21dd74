"""The Tamm-Dancoff ISDF BSE/Casida block action on the GPU.

Drop-in for `isdf_block_action` (src/SingleReference/LinearResponse/
block_actions.py:365) and its dispatcher `block_action` (:73): same arguments,
same `(apply_AB, diag_d)` return, same attributes on `apply_AB`
(`screened_diagonal`, `held_bytes`, `vector_bytes`, `scratch_bytes`,
`comm_clock`, `piece_clock`), same exceptions.  The arithmetic runs in
libmbpt_hip.so (sigma.hip); this module only converts the reference's factor
gauge into the grid one the kernels take and moves arrays.

Gauge: the reference's factors are (X_mo, D) with V = D D^T and a screened
interaction W_aux in D's auxiliary basis, so the exchange kernel on the grid
is Zt = D W_aux D^T (block_actions.py:275-294).  The kernels take Zt and V on
the grid directly; `grid_block_action` is that entry, and
`isdf_block_action` forms Zt and V once (FP64, on the device) and calls it.
"""
import ctypes
import time

import numpy as np
import torch

from . import _lib
from .constants import BYTES_PER_GB, ISDF_TILE_GB, KAPPA
from .orbital_spaces import bse_pair_diagonal

COMM_KINDS = ('lockstep', 'gather', 'reduce_grid', 'reduce_pairs')
ACTION_PIECES = ('gemm', 'hadamard', 'hartree', 'owner', 'batch')


class _CommClock:
    """Seconds in collectives (block_actions.py:44-53); a one-GPU action has
    none, so every entry stays 0."""
    __slots__ = ('seconds', 'by_kind')

    def __init__(self):
        self.seconds = 0.0
        self.by_kind = dict.fromkeys(COMM_KINDS, 0.0)


class _PieceClock:
    """Per-piece seconds (block_actions.py:56-70).  The pieces run fused in
    device kernels, so the whole device time of an application is booked
    under 'gemm' (measured with stream events when `timed` is set)."""
    __slots__ = ('by_piece',)

    def __init__(self):
        self.by_piece = dict.fromkeys(ACTION_PIECES, 0.0)


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError('the ISDF block action runs on a ROCm GPU; none is visible')
    return torch.device('cuda', torch.cuda.current_device())


def _as_device(a, dev):
    if isinstance(a, torch.Tensor):
        return a.to(device=dev, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)


def _add_counters(a, b):
    """The sum of two `profile_read` records (a may be None)."""
    if a is None:
        return b
    return {'ms': {k: a['ms'][k] + b['ms'][k] for k in b['ms']},
            'launches': a['launches'] + b['launches'],
            'exchange_flops': a['exchange_flops'] + b['exchange_flops']}


class SigmaOperator:
    """A z for batches of Tamm-Dancoff trial vectors, on one GPU.

    eps (norb,), X (M, norb) collocation, W (M, M) screened and V (M, M) bare
    interaction on the grid, kappa the Hartree prefactor, precision 'fp64'
    (the default, as the reference computes), 'fp32' or 'bf16x3' (the bulk
    GEMMs; d z and the output stay FP64).  workspace_bytes bounds the
    per-apply device workspace: a batch of trial vectors whose workspace
    would exceed it is applied in consecutive slices (None: no bound).
    """

    def __init__(self, eps, nocc, X, W, V, kappa=2.0, precision='fp64',
                 device=None, workspace_bytes=None):
        dev = _device() if device is None else torch.device(device)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        X = _as_device(X, dev)
        M, norb = X.shape
        if norb != len(eps):
            raise ValueError(f'X has {norb} orbitals, eps has {len(eps)}.')
        if not 0 < nocc < norb:
            raise ValueError(f'nocc={nocc} outside (0, {norb})')
        W = _as_device(W, dev)
        if W.shape != (M, M):
            raise ValueError(f'W is {tuple(W.shape)}, need ({M}, {M})')
        Vt = None
        if kappa:
            Vt = _as_device(V, dev)
            if Vt.shape != (M, M):
                raise ValueError(f'V is {tuple(Vt.shape)}, need ({M}, {M})')
        self.M, self.nocc, self.nvir = M, nocc, norb - nocc
        self.npair = self.nocc * self.nvir
        self.kappa = float(kappa)
        self.precision = precision
        self.device = dev
        self.diag_d = bse_pair_diagonal(eps, nocc)
        self._lib = _lib.lib()
        h = ctypes.c_void_p()
        _, eps_p = None, None
        eps_arr, eps_p = _lib.host_doubles(eps)
        with torch.cuda.device(dev):
            _lib.check(self._lib.mbpt_sigma_create(
                ctypes.byref(h), M, nocc, self.nvir, _lib.precision_code(precision),
                self.kappa, _lib.dptr(X), _lib.dptr(W),
                _lib.dptr(Vt) if Vt is not None else None, eps_p,
                _lib.stream_ptr()), 'mbpt_sigma_create')
            torch.cuda.current_stream().synchronize()
        self._h = h
        self.applied_vectors = 0
        self.applied_calls = 0
        self.workspace_bytes = workspace_bytes
        self._kcap = {}

    def vectors_per_apply(self, k, with_b=False):
        """Largest slice of a k-vector batch whose workspace fits
        `workspace_bytes` (at least 1)."""
        if self.workspace_bytes is None or k <= 1:
            return max(k, 1)
        key = (k, bool(with_b))
        if key not in self._kcap:
            fits = lambda j: self._lib.mbpt_sigma_workspace_bytes(self._h, j, int(with_b)) \
                <= self.workspace_bytes
            if fits(k):
                cap = k
            else:
                lo, hi = 1, k          # fits(hi) is False; find the largest j that fits
                while hi - lo > 1:
                    mid = (lo + hi) // 2
                    lo, hi = (mid, hi) if fits(mid) else (lo, mid)
                cap = lo
            self._kcap[key] = cap
        return self._kcap[key]

    def workspace_for(self, k, with_b=False):
        """Device bytes of the per-apply workspace an apply of k vectors uses."""
        return int(self._lib.mbpt_sigma_workspace_bytes(self._h, k, int(with_b)))

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            try:
                torch.cuda.synchronize(self.device)
            except Exception:
                pass
            self._lib.mbpt_sigma_destroy(h)
            self._h = None

    @property
    def held_bytes(self):
        return int(self._lib.mbpt_sigma_held_bytes(self._h))

    def apply(self, z, out=None):
        """A z for a device tensor z (k, nocc, nvir) or (k, npair), FP64,
        on the current stream; returns a fresh tensor of z's shape (or `out`)."""
        if z.dtype != torch.float64 or not z.is_cuda:
            raise ValueError('z must be a float64 device tensor')
        z = z.contiguous()
        k = z.numel() // self.npair
        if k * self.npair != z.numel():
            raise ValueError(f'z has {z.numel()} elements, not a multiple of '
                             f'the {self.npair} pairs')
        if out is None:
            out = torch.empty_like(z)
        zf, of = z.reshape(k, self.npair), out.view(k, self.npair)
        step = self.vectors_per_apply(k)
        for j in range(0, k, step):
            kj = min(step, k - j)
            _lib.check(self._lib.mbpt_sigma_apply(self._h, kj, _lib.dptr(zf[j:j + kj]),
                                                   _lib.dptr(of[j:j + kj]),
                                                   _lib.stream_ptr()),
                       'mbpt_sigma_apply')
        self.applied_vectors += k
        self.applied_calls += 1
        return out

    def apply_ab(self, z, out_a=None, out_b=None):
        """(A z, B z) of the full Casida action for a device tensor z (k,
        nocc, nvir) or (k, npair), FP64; fresh tensors of z's shape unless
        given."""
        if z.dtype != torch.float64 or not z.is_cuda:
            raise ValueError('z must be a float64 device tensor')
        z = z.contiguous()
        k = z.numel() // self.npair
        if k * self.npair != z.numel():
            raise ValueError(f'z has {z.numel()} elements, not a multiple of '
                             f'the {self.npair} pairs')
        out_a = torch.empty_like(z) if out_a is None else out_a
        out_b = torch.empty_like(z) if out_b is None else out_b
        zf = z.reshape(k, self.npair)
        af, bf = out_a.view(k, self.npair), out_b.view(k, self.npair)
        step = self.vectors_per_apply(k, with_b=True)
        for j in range(0, k, step):
            kj = min(step, k - j)
            _lib.check(self._lib.mbpt_sigma_apply_ab(
                self._h, kj, _lib.dptr(zf[j:j + kj]), _lib.dptr(af[j:j + kj]),
                _lib.dptr(bf[j:j + kj]), _lib.stream_ptr()), 'mbpt_sigma_apply_ab')
        self.applied_vectors += k
        self.applied_calls += 1
        return out_a, out_b

    def profile(self, enable=True):
        """Start (and zero) or stop per-stage device timing of `apply`."""
        _lib.check(self._lib.mbpt_sigma_profile(self._h, int(enable)),
                   'mbpt_sigma_profile')
        self._prof = {'on': bool(enable), 'held': None}

    def _profile_counters(self):
        ms = (ctypes.c_double * 4)()
        n = ctypes.c_long()
        fl = ctypes.c_double()
        _lib.check(self._lib.mbpt_sigma_profile_read(
            self._h, ms, ctypes.byref(n), ctypes.byref(fl)), 'mbpt_sigma_profile_read')
        return {'ms': dict(zip(('head', 'hartree', 'exchange', 'tail'), list(ms))),
                'launches': int(n.value), 'exchange_flops': float(fl.value)}

    def profile_pause(self):
        """Stop timing while keeping what was timed: applies that share the
        device with other work (a check running on a second stream) are
        left out of the figures.  `profile_resume` goes on."""
        st = getattr(self, '_prof', None)
        if not st or not st['on'] or st['held'] is not None:
            return
        st['held'] = self._profile_counters()
        _lib.check(self._lib.mbpt_sigma_profile(self._h, 0), 'mbpt_sigma_profile')

    def profile_resume(self):
        st = getattr(self, '_prof', None)
        if not st or st['held'] is None:
            return
        held = st['held']
        _lib.check(self._lib.mbpt_sigma_profile(self._h, 1), 'mbpt_sigma_profile')
        st['held'] = None
        st['base'] = _add_counters(st.get('base'), held)

    def profile_read(self):
        """{'ms': {head, hartree, exchange, tail}, 'launches', 'exchange_flops'}
        of the timed applies since `profile(True)`."""
        st = getattr(self, '_prof', None) or {}
        now = st['held'] if st.get('held') is not None else self._profile_counters()
        return _add_counters(st.get('base'), now)

    def screened_diagonal_device(self):
        """d - (ii|W|aa), (nocc, nvir) FP64 on the device; formed once per
        operator (W and X do not change) and kept."""
        cached = getattr(self, '_sdiag', None)
        if cached is not None:
            return cached
        out = torch.empty((self.nocc, self.nvir), dtype=torch.float64,
                          device=self.device)
        _lib.check(self._lib.mbpt_sigma_screened_diag(self._h, _lib.dptr(out),
                                                       _lib.stream_ptr()),
                   'mbpt_sigma_screened_diag')
        self._sdiag = out
        return out


def grid_block_action(eps, nocc, X, W, V, kappa=2.0, precision='fp64',
                      device=None, tda=True, tile_memory_gb=None):
    """(apply_AB, diag_d) from grid-gauge W and V: apply_AB(z) -> (A z, None)
    with z (k, nocc, nvir), NumPy or device FP64; with tda=False the full
    Casida action, (A z, B z).  precision 'fp64' (default), 'fp32' or
    'bf16x3'; tile_memory_gb bounds the per-apply workspace (None: no bound)."""
    ws = None if tile_memory_gb is None else int(tile_memory_gb * BYTES_PER_GB)
    op = SigmaOperator(eps, nocc, X, W, V, kappa=kappa, precision=precision,
                       device=device, workspace_bytes=ws)
    return _wrap(op, tda=tda)


def _wrap(op, tda=True):
    comm_clock, piece_clock = _CommClock(), _PieceClock()

    def apply_AB(z):
        t0 = time.time()
        host = not isinstance(z, torch.Tensor)
        if host:
            z = np.asarray(z, dtype=np.float64)
            zz = torch.as_tensor(np.ascontiguousarray(z), device=op.device)
        else:
            zz = z.to(device=op.device, dtype=torch.float64)
        zk = zz.reshape(-1, op.nocc, op.nvir)
        if tda:
            outs = (op.apply(zk), None)
        else:
            outs = op.apply_ab(zk)
        outs = tuple(None if o is None else
                     (o.cpu().numpy() if host else o).reshape(z.shape) for o in outs)
        piece_clock.by_piece['gemm'] += time.time() - t0
        return outs

    def screened_diagonal():
        return op.screened_diagonal_device().cpu().numpy()

    apply_AB.operator = op
    apply_AB.comm_clock = comm_clock
    apply_AB.piece_clock = piece_clock
    apply_AB.screened_diagonal = screened_diagonal
    apply_AB.held_bytes = op.held_bytes
    apply_AB.vector_bytes = int(8 * op.npair)
    apply_AB.scratch_bytes = int(8 * (2 * op.nocc * op.M + op.npair))
    return apply_AB, op.diag_d


def isdf_block_action(lr_solver, nocc, lBSE, W_aux, isdf_factors,
                      tile_memory_gb=ISDF_TILE_GB, spin='singlet', comm=None,
                      tda=False, precision='fp64'):
    """The reference's `isdf_block_action` (block_actions.py:365) on the GPU.

    lr_solver: anything with `.eps` and `.spin_mode`.  isdf_factors: (X_mo, D)
    or (X_mo, D, ...).  W_aux: screened interaction in D's gauge, None for
    TDHF (bare exchange), unused when lBSE is False (RPA).  tda=False gives
    the full Casida action, apply_AB(z) -> (A z, B z) (the B block,
    block_actions.py:632-652, in the handle's precision).  comm: one GPU
    only, so None or a one-rank communicator.  tile_memory_gb: the bound on
    the action's per-apply scratch, as in the reference (whose row tiles it
    sizes, block_actions.py:526-543); here trial vectors are applied in slices
    whose device workspace fits it.  precision: 'fp64' (the default, the
    reference's arithmetic), or 'fp32' / 'bf16x3' for the bulk GEMMs as an
    explicit opt-in.
    """
    if getattr(lr_solver, 'spin_mode', 'restricted') == 'unrestricted':
        raise NotImplementedError('The ISDF Davidson route is restricted-spin only.')
    if comm is not None and comm.Get_size() > 1:
        raise NotImplementedError('this action runs on one GPU; a rank split '
                                  'of W is not built')
    if spin not in KAPPA:
        raise ValueError(f"spin {spin!r}: one of {', '.join(sorted(KAPPA))}")
    if not tile_memory_gb > 0:
        raise ValueError(f'tile_memory_gb={tile_memory_gb} must be positive')
    X_mo, D = isdf_factors[0], isdf_factors[1]
    eps = np.asarray(lr_solver.eps, float)
    nmo = np.shape(X_mo)[1]
    naux = np.shape(D)[1]
    if nmo != len(eps):
        raise ValueError(f'isdf_factors X has {nmo} orbitals, lr_solver.eps has '
                         f'{len(eps)}.')
    if W_aux is not None and np.shape(W_aux)[0] != naux:
        raise ValueError(f'W_aux is ({np.shape(W_aux)[0]}, ...) but D has {naux} '
                         'auxiliary functions; they are not from the same fit '
                         '(see isdf_bse_factors).')
    dev = _device()
    Dt = _as_device(D, dev)
    V = Dt @ Dt.T
    if lBSE:
        if W_aux is not None:
            Wa = _as_device(W_aux, dev)
            asym = (Wa - Wa.T).abs().max().item()
            if asym > 1e-10 * Wa.abs().max().item():
                raise ValueError(f'W_aux is not symmetric (max asymmetry {asym:.2e}); '
                                 'the B block below relies on it.')
            Zt = Dt @ (Wa @ Dt.T)
            Zt = 0.5 * (Zt + Zt.T)
        else:
            Zt = V
    else:
        Zt = torch.zeros_like(V)
    return grid_block_action(eps, nocc, X_mo, Zt, V, kappa=KAPPA[spin],
                             precision=precision, device=dev, tda=tda,
                             tile_memory_gb=tile_memory_gb)


def block_action(lr_solver, nocc, polarizability, W_aux, isdf_factors,
                 spin='singlet', comm=None, tda=False, precision='fp64'):
    """Mode dispatch (block_actions.py:73-96): 'RPA', 'TDHF' or 'BSE' on ISDF
    factors.  The DF route (isdf_factors None) needs pyscf's three-index
    factor and is not part of this package."""
    if spin not in KAPPA:
        raise ValueError(f"spin {spin!r}: one of {', '.join(sorted(KAPPA))}")
    mode = polarizability.upper()
    if mode == 'RPA':
        lBSE, w = False, None
    elif mode == 'TDHF':
        lBSE, w = True, None
    elif mode == 'BSE':
        if W_aux is None:
            raise ValueError("polarizability='BSE' requires W_aux (see "
                             'static_screening_matrix).')
        lBSE, w = True, W_aux
    else:
        raise ValueError(f"Unknown polarizability '{polarizability}'; choose "
                         "'RPA', 'TDHF', or 'BSE'.")
    if isdf_factors is None:
        raise NotImplementedError('the DF block action needs pyscf integrals; '
                                  'pass isdf_factors=(X_mo, D)')
    return isdf_block_action(lr_solver, nocc, lBSE, w, isdf_factors, spin=spin,
                             comm=comm, tda=tda, precision=precision)
