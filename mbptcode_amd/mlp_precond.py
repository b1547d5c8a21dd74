"""The learned Davidson preconditioner.

A 3 -> 256 -> 256 -> 1 SiLU network (BF16 weights, FP32 accumulation, the
forward, backward and Adam kernels in libmbpt_hip.so, mlp.hip) maps each pair's
[r_ia / rms(r), d_ia - theta, theta] -- d the pair energy on the Davidson's
preconditioner diagonal -- to t_ia / rms(r).  It stands in for the diagonal
correction r / (d - theta) (LR/davidson_shared.py:157-163, Solvers/davidson.py:
43-77): the correction applied is the Jacobi-Davidson one on that diagonal
plus rms(r) times the network's output, a learned remainder (the output
layer starts at zero, so an untrained network is the diagonal correction).
`run_davidson(preconditioner='learned')` falls back to the screened
Jacobi-Davidson correction alone when the residual stalls.

Training targets come from plain FP64 Davidson runs: for a Ritz vector u with
residual r, the exact eigenvector x of the root it approximates (the tracked
exact root of largest overlap) gives the ideal correction
t* = x / (u . x) - u, orthogonal to u and making u + t* parallel to x; the
network learns (t* - t_JD) / rms(r) on pairs drawn half uniformly and half in
proportion to |r|.
"""
import ctypes
import os
import time

import numpy as np
import torch

from . import _lib

HIDDEN = 256
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data',
                               'mlp_precond.npz')


def nparams():
    return int(_lib.lib().mbpt_mlp_nparams())


def init_params(seed=0, hidden=HIDDEN, zero_output=False):
    """He-style initial parameters, flattened FP32 (W1, b1, W2, b2, w3, b3);
    zero_output starts the output layer at 0 (the net adds nothing to the
    Jacobi-Davidson correction before it is trained)."""
    rng = np.random.default_rng(seed)
    w3 = (np.zeros(hidden) if zero_output
          else rng.normal(0, np.sqrt(1.0 / hidden), hidden))
    parts = [rng.normal(0, np.sqrt(2.0 / 3), (3, hidden)), np.zeros(hidden),
             rng.normal(0, np.sqrt(2.0 / hidden), (hidden, hidden)), np.zeros(hidden),
             w3, np.zeros(1)]
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)


def unflatten(flat, hidden=HIDDEN):
    """dict W1 (3, H), b1, W2 (H, H), b2, w3, b3 () of a flat parameter vector."""
    H = hidden
    o = np.cumsum([0, 3 * H, H, H * H, H, H, 1])
    f = np.asarray(flat)
    return {'W1': f[o[0]:o[1]].reshape(3, H), 'b1': f[o[1]:o[2]],
            'W2': f[o[2]:o[3]].reshape(H, H), 'b2': f[o[3]:o[4]],
            'w3': f[o[4]:o[5]], 'b3': f[o[5]:o[6]].reshape(())}


class LearnedPreconditioner:
    """The network on one GPU; `correction` is the Davidson hook."""

    def __init__(self, params=None, d=None, seed=0, device=None):
        self.device = (torch.device('cuda', torch.cuda.current_device())
                       if device is None else torch.device(device))
        self._lib = _lib.lib()
        flat = init_params(seed) if params is None else np.ascontiguousarray(params, np.float32)
        if flat.size != nparams():
            raise ValueError(f'{flat.size} parameters, the network has {nparams()}')
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self._lib.mbpt_mlp_create(
                ctypes.byref(h), flat.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                _lib.stream_ptr()), 'mbpt_mlp_create')
        self._h = h
        self.d = None if d is None else torch.as_tensor(
            np.asarray(d, float).ravel(), device=self.device)
        self.calls = 0
        #: the network predicts the remainder t - t_JD on top of the
        #: Jacobi-Davidson correction on the same diagonal (how it is trained)
        self.residual_of_jd = True

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            try:
                torch.cuda.synchronize(self.device)
            except Exception:
                pass
            self._lib.mbpt_mlp_destroy(h)
            self._h = None

    # -- parameters -------------------------------------------------------
    def params(self):
        out = np.empty(nparams(), np.float32)
        _lib.check(self._lib.mbpt_mlp_get_params(
            self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
            _lib.stream_ptr()), 'mbpt_mlp_get_params')
        return out

    def save(self, path=DEFAULT_WEIGHTS, **meta):
        np.savez(path, params=self.params(), residual_of_jd=np.asarray(self.residual_of_jd),
                 **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS, d=None):
        with np.load(path, allow_pickle=False) as f:
            net = cls(params=f['params'], d=d)
            if 'residual_of_jd' in f:
                net.residual_of_jd = bool(f['residual_of_jd'])
        return net

    # -- network ------------------------------------------------------------
    def forward(self, X):
        """net(X) for X (N, 3) float32 on the device."""
        X = X.to(self.device, torch.float32).contiguous()
        y = torch.empty(X.shape[0], dtype=torch.float32, device=self.device)
        _lib.check(self._lib.mbpt_mlp_forward(self._h, X.shape[0], _lib.dptr(X),
                                               _lib.dptr(y), _lib.stream_ptr()),
                   'mbpt_mlp_forward')
        return y

    def gradient(self, X, T):
        """(loss, flat gradient) of the mean-squared error."""
        X = X.to(self.device, torch.float32).contiguous()
        T = T.to(self.device, torch.float32).contiguous()
        g = torch.empty(nparams(), dtype=torch.float32, device=self.device)
        loss = ctypes.c_float()
        _lib.check(self._lib.mbpt_mlp_gradient(self._h, X.shape[0], _lib.dptr(X),
                                                _lib.dptr(T), _lib.dptr(g),
                                                ctypes.byref(loss), _lib.stream_ptr()),
                   'mbpt_mlp_gradient')
        return float(loss.value), g

    def train_step(self, X, T, lr):
        loss = ctypes.c_float()
        _lib.check(self._lib.mbpt_mlp_train_step(self._h, X.shape[0], _lib.dptr(X),
                                                  _lib.dptr(T), ctypes.c_float(lr),
                                                  ctypes.byref(loss), _lib.stream_ptr()),
                   'mbpt_mlp_train_step')
        return float(loss.value)

    def fit(self, X, T, steps=2000, batch=65536, lr=1e-3, lr_final=1e-4, seed=0,
            log_every=0):
        """Adam on minibatches drawn from (X (N, 3), T (N)) device tensors;
        returns the loss history (every step)."""
        X = X.to(self.device, torch.float32).contiguous()
        T = T.to(self.device, torch.float32).contiguous()
        gen = torch.Generator(device=self.device).manual_seed(seed)
        hist = []
        N = X.shape[0]
        for s in range(steps):
            a = lr * (lr_final / lr) ** (s / max(1, steps - 1))
            if N > batch:
                idx = torch.randint(0, N, (batch,), device=self.device, generator=gen)
                xb, tb = X[idx].contiguous(), T[idx].contiguous()
            else:
                xb, tb = X, T
            hist.append(self.train_step(xb, tb, a))
            if log_every and s % log_every == 0:
                print(f'[mlp] step {s} loss {hist[-1]:.4e}', flush=True)
        return hist

    # -- the Davidson hook -----------------------------------------------------
    def correction(self, R, U, theta, diag=None, jd=None):
        """Corrections T (n, k) for residual columns R (n, k) at Ritz values
        theta (k): rms(r) * net(r / rms(r), d - theta, theta).  d is the pair
        energy on the Davidson's preconditioner diagonal (`diag`, the screened
        d - (ii|W|aa) by default, the features the network was trained on),
        else the pair energies this object was built with.  A network trained
        as a remainder over Jacobi-Davidson gets that correction added (`jd`
        when the caller has it, else computed here)."""
        dvec = diag if diag is not None else self.d
        if dvec is None:
            raise ValueError('LearnedPreconditioner needs the pair energies d')
        dvec = dvec.reshape(-1).to(torch.float64).contiguous()
        n, k = R.shape
        R = R.contiguous()
        T = torch.empty_like(R)
        th = theta.to(torch.float64).contiguous()
        _lib.check(self._lib.mbpt_mlp_correction(self._h, n, k, _lib.dptr(R), k,
                                                  _lib.dptr(dvec), _lib.dptr(th),
                                                  _lib.dptr(T), k, _lib.stream_ptr()),
                   'mbpt_mlp_correction')
        self.calls += 1
        if self.residual_of_jd and U is not None:
            if jd is None:
                from .davidson import jacobi_davidson_correction
                jd = jacobi_davidson_correction(R, U, dvec, th)
            T += jd
        return T

    def correction_batched(self, R, U, theta, diag, npair, jd=None):
        """Corrections (B, P, k) for a batch of Davidson runs: R, U (B, P, k),
        theta (B, k), diag (B, P), npair (B,) real pairs per instance (the
        rest of each P block is padding and stays zero).  With the network
        trained as a remainder over Jacobi-Davidson, `jd` (the JD corrections,
        computed when not given) is added."""
        B, P, k = R.shape
        R = R.contiguous()
        T = torch.zeros_like(R)
        key = (R.device, tuple(int(n) for n in npair))
        if getattr(self, '_off_key', None) != key:
            off = np.concatenate([[0], np.cumsum(np.asarray(npair, np.int64))])
            self._off = torch.as_tensor(off.astype(np.int32), device=R.device)
            self._ntot = int(off[-1])
            self._off_key = key
        ntot = self._ntot
        d = diag.to(torch.float64).contiguous()
        th = theta.to(torch.float64).contiguous()
        _lib.check(self._lib.mbpt_mlp_correction_batched(
            self._h, B, P, k, _lib.dptr(self._off), ntot, _lib.dptr(R), k, _lib.dptr(d),
            _lib.dptr(th), _lib.dptr(T), k, _lib.stream_ptr()), 'mbpt_mlp_correction_batched')
        self.calls += 1
        if self.residual_of_jd and U is not None:
            if jd is None:
                from .batched import jd_correction_batched
                jd = jd_correction_batched(R, U, d, th)
            T += jd
        return T


# ---------------------------------------------------------------------------
# training data from plain FP64 Davidson runs
# ---------------------------------------------------------------------------

def harvest(inst, W_grid=None, pairs_per_root=256, seed=0, tol=1e-5, tight=1e-9,
            max_cycle=200, target='jd'):
    """(X (N, 3), T (N,)) float32 device tensors from one instance.

    A tight FP64 solve gives the exact tracked eigenvectors; a plain FP64
    solve at `tol` with the screened Jacobi-Davidson correction then records,
    every cycle and every unconverged root, the features of `pairs_per_root`
    pairs and the scaled target: target 'jd' the remainder
    (t* - t_JD) / rms(r) of the ideal correction over the Jacobi-Davidson one;
    target 'span' the part of the exact vector outside the current trial
    space, (x - V V^T x) / (u . x) / rms(r), the direction the next
    expansion should add.
    """
    from .block_actions import grid_block_action
    from .davidson import davidson, guess_indices, tracked_roots, jacobi_davidson_correction
    from .grids import static_screening_grid
    from .space_time import screened_interaction_grid
    if W_grid is None:
        g = static_screening_grid(inst.eps, inst.nocc, ntau=inst.ntau)
        W_grid, _ = screened_interaction_grid(inst.X, inst.V, inst.eps, inst.nocc, g,
                                              precision='fp32')
    act, dd = grid_block_action(inst.eps, inst.nocc, inst.X, W_grid, inst.V,
                                precision='fp64')
    op = act.operator
    no, nv = inst.nocc, inst.nvir
    npair = no * nv
    dev = op.device
    pdiag = op.screened_diagonal_device().reshape(-1)
    track = tracked_roots(inst.nroots, npair)
    order = guess_indices(dd, inst.nroots)
    x0 = torch.zeros((npair, len(order)), dtype=torch.float64, device=dev)
    x0[torch.as_tensor(order, device=dev), torch.arange(len(order), device=dev)] = 1.0

    def mm(V):
        k = V.shape[1]
        return op.apply(V.T.contiguous().reshape(k, no, nv)).reshape(k, npair).T

    exact = davidson(mm, pdiag, track, x0, tol=tight, max_cycle=max_cycle)
    Xe = exact.X                                        # (npair, track)
    gen = torch.Generator(device=dev).manual_seed(seed)
    feats, targs = [], []

    def record(cycle, V, R, U, theta):
        ov = (U.T @ Xe).abs()                           # (k, track)
        j = ov.argmax(dim=1)
        x = Xe[:, j]
        dots = (U * x).sum(0)
        ok = dots.abs() > 0.5
        if target == 'span':
            # the part of the exact vector the trial space V still lacks
            tstar = (x - V @ (V.T @ x)) / dots[None, :]
        else:
            tstar = x / dots[None, :] - U
        rms = R.pow(2).mean(0).sqrt().clamp_min(1e-300)
        k = R.shape[1]
        # half the pairs uniform, half in proportion to |r| (where the
        # correction matters)
        half = pairs_per_root // 2
        idx = torch.cat([
            torch.randint(0, npair, (half, k), device=dev, generator=gen),
            torch.multinomial(R.abs().T + 1e-300, pairs_per_root - half,
                              replacement=True, generator=gen).T], dim=0)
        jd = jacobi_davidson_correction(R, U, pdiag, theta)
        cols = torch.arange(k, device=dev)[None, :].expand_as(idx)
        f = torch.stack([R[idx, cols] / rms[None, :], pdiag[idx] - theta[None, :],
                         theta[None, :].expand_as(idx)], dim=-1)
        base = jd[idx, cols] if target == 'jd' else 0.0
        t = (tstar[idx, cols] - base) / rms[None, :]
        keep = ok[None, :].expand_as(idx)
        feats.append(f[keep].reshape(-1, 3).float())
        targs.append(t[keep].reshape(-1).float())

    davidson(mm, pdiag, track, x0, tol=tol, max_cycle=max_cycle, observe=record)
    if not feats:
        return (torch.empty((0, 3), dtype=torch.float32, device=dev),
                torch.empty(0, dtype=torch.float32, device=dev))
    return torch.cat(feats), torch.cat(targs)


def train_on_instances(instances, steps=2000, batch=65536, lr=2e-3, seed=0, clip=50.0,
                       log_every=0, target='jd'):
    """Harvest every instance, fit a fresh network, return (network, info).
    target: 'jd' (the network adds to the Jacobi-Davidson correction) or
    'span' (the network is the whole correction)."""
    t0 = time.time()
    Xs, Ts = [], []
    for i, inst in enumerate(instances):
        X, T = harvest(inst, seed=seed + i, target=target)
        Xs.append(X)
        Ts.append(T)
    X = torch.cat(Xs)
    T = torch.cat(Ts).clamp(-clip, clip)
    t_harvest = time.time() - t0
    net = LearnedPreconditioner(params=init_params(seed, zero_output=(target == 'jd')))
    net.residual_of_jd = target == 'jd'
    t0 = time.time()
    hist = net.fit(X, T, steps=steps, batch=batch, lr=lr, lr_final=lr / 20, seed=seed,
                   log_every=log_every)
    torch.cuda.synchronize()
    return net, {'rows': int(X.shape[0]), 'harvest_s': t_harvest,
                 'train_s': time.time() - t0, 'loss_first': hist[0], 'loss_last': hist[-1],
                 'target_var': float(T.var()), 'target': target}
