"""Many independent Tamm-Dancoff BSE problems solved together on one GPU.

The reference solves one ISDF BSE problem at a time: each `apply_AB`
closure (src/SingleReference/LinearResponse/block_actions.py:563-661) and
each Davidson (LR/davidson.py:462, :631) serves one instance.  Here a batch of
instances of uneven size (n_occ, n_vir, M) advances in lockstep:

* `BatchedSigma` -- the sigma-vectors of every instance in one pass of
  libmbpt_hip's batched kernels (batched.hip: one launch per stage over all
  instances, driven by device work tables), FP32 or FP64 MFMA;
* `SigmaGraph` -- that pass captured once in a hipGraph and replayed each
  cycle; per-cycle inputs (trial vectors and live-vector counts) sit in
  fixed device buffers;
* `autotune` -- the work decomposition (grid rows per exchange item, K chunks
  of the tail and Hartree products, persistent grid size) picked by timing
  the captured pass, cached as JSON per batch signature;
* `davidson_batched` -- block Davidson for all instances at once, FP64
  subspace on the device: trial spaces (B, P, cap), Rayleigh matrices
  (B, cap, cap), the m x m eigensolves in one batched Jacobi launch
  (`syev_batched`), one small device-to-host copy per cycle (Ritz values and
  residual norms).  Converged instances stop adding vectors and stop costing
  sigma-vectors (their live count drops to zero);
* `solve_batched` -- the drop-in driver: the reference's per-instance
  unit-vector guess and tracked roots, the screened Jacobi-Davidson
  correction (or the learned one), a lower-precision pass followed by the
  FP64 recheck and polish of every root.

Vectors of instance b use the first npair[b] = nocc[b] nvir[b] entries of a
row of length P = max npair (occupied index slow); the rest is zero.
"""
import ctypes
import hashlib
import json
import os
import time

import numpy as np
import torch

from . import _lib
from .constants import KAPPA
from .davidson import guess_indices, lindep_for, tracked_roots

#: Diagonal given to unused trial-space slots so their Ritz values sort last.
_SLOT_SHIFT = 1e10
#: Pair energy given to the padding so the diagonal corrections vanish there.
_PAD_DIAG = 1e6
MAX_SPACE = 64

DEFAULT_PLAN = dict(rows_per_split=1024, tail_kchunk=512, hartree_kchunk=512, grid_cap=0,
                    bucket_width=16)
PLAN_SPACE = dict(rows_per_split=(256, 512, 1024, 2048, 4096),
                  tail_kchunk=(128, 256, 512, 1024, 4096),
                  hartree_kchunk=(256, 1024, 4096),
                  grid_cap=(0, 2048, 4096, 8192),
                  bucket_width=(16, 32))
SHIPPED_TUNING = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data',
                              'batch_tune.json')


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError('the batched BSE solver runs on a ROCm GPU; none is visible')
    return torch.device('cuda', torch.cuda.current_device())


# ---------------------------------------------------------------------------
# batched small eigensolver
# ---------------------------------------------------------------------------

def syev_batched(A, max_sweeps=30, scale=1.0):
    """(w, V) of the symmetric parts of A (B, m, m) FP64: ascending eigenvalues
    (B, m) and eigenvectors as columns (B, m, m).  On the device through the
    batched Jacobi kernel (m <= 64; `scale` the size of the spectrum, below
    1e-15 scale an off-diagonal entry counts as zero); elsewhere
    torch.linalg.eigh."""
    if not A.is_cuda:
        return torch.linalg.eigh(0.5 * (A + A.transpose(1, 2)))
    B, m, m2 = A.shape
    if m != m2:
        raise ValueError(f'syev_batched: square matrices expected, got {tuple(A.shape)}')
    if A.stride(2) != 1:
        A = A.contiguous()
    w = torch.empty((B, m), dtype=torch.float64, device=A.device)
    V = torch.empty((B, m, m), dtype=torch.float64, device=A.device)
    _lib.check(_lib.lib().mbpt_syev_batched_f64(
        B, m, ctypes.c_void_p(A.data_ptr()), A.stride(1), A.stride(0),
        ctypes.c_void_p(w.data_ptr()), m, ctypes.c_void_p(V.data_ptr()), m, m * m,
        max_sweeps, float(scale), _lib.stream_ptr()), 'mbpt_syev_batched_f64')
    return w, V


# ---------------------------------------------------------------------------
# batched subspace products over the real rows of each instance
# ---------------------------------------------------------------------------

_work_cache = {}


def _rows_ok(t):
    return t if t.stride(-1) == 1 else t.contiguous()


def bnn(A, S, rows, out=None, subtract=False):
    """A_b S_b for every instance (A (B, P, m), S (B, m, c)) over the first
    rows[b] rows; the result's padding rows are zero.  With `subtract`,
    out -= A S in place.  Device tensors go through the batched FP64 kernel,
    host tensors through torch.bmm."""
    B, P, m = A.shape
    c = S.shape[2]
    if not A.is_cuda or rows is None or m > 64 or m == 0:
        R = torch.bmm(A, S) if m else torch.zeros((B, P, c), dtype=A.dtype, device=A.device)
        if subtract:
            out -= R
            return out
        if out is None:
            return R
        out.copy_(R)
        return out
    A, S = _rows_ok(A), _rows_ok(S)
    if out is None:
        out = torch.empty((B, P, c), dtype=torch.float64, device=A.device)
    if out.stride(-1) != 1:
        raise ValueError('bnn: the output rows must be unit-stride')
    _lib.check(_lib.lib().mbpt_bnn_f64(
        B, P, m, c, ctypes.c_void_p(rows.data_ptr()), ctypes.c_void_p(A.data_ptr()),
        A.stride(0), A.stride(1), ctypes.c_void_p(S.data_ptr()), S.stride(0), S.stride(1),
        ctypes.c_void_p(out.data_ptr()), out.stride(0), out.stride(1), int(subtract),
        _lib.stream_ptr()), 'mbpt_bnn_f64')
    return out


def btn(A, Bm, rows, out=None):
    """A_b^T B_b for every instance (A (B, P, a), Bm (B, P, c)) over the
    first rows[b] rows, into `out` (B, a, c; any row stride) or a new tensor."""
    B, P, a = A.shape
    c = Bm.shape[2]
    if not A.is_cuda or rows is None or a > 64 or c > 64:
        R = torch.bmm(A.transpose(1, 2), Bm)
        if out is None:
            return R
        out.copy_(R)
        return out
    A, Bm = _rows_ok(A), _rows_ok(Bm)
    if out is None:
        out = torch.empty((B, a, c), dtype=torch.float64, device=A.device)
    if out.stride(-1) != 1:
        raise ValueError('btn: the output rows must be unit-stride')
    L = _lib.lib()
    need = int(L.mbpt_btn_f64_work(B, P, a, c))
    key = A.device
    w = _work_cache.get(key)
    if w is None or w.numel() < need:
        w = torch.empty(max(need, 1), dtype=torch.float64, device=A.device)
        _work_cache[key] = w
    _lib.check(L.mbpt_btn_f64(
        B, P, a, c, ctypes.c_void_p(rows.data_ptr()), ctypes.c_void_p(A.data_ptr()),
        A.stride(0), A.stride(1), ctypes.c_void_p(Bm.data_ptr()), Bm.stride(0), Bm.stride(1),
        ctypes.c_void_p(out.data_ptr()), out.stride(0), out.stride(1),
        ctypes.c_void_p(w.data_ptr()), w.numel(), _lib.stream_ptr()), 'mbpt_btn_f64')
    return out


# ---------------------------------------------------------------------------
# batched sigma-vectors
# ---------------------------------------------------------------------------

class BatchedSigma:
    """A z for every instance of a batch, one pass of device launches.

    X, W, V: per-instance device (or host) FP64 arrays, (M, norb), (M, M),
    (M, M); eps per-instance orbital energies; nocc per instance; kmax trial
    vectors per instance per pass.
    """

    def __init__(self, eps, nocc, X, W, V, kappa=KAPPA['singlet'], precision='fp32',
                 kmax=8, device=None):
        dev = _device() if device is None else torch.device(device)
        n = len(X)
        if not (len(eps) == len(nocc) == len(W) == n and (kappa == 0 or len(V) == n)):
            raise ValueError('eps, nocc, X, W and V must list the same instances')
        L = _lib.lib()
        self.device = dev
        self.n = n
        self.kmax = int(kmax)
        self.precision = precision
        self.kappa = float(kappa)
        keep = []

        def dv(a):
            t = (a.to(device=dev, dtype=torch.float64) if isinstance(a, torch.Tensor)
                 else torch.as_tensor(np.asarray(a, np.float64), device=dev)).contiguous()
            keep.append(t)
            return t.data_ptr()

        Ms, nos, nvs, eps_all = [], [], [], []
        xp, wp, vp = (ctypes.c_void_p * n)(), (ctypes.c_void_p * n)(), (ctypes.c_void_p * n)()
        for b in range(n):
            e = np.asarray(eps[b], np.float64).ravel()
            M, norb = X[b].shape
            no = int(nocc[b])
            if norb != e.size or not 0 < no < norb:
                raise ValueError(f'instance {b}: X has {norb} orbitals, eps {e.size}, nocc {no}')
            if tuple(W[b].shape) != (M, M):
                raise ValueError(f'instance {b}: W must be ({M}, {M})')
            Ms.append(M); nos.append(no); nvs.append(norb - no); eps_all.append(e)
            xp[b] = dv(X[b])
            wp[b] = dv(W[b])
            vp[b] = dv(V[b]) if kappa != 0 else None
        self.M = np.asarray(Ms, np.int32)
        self.nocc = np.asarray(nos, np.int32)
        self.nvir = np.asarray(nvs, np.int32)
        self.npair = self.nocc.astype(np.int64) * self.nvir
        eps_c, eps_p = _lib.host_doubles(np.concatenate(eps_all))
        self._eps = [np.asarray(e) for e in eps_all]
        ip = ctypes.POINTER(ctypes.c_int)
        h = ctypes.c_void_p()
        _lib.check(L.mbpt_batch_create(
            ctypes.byref(h), n, self.M.ctypes.data_as(ip), self.nocc.ctypes.data_as(ip),
            self.nvir.ctypes.data_as(ip), _lib.precision_code(precision), self.kappa,
            xp, wp, vp, eps_p, self.kmax, _lib.stream_ptr()), 'mbpt_batch_create')
        torch.cuda.synchronize(dev)
        del keep
        self._h = h
        self.P = int(L.mbpt_batch_npair_max(h))
        self.plan_args = dict(DEFAULT_PLAN)
        self.generation = 0   # bumped by plan(): captured graphs go stale

    # -- properties ----------------------------------------------------------
    @property
    def held_bytes(self):
        return int(_lib.lib().mbpt_batch_held_bytes(self._h))

    def pair_mask(self):
        """(n, P) bool: the real pairs of each instance."""
        idx = torch.arange(self.P, device=self.device)
        return idx[None, :] < torch.as_tensor(self.npair, device=self.device)[:, None]

    def pair_energies(self):
        """(n, P) d_ia = eps_a - eps_i, padding set to a large value (a fresh
        tensor; the operator keeps its own copy)."""
        if getattr(self, '_dpair', None) is None:
            d = np.full((self.n, self.P), _PAD_DIAG)
            for b in range(self.n):
                e, no = self._eps[b], int(self.nocc[b])
                d[b, :self.npair[b]] = (e[no:][None, :] - e[:no][:, None]).ravel()
            self._dpair = torch.as_tensor(d, device=self.device)
        return self._dpair.clone()

    def screened_diagonal(self):
        """(n, P) d - (ii|W|aa) in the handle's precision, padding large (a
        fresh tensor).  Computed on the first call and kept: it depends on
        the operator only, not on the plan."""
        if getattr(self, '_sdiag', None) is None:
            out = torch.full((self.n, self.P), _PAD_DIAG, dtype=torch.float64,
                             device=self.device)
            _lib.check(_lib.lib().mbpt_batch_screened_diag(self._h, _lib.dptr(out),
                                                           _lib.stream_ptr()),
                       'mbpt_batch_screened_diag')
            self._sdiag = out
        return self._sdiag.clone()

    def sigma_graph(self, graph=True):
        """The operator's SigmaGraph (one per `graph` setting), kept across
        solves so that a pass is captured once per plan, not once per solve.
        Solves that share it must not overlap."""
        graphs = self.__dict__.setdefault('_graphs', {})
        if graph not in graphs:
            graphs[graph] = SigmaGraph(self, graph=graph)
        return graphs[graph]

    # -- work -----------------------------------------------------------------
    def plan(self, rows_per_split, tail_kchunk, hartree_kchunk, grid_cap=0, bucket_width=16):
        _lib.check(_lib.lib().mbpt_batch_plan(self._h, int(rows_per_split), int(tail_kchunk),
                                              int(hartree_kchunk), int(grid_cap),
                                              int(bucket_width), _lib.stream_ptr()),
                   'mbpt_batch_plan')
        self.plan_args = dict(rows_per_split=int(rows_per_split), tail_kchunk=int(tail_kchunk),
                              hartree_kchunk=int(hartree_kchunk), grid_cap=int(grid_cap),
                              bucket_width=int(bucket_width))
        self.generation += 1

    def apply_into(self, z, out, kact=None):
        """out (n, kmax, P) <- A z (n, kmax, P); rows at and past kact[b]
        (an int32 device tensor, or None for all) are left untouched."""
        shape = (self.n, self.kmax, self.P)
        for name, t in (('z', z), ('out', out)):
            if tuple(t.shape) != shape or t.dtype != torch.float64 or not t.is_contiguous():
                raise ValueError(f'{name} must be contiguous float64 {shape}')
        kp = None
        if kact is not None:
            if kact.dtype != torch.int32 or kact.numel() != self.n or not kact.is_cuda:
                raise ValueError('kact must be an int32 device tensor of n entries')
            kp = ctypes.c_void_p(kact.data_ptr())
        _lib.check(_lib.lib().mbpt_batch_apply(self._h, _lib.dptr(z), _lib.dptr(out), kp,
                                               _lib.stream_ptr()), 'mbpt_batch_apply')
        return out

    def apply(self, z, kact=None):
        out = torch.zeros_like(z)
        return self.apply_into(z, out, kact)

    def close(self):
        if getattr(self, '_h', None):
            _lib.lib().mbpt_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SigmaGraph:
    """One batched pass captured in a hipGraph on fixed buffers `z`, `out`
    and `kact`: fill z (and kact), call `run()`, read out.  Rows of `out` at
    and past kact[b] hold stale values."""

    def __init__(self, op, graph=True):
        self.op = op
        shape = (op.n, op.kmax, op.P)
        self.z = torch.zeros(shape, dtype=torch.float64, device=op.device)
        self.out = torch.zeros_like(self.z)
        self.kact = torch.full((op.n,), op.kmax, dtype=torch.int32, device=op.device)
        self.use_graph = graph
        self.graph = None
        self.generation = None

    def _capture(self):
        op = self.op
        s = torch.cuda.Stream(device=op.device)
        s.wait_stream(torch.cuda.current_stream(op.device))
        with torch.cuda.stream(s):
            op.apply_into(self.z, self.out, self.kact)   # warm-up outside the capture
        torch.cuda.current_stream(op.device).wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            op.apply_into(self.z, self.out, self.kact)
        self.graph = g
        self.generation = op.generation

    def run(self):
        if not self.use_graph:
            return self.op.apply_into(self.z, self.out, self.kact)
        if self.graph is None or self.generation != self.op.generation:
            self._capture()
        self.graph.replay()
        return self.out


# ---------------------------------------------------------------------------
# autotuner
# ---------------------------------------------------------------------------

def _signature(op):
    h = hashlib.sha1()
    h.update(np.stack([op.M, op.nocc, op.nvir]).astype(np.int64).tobytes())
    name = torch.cuda.get_device_name(op.device) if op.device.type == 'cuda' else 'cpu'
    return f'{name}|{op.precision}|k{op.kmax}|n{op.n}|{h.hexdigest()[:16]}'


def _cache_path():
    return os.environ.get('MBPT_TUNE_CACHE',
                          os.path.join(os.path.expanduser('~'), '.cache', 'mbptcode_amd',
                                       'batch_tune.json'))


def _load_cache(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def time_pass(op, reps=5, kact=None):
    """Median milliseconds of one captured batched pass on random vectors."""
    sg = SigmaGraph(op)
    gen = torch.Generator(device=op.device).manual_seed(0)
    sg.z.normal_(generator=gen)
    sg.z.mul_(op.pair_mask()[:, None, :])
    if kact is not None:
        sg.kact.copy_(kact)
    sg.run()
    ts = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        sg.run()
        b.record()
        b.synchronize()
        ts.append(a.elapsed_time(b))
    return float(np.median(ts))


def autotune(op, space=None, reps=5, cache=True, sweeps=2, log=None):
    """Pick the plan of `op` by coordinate descent over `space` (each knob in
    turn, the others held), timing the captured pass; the winner is applied
    and, with `cache`, stored under the batch's signature.  A cached or
    shipped entry is used without timing.  Returns (plan, info)."""
    space = PLAN_SPACE if space is None else space
    key = _signature(op)
    path = _cache_path()
    if cache:
        for p in (path, SHIPPED_TUNING):
            hit = _load_cache(p).get(key)
            if hit:
                op.plan(**hit['plan'])
                return dict(hit['plan']), {'source': p, 'ms': hit.get('ms'), 'key': key}
    best = dict(DEFAULT_PLAN)
    op.plan(**best)
    best_ms = time_pass(op, reps)
    trials = 1
    for _ in range(sweeps):
        changed = False
        for knob, values in space.items():
            for v in values:
                if v == best[knob]:
                    continue
                cand = dict(best, **{knob: v})
                op.plan(**cand)
                ms = time_pass(op, reps)
                trials += 1
                if log:
                    log(f'[tune] {cand} {ms:.3f} ms')
                if ms < best_ms * 0.98:
                    best, best_ms, changed = cand, ms, True
        if not changed:
            break
    op.plan(**best)
    info = {'source': 'timed', 'ms': best_ms, 'trials': trials, 'key': key}
    if cache:
        data = _load_cache(path)
        data[key] = {'plan': best, 'ms': best_ms}
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path, 'w') as f:
                json.dump(data, f, indent=1, sort_keys=True)
        except OSError:
            pass
    return best, info


# ---------------------------------------------------------------------------
# batched Davidson
# ---------------------------------------------------------------------------

def jd_correction_batched(R, U, diag, theta, floor=1e-8):
    """Jacobi-Davidson corrections t = K^-1 r - alpha K^-1 u for R, U (B, P, r),
    diag (B, P), theta (B, r); K = diag - theta floored at `floor`
    (Solvers/davidson.py:43-77), per instance and root."""
    den = diag[:, :, None] - theta[:, None, :]
    den = torch.where(den.abs() < floor, torch.full_like(den, floor), den)
    Kr = R / den
    Ku = U / den
    uKu = (U * Ku).sum(1)
    uKr = (U * Kr).sum(1)
    alpha = torch.where(uKu.abs() > 1e-300, uKr / uKu, torch.zeros_like(uKu))
    return Kr - alpha[:, None, :] * Ku


def orthonormalize_batched(T, V, drop, rows=None):
    """Columns of T (B, P, c) projected off the orthonormal columns V (B, P, m)
    twice and orthonormalised among themselves (eigendecomposition of the
    Gram matrix), per instance.  Columns whose norm after projection is at or
    below `drop` relative to before are discarded.  Returns (T', live): the
    kept directions first, zero columns after, and their count (B,) int32."""
    T = T / T.norm(dim=1, keepdim=True).clamp_min(1e-300)
    if V is not None and V.shape[2]:
        for _ in range(2):
            bnn(V, btn(V, T, rows), rows, out=T, subtract=True)
    keep = T.norm(dim=1) > drop
    T = T * keep[:, None, :]
    G = btn(T, T, rows)
    s, U = syev_batched(G)
    smax = s.max(dim=1, keepdim=True).values
    good = (s > 1e-10 * smax) & (s > 1e-200)
    scale = torch.where(good, s.clamp_min(1e-300).rsqrt(), torch.zeros_like(s))
    C = (U * scale[:, None, :]).flip(2).contiguous()   # largest Gram eigenvalues first
    return bnn(T, C, rows), good.sum(1).to(torch.int32)


class BatchedResult:
    __slots__ = ('e', 'X', 'conv', 'rnorm', 'cycles', 'vectors', 'trace', 'seconds',
                 'sync_s')

    def __init__(self, **kw):
        for k in self.__slots__:
            setattr(self, k, kw.get(k))


def davidson_batched(apply, diag, nroots, x0, live0, tol, max_cycle=100,
                     max_space=MAX_SPACE, precond=None, lindep=None, ax0=None,
                     observe=None, rows=None):
    """Block Davidson for the `nroots` lowest eigenpairs of B symmetric
    operators at once, FP64 on the device.

    apply(T (B, P, c), live (B,) int32) -> A T (B, P, c), zero in the
    columns at and past live[b].  diag (B, P) the correction diagonal.
    x0 (B, P, c0) orthonormal guess columns, live0 (B,) how many are real;
    ax0 optionally their image.  precond(R, U, theta, cycle) -> corrections
    (B, P, nroots); default Jacobi-Davidson on `diag`.  The space collapses
    onto the Ritz vectors when it would exceed max_space (<= 64).  rows
    (B,) int32 on the device: real pairs per instance (the subspace products
    skip the padding); None for all P.
    Returns a BatchedResult with e (B, nroots) and rnorm host arrays,
    X (B, P, nroots) device.
    """
    dev = diag.device
    B, P = diag.shape
    r = int(nroots)
    max_space = int(min(max_space, MAX_SPACE))
    if x0.shape[2] > max_space or r > max_space:
        raise ValueError('guess or root count exceeds the trial-space limit')
    lindep = lindep_for(tol) if lindep is None else lindep
    drop = float(np.sqrt(lindep))
    if precond is None:
        def precond(R, U, theta, cycle):
            return jd_correction_batched(R, U, diag, theta)
    t0 = time.time()
    # the Rayleigh matrices' eigenvalues are bounded by the diagonal's size
    dscale = float(torch.where(diag.abs() < _PAD_DIAG, diag.abs(),
                               torch.zeros_like(diag)).max().clamp_min(1.0))
    cap = max_space
    Vb = torch.zeros((B, P, cap), dtype=torch.float64, device=dev)
    AVb = torch.zeros_like(Vb)
    H = torch.zeros((B, cap, cap), dtype=torch.float64, device=dev)
    valid = torch.zeros((B, cap), dtype=torch.bool, device=dev)
    m = 0
    vectors = 0

    def append(T, AT, live):
        nonlocal m
        c = T.shape[2]
        Vb[:, :, m:m + c] = T
        AVb[:, :, m:m + c] = AT
        btn(Vb[:, :, :m + c], AT, rows, out=H[:, :m + c, m:m + c])
        if m:
            btn(T, AVb[:, :, :m], rows, out=H[:, m:m + c, :m])
        valid[:, m:m + c] = torch.arange(c, device=dev)[None, :] < live[:, None]
        m += c

    live0 = live0.to(device=dev, dtype=torch.int32)
    append(x0, apply(x0, live0) if ax0 is None else ax0, live0)
    vectors += int(live0.sum())
    trace = []
    sync_s = 0.0
    cycles = 0
    for cycle in range(max_cycle):
        cycles = cycle + 1
        Hm = H[:, :m, :m] + torch.diag_embed(_SLOT_SHIFT * (~valid[:, :m]).double())
        w_all, Y = syev_batched(Hm, scale=dscale)
        y = Y[:, :, :r]
        w = w_all[:, :r].contiguous()
        X = bnn(Vb[:, :, :m], y, rows)
        AX = bnn(AVb[:, :, :m], y, rows)
        R = AX - X * w[:, None, :]
        rn = R.norm(dim=1)
        ts = time.time()
        host = torch.cat([w, rn], dim=1).cpu().numpy()
        sync_s += time.time() - ts
        w_h, rn_h = host[:, :r], host[:, r:]
        conv = rn_h <= tol
        trace.append({'cycle': cycle, 'space': m, 'max_residual': float(rn_h.max()),
                      'converged': int(conv.sum()), 'instances_done': int(conv.all(1).sum())})
        if conv.all():
            break
        if observe is not None:
            observe(cycle, Vb[:, :, :m], R, X, w, conv)
        open_ = torch.as_tensor(~conv, device=dev).to(torch.float64)
        Tc = precond(R, X, w, cycle)
        Tc = Tc * open_.repeat(1, Tc.shape[2] // r)[:, None, :]
        T, live = orthonormalize_batched(Tc, Vb[:, :, :m], drop, rows)
        ts = time.time()
        cmax = int(live.max())
        sync_s += time.time() - ts
        if cmax == 0:
            break
        T = T[:, :, :cmax]
        if m + T.shape[2] > max_space:
            Hc = torch.bmm(torch.bmm(y.transpose(1, 2), H[:, :m, :m]), y)
            Vb[:, :, :r] = X
            AVb[:, :, :r] = AX
            H.zero_()
            H[:, :r, :r] = Hc
            valid.zero_()
            valid[:, :r] = True
            m = r
        AT = apply(T, live)
        vectors += int(live.sum())
        append(T, AT, live)
    else:
        # cycle budget ran out after an append: one more Rayleigh-Ritz step
        Hm = H[:, :m, :m] + torch.diag_embed(_SLOT_SHIFT * (~valid[:, :m]).double())
        w_all, Y = syev_batched(Hm, scale=dscale)
        y = Y[:, :, :r]
        w = w_all[:, :r].contiguous()
        X = bnn(Vb[:, :, :m], y, rows)
        AX = bnn(AVb[:, :, :m], y, rows)
        rn = (AX - X * w[:, None, :]).norm(dim=1)
        w_h, rn_h = w.cpu().numpy(), rn.cpu().numpy()
        conv = rn_h <= tol
    return BatchedResult(e=w_h, X=X, conv=conv, rnorm=rn_h, cycles=cycles, vectors=vectors,
                         trace=trace, seconds=time.time() - t0, sync_s=sync_s)


def _apply_via(sg):
    """apply(T (B, P, c), live) through a SigmaGraph, kmax columns per replay."""
    kmax = sg.op.kmax

    def apply(T, live):
        c = T.shape[2]
        out = torch.empty_like(T)
        for c0 in range(0, c, kmax):
            c1 = min(c, c0 + kmax)
            sg.z[:, :c1 - c0].copy_(T[:, :, c0:c1].transpose(1, 2))
            if c1 - c0 < kmax:
                sg.z[:, c1 - c0:].zero_()
            sg.kact.copy_((live - c0).clamp(0, c1 - c0))
            out[:, :, c0:c1] = sg.run()[:, :c1 - c0].transpose(1, 2)
        mask = (torch.arange(c, device=T.device)[None, :] < live[:, None]).to(torch.float64)
        return out * mask[:, None, :]
    return apply


def batched_guess(diag_host, npair, nroots, P, device):
    """Unit-vector guesses per instance (guess_indices on each instance's
    pair energies): x0 (B, P, c0) and live counts (B,)."""
    idx = [np.asarray(guess_indices(diag_host[b][:npair[b]], nroots), np.int64)
           for b in range(len(npair))]
    c0 = max(len(i) for i in idx)
    x0 = torch.zeros((len(npair), P, c0), dtype=torch.float64, device=device)
    # one scatter for the whole batch: (instance, pair, column) of every unit entry
    where = np.stack([np.concatenate([np.full(len(i), b) for b, i in enumerate(idx)]),
                      np.concatenate(idx),
                      np.concatenate([np.arange(len(i)) for i in idx])])
    b_, p_, c_ = torch.as_tensor(where, device=device)
    x0[b_, p_, c_] = 1.0
    return x0, torch.as_tensor([len(i) for i in idx], dtype=torch.int32, device=device)


def solve_batched(op, nroots, tol, max_cycle=100, preconditioner='screened', refine_op=None,
                  max_space=MAX_SPACE, graph=True, learned=None, learned_until=1000.0,
                  learned_mode='augment', pdiag=None, timings=None):
    """The `nroots` lowest TDA roots of every instance of `op` (a
    BatchedSigma).  refine_op: an FP64 BatchedSigma of the same batch; the
    roots found with `op` are then re-checked (A x - theta x in FP64) and
    iterated on it until every FP64 residual meets `tol`.

    Returns (e (B, nroots) host, X (B, P, nroots) device, info)."""
    t_all = time.time()
    dev = op.device
    dpair = op.pair_energies()
    d_host = dpair.cpu().numpy()
    track = max(tracked_roots(nroots, int(n)) for n in op.npair)
    if pdiag is None:
        pdiag = dpair if preconditioner == 'bare' else op.screened_diagonal()
    x0, live0 = batched_guess(d_host, op.npair, nroots, op.P, dev)
    apply = _apply_via(op.sigma_graph(graph))
    precond = None
    if preconditioner == 'learned':
        if learned is None:
            raise ValueError("preconditioner='learned' needs `learned`")

        def precond(R, U, theta, cycle):
            # the learned direction while the residuals are large; the
            # Jacobi-Davidson one alone near convergence, where the network's
            # BF16 output noise would dominate the useful correction
            jd = jd_correction_batched(R, U, pdiag, theta)
            if learned_until is not None and float(R.norm(dim=1).max()) < learned_until * tol:
                return jd
            t = learned.correction_batched(R, U, theta, pdiag, op.npair, jd)
            return torch.cat([jd, t], dim=2) if learned_mode == 'augment' else t
    rows = torch.as_tensor(np.asarray(op.npair, np.int32), device=dev)
    res = davidson_batched(apply, pdiag, track, x0, live0, tol, max_cycle=max_cycle,
                           max_space=max_space, precond=precond, rows=rows)
    info = {'cycles': res.cycles, 'vectors': res.vectors, 'trace': res.trace,
            'sync_s': res.sync_s, 'track': track}
    e, X, rn = res.e, res.X, res.rnorm
    if refine_op is not None:
        apply64 = _apply_via(refine_op.sigma_graph(graph))
        live = torch.full((op.n,), track, dtype=torch.int32, device=dev)
        AX = apply64(X, live)
        theta = (X * AX).sum(1)
        rn64 = (AX - X * theta[:, None, :]).norm(dim=1).cpu().numpy()
        info['rnorm_fp64_before'] = float(rn64.max())
        if not (rn64 <= tol).all():
            res2 = davidson_batched(apply64, pdiag, track, X, live, tol, max_cycle=max_cycle,
                                    max_space=max_space, ax0=AX, rows=rows)
            e, X, rn = res2.e, res2.X, res2.rnorm
            info['refine_cycles'] = res2.cycles
            info['refine_vectors'] = res2.vectors
        else:
            e, rn = theta.cpu().numpy(), rn64
            info['refine_cycles'] = 0
        info['rnorm_fp64_after'] = float(np.max(rn))
    info['converged'] = bool((rn[:, :nroots] <= tol).all())
    info['seconds'] = time.time() - t_all
    if timings is not None:
        timings.update(info)
    return np.asarray(e)[:, :nroots], X[:, :, :nroots], info
