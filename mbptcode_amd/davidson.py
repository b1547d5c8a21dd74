"""The Tamm-Dancoff BSE Davidson on the GPU.

Drop-ins for the reference's Hermitian Davidson stack:

* `solve_symmetric` (src/Solvers/davidson.py:179) -- lowest eigenpairs of a
  symmetric matrix-free operator, Jacobi-Davidson preconditioner (:43-77),
  linear-dependence threshold scaled to the tolerance (:156-176);
* `run_davidson` / `_tda_davidson_root` (LR/davidson.py:631, :462) -- the
  TDA loop on an `apply_AB`: unit-vector guess (`guess_indices`,
  davidson_shared.py:19-54), tracked roots above the requested ones
  (`tracked_roots` :146-154), 'bare' / 'screened' preconditioner diagonal
  (:125-143), unconfirmed-root bound (:538-546);
* `solve_casida_davidson` (LR/davidson.py:252) and `check_residual_floor`
  (:428-450); with `tda=False` the full Casida roots come from the paired
  loop in `casida.py` (`_davidson_root` :998, `casida_rows`).

Everything the iteration does is FP64 and stays on the device: the trial
space V and its image AV, the Rayleigh-Ritz matrix, residuals, corrections
and their orthogonalisation.  Only the m x m Ritz matrix and the residual
norms cross to the host once per cycle (the m x m eigensolve is done in
LAPACK there).  The operator may apply A in FP64, FP32 or split BF16; with
`refine=True` the roots converged on a lower-precision operator are
re-checked with an FP64 operator (A x - theta x recomputed in FP64) and
iterated on it until the FP64 residual meets the tolerance.
"""
import contextlib
import time
import warnings

import numpy as np
import torch

from .subspace import (Ranges, block_scale, jd_correction, nn_ranges, ritz_eigh, ritz_max_m,
                       ritz_residual, tn, tn_ranges)
from .constants import (DAVIDSON_DENOMINATOR_FLOOR, DAVIDSON_FLOOR_EPS_MULTIPLE,
                        DAVIDSON_GUESS_DEG_TOL, DAVIDSON_GUESS_FACTOR,
                        DAVIDSON_LINDEP_CEILING, DAVIDSON_LINDEP_FLOOR,
                        DAVIDSON_MAX_CYCLE, DAVIDSON_PRECONDITIONER,
                        DAVIDSON_PRECONDITIONERS, DAVIDSON_TOL_RESIDUAL,
                        DAVIDSON_EXTRA_ROOTS_DIVISOR, DAVIDSON_EXTRA_ROOTS_FLOOR,
                        DAVIDSON_COMPLETION_SEED, DAVIDSON_FOLLOW_OVERLAP_TIE,
                        DAVIDSON_NULL_DIRECTION, BSE_DENSE_MAX_NOV, DAVIDSON_DEVICE_RITZ_MIN)


_BLAS = {}
_SIDE = {}


def _side_stream(dev):
    """The second stream of a device for the early FP64 check, made once and
    kept: the caching allocator holds freed blocks per stream, so a stream
    made anew for every solve would meet its first allocations (the check's
    operands, tens of MB) with fresh device allocations each time."""
    key = dev.index if dev.index is not None else torch.cuda.current_device()
    s = _SIDE.get(key)
    if s is None:
        s = _SIDE[key] = torch.cuda.Stream(device=dev)
    return s


def _host_lapack():
    """Context running host LAPACK on one thread: the Ritz and Gram
    eigensolves here are at most a few hundred wide, where a single thread
    is ~20% faster than the BLAS pool's default."""
    ctl = _BLAS.get('ctl')
    if ctl is None:
        try:
            from threadpoolctl import ThreadpoolController
            ctl = ThreadpoolController()
        except ImportError:
            ctl = False
        _BLAS['ctl'] = ctl
    if ctl is False:
        return contextlib.nullcontext()
    return ctl.limit(limits=1, user_api='blas')


class DavidsonNotConverged(RuntimeError):
    """Roots left above the residual tolerance with refuse_unconverged set
    (src/Base/step_failed.py:109)."""


class CasidaInstability(RuntimeError):
    """The Casida problem has no physical roots: (A - B) or omega^2 is not
    positive on the pair space (the dense solver's refusal)."""


class UnstableReference(CasidaInstability):
    """The Casida problem refused for an unstable reference: min eig(A - B)
    <= 0 measured (LR/davidson.py:232-241).  `min_eig_amb` is the measured
    minimum, or None."""

    def __init__(self, message, min_eig_amb=None):
        super().__init__(message)
        self.min_eig_amb = min_eig_amb


class CasidaBreakdown(RuntimeError):
    """A projected (A - B) block that is not positive definite, with the
    min eig(A - B) the probe measured on the whole pair space, or None where
    the probe did not converge (LR/davidson.py:221-229)."""

    def __init__(self, message, min_eig_amb):
        super().__init__(message)
        self.min_eig_amb = min_eig_amb


def unstable_reference(amb, when='before the solve returned its roots'):
    """The UnstableReference for min eig(A - B) = amb <= 0
    (LR/davidson.py:1282-1293)."""
    return UnstableReference(
        f'BSE refused {when}: min eig(A-B) = {amb:.6f} Ha <= 0 -- the mean-field '
        'reference is singlet/triplet unstable and the Casida omega^2 reduction is '
        'invalid there; stabilise the reference (check that the SCF is its lowest '
        'solution) or solve the non-Hermitian problem', amb)


def casida_breakdown(apply_AB, diag_d, conv_tol):
    """The exception for a projected (A - B) block that is not positive
    definite (LR/davidson.py:1243-1279): the cold (A - B) probe on the same
    action tells a stable reference's numerical breakdown (CasidaBreakdown,
    min eig > 0) from an unstable reference (UnstableReference); a probe that
    does not converge gives a CasidaBreakdown with min_eig_amb None."""
    from .probe import AmbProbeNotConverged, lowest_amb_from_action
    try:
        amb = float(lowest_amb_from_action(apply_AB, diag_d)[0])
        detail = f'measured min eig(A-B) = {amb:.6f} Ha'
    except AmbProbeNotConverged:
        amb = None
        detail = ('min eig(A-B) probe did not converge; smallest orbital-energy '
                  f'difference {np.min(diag_d):.4f} Ha')
    if amb is not None and amb > 0:
        return CasidaBreakdown(
            'Casida Davidson broke down numerically: the projected (A-B) block is '
            f'not positive definite although (A-B) itself is ({detail}); the trial '
            f'space lost accuracy as conv_tol ({conv_tol:g} Ha) neared round-off '
            f'({residual_floor(diag_d):.1e} Ha here). Loosen conv_tol.', amb)
    if amb is not None:
        return unstable_reference(amb)
    return CasidaBreakdown(
        'Casida Davidson failed: the projected (A-B) block is not positive '
        f'definite, the signature of an unstable mean-field reference ({detail}); '
        'the omega^2 reduction is invalid in this regime.', amb)


def probe_after_solve(apply_AB, diag_d, omega, X, Y, probe, stats=None, comm=None):
    """min eig(A - B) on the action a Casida solve just ran on
    (LR/davidson.py:356-373): the converged probe from 1/d for probe=True, the
    sign alone for probe='sign', tried first on the span of the roots' X - Y.
    A value <= 0 refuses the roots (UnstableReference).  Returns the value."""
    from .probe import lowest_amb_from_action
    sign_only = isinstance(probe, str) and probe.lower() == 'sign'
    order = np.argsort(omega, kind='stable')
    amb = float(lowest_amb_from_action(
        apply_AB, diag_d, sign_only=sign_only, stats=stats, comm=comm,
        start=(np.asarray(X) - np.asarray(Y))[:, order] if sign_only else None)[0])
    if amb <= 0:
        raise unstable_reference(amb)
    return amb


# ---------------------------------------------------------------------------
# shared pieces (LR/davidson_shared.py)
# ---------------------------------------------------------------------------

def guess_indices(diag_d, nroots, guess_factor=DAVIDSON_GUESS_FACTOR,
                  deg_tol=DAVIDSON_GUESS_DEG_TOL, counters=None):
    """Pairs carrying the unit-vector guess: max(guess_factor k, k + 4) lowest,
    extended through a degenerate cluster up to max(4k, 20)
    (davidson_shared.py:19-54)."""
    d = np.asarray(diag_d, float).ravel()
    order = np.argsort(d)
    n = min(max(guess_factor * nroots, nroots + 4), d.size)
    cut = d[order[n - 1]] + deg_tol
    n_max = min(max(4 * nroots, 20, n), d.size)
    while n < n_max and d[order[n]] <= cut:
        n += 1
    if n == n_max < d.size and d[order[n]] <= cut:
        if counters is not None:
            counters['guess_truncated'] = True
        warnings.warn(f'Davidson guess truncated at {n} vectors inside a cluster '
                      f'of pair energies degenerate to {deg_tol:g} Ha',
                      RuntimeWarning, stacklevel=3)
    return order[:n]


def tracked_roots(nroots, n_pair, extra=None):
    """nroots + max(2, nroots // 2), at most the pair space (:146-154)."""
    if extra is None:
        extra = max(DAVIDSON_EXTRA_ROOTS_FLOOR, nroots // DAVIDSON_EXTRA_ROOTS_DIVISOR)
    return min(nroots + int(extra), n_pair)


def lindep_for(tol_residual):
    """Squared-norm linear-dependence threshold: clamp(1e-2 tol^2, 1e-22, 1e-14)
    (Solvers/davidson.py:156-176)."""
    return float(max(DAVIDSON_LINDEP_FLOOR,
                     min(DAVIDSON_LINDEP_CEILING, 1e-2 * tol_residual ** 2)))


def residual_floor(diag_d):
    """1e3 eps_FP64 max|d| (LR/davidson.py:453-457)."""
    return DAVIDSON_FLOOR_EPS_MULTIPLE * np.finfo(float).eps * float(np.abs(diag_d).max())


def check_residual_floor(conv_tol, diag_d):
    """ValueError when conv_tol is below the FP64 residual floor (:428-450)."""
    floor = residual_floor(diag_d)
    if not conv_tol >= floor:
        raise ValueError(f'Davidson conv_tol={conv_tol:g} is below the residual '
                         f'this pair space can reach in double precision, '
                         f'{floor:.1e} Ha; ask for at least that.')
    return floor


def check_preconditioner(preconditioner):
    if preconditioner not in DAVIDSON_PRECONDITIONERS:
        raise ValueError(f'preconditioner={preconditioner!r}; choose one of '
                         f"{', '.join(map(repr, DAVIDSON_PRECONDITIONERS))}.")


def jacobi_davidson_correction(R, U, diag, theta, floor=DAVIDSON_DENOMINATOR_FLOOR):
    """Columns t = K^-1 r - alpha K^-1 u, alpha = u.K^-1 r / u.K^-1 u,
    K = diag - theta floored at `floor` in magnitude (Solvers/davidson.py:43-77),
    for a block of residuals R and Ritz vectors U (n, m) on the device."""
    den = diag[:, None] - theta[None, :]
    den = torch.where(den.abs() < floor, torch.full_like(den, floor), den)
    Kr = R / den
    Ku = U / den
    uKu = (U * Ku).sum(0)
    uKr = (U * Kr).sum(0)
    alpha = torch.where(uKu.abs() > 1e-300, uKr / uKu, torch.zeros_like(uKu))
    return Kr - alpha[None, :] * Ku


# ---------------------------------------------------------------------------
# the iteration
# ---------------------------------------------------------------------------

def _to_host(t):
    """A device tensor as a NumPy array, through page-locked memory (the
    copy of the returned eigenvectors runs at the link's rate)."""
    if not t.is_cuda:
        return t.numpy()
    out = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    out.copy_(t)
    return out.numpy()


def _tn(A, B):
    """A^T B: on the device through the split-contraction FP64 kernel
    (the subspace reductions over the pair index), else plain torch."""
    if A.is_cuda:
        return tn(A, B)
    return A.T @ B


def _orthonormal_new(T, V, drop):
    """T projected off the orthonormal columns V (twice) and orthonormalised
    among itself by the eigendecomposition of its Gram matrix (solved on the
    host: the matrix is at most a few dozen wide).  Columns whose norm after
    the projection is below `drop` relative to before are discarded.  FP64."""
    T = T / T.norm(dim=0).clamp_min(1e-300)
    if V is not None and V.shape[1]:
        T = T - V @ _tn(V, T)
        T = T - V @ _tn(V, T)
    G = _tn(T, T).cpu().numpy()
    keep = np.flatnonzero(np.sqrt(np.clip(np.diag(G), 0, None)) > drop)
    if keep.size == 0:
        return T[:, :0]
    G = G[np.ix_(keep, keep)]
    with _host_lapack():
        s, U = np.linalg.eigh(G)
    good = s > 1e-10 * s.max()
    C = np.zeros((T.shape[1], int(good.sum())))
    C[keep] = U[:, good] / np.sqrt(s[good])
    return T @ torch.as_tensor(C, device=T.device)


class DavidsonResult:
    __slots__ = ('e', 'X', 'conv', 'rnorm', 'cycles', 'vectors', 'seconds',
                 'trace', 'stalled', 'exhausted')

    def __init__(self, **kw):
        for k in self.__slots__:
            setattr(self, k, kw.get(k))


# The iteration runs on device tensors through dvd.hip; the same steps on
# CPU tensors (plain torch) let its host logic be tested without a GPU.

def _cols(k):
    """Column blocks of at most 64 (the width of the device update kernels)."""
    return [(a, min(64, k - a)) for a in range(0, k, 64)]


def _tnr(A, B, out=None):
    """A^T B for two `Ranges`."""
    if A.buf.is_cuda:
        return tn_ranges(A, B, out=out)
    r = A.columns().T @ B.columns()
    if out is None:
        return r
    out.copy_(r)
    return out


def _nn(A, Y, C, c_off, alpha=1.0, beta=0.0):
    """C[:, c_off : c_off + c] = beta C + alpha A Y, in blocks of 64 columns."""
    if not A.buf.is_cuda:
        tgt = C[:, c_off:c_off + Y.shape[1]]
        r = alpha * (A.columns() @ Y)
        tgt.copy_(r if beta == 0.0 else beta * tgt + r)
        return
    for a, c in _cols(Y.shape[1]):
        nn_ranges(A, Y[:, a:a + c].contiguous() if Y.shape[1] > 64 else Y, C, c_off + a,
                  alpha, beta)


def _ritz(VA, m, cap, Y, theta):
    """(X, R) = (V Y, AV Y - X diag(theta)) for the m columns of V and AV in
    the trial-space buffer VA (n, 2 cap): (n, k) each."""
    n, k = VA.shape[0], Y.shape[1]
    if not VA.is_cuda:
        X = VA[:, :m] @ Y
        return X, VA[:, cap:cap + m] @ Y - X * theta[None, :]
    X = torch.empty((n, k), dtype=torch.float64, device=VA.device)
    R = torch.empty_like(X)
    for a, c in _cols(k):
        if k <= 64:
            ritz_residual(VA, m, 0, cap, Y, theta, X, R)
        else:
            Xc = torch.empty((n, c), dtype=torch.float64, device=VA.device)
            Rc = torch.empty_like(Xc)
            ritz_residual(VA, m, 0, cap, Y[:, a:a + c].contiguous(),
                          theta[a:a + c].contiguous(), Xc, Rc)
            X[:, a:a + c], R[:, a:a + c] = Xc, Rc
    return X, R


def _jd(U, R, diag, theta):
    """The Jacobi-Davidson corrections of the columns of R (n, k)."""
    if not R.is_cuda:
        return jacobi_davidson_correction(R, U, diag, theta)
    out = torch.empty_like(R)
    jd_correction(U, R, diag, theta, out, floor=DAVIDSON_DENOMINATOR_FLOOR)
    return out


def davidson(matmat, diag, nroots, x0, tol=DAVIDSON_TOL_RESIDUAL,
             max_cycle=DAVIDSON_MAX_CYCLE, max_space=None, precond=None,
             lindep=None, stall_cycles=None, on_stall=None, observe=None, phases=None,
             keep=None, pick=None, early=None, early_roots=0, precond_roots=False,
             ritz_min=DAVIDSON_DEVICE_RITZ_MIN):
    """Block Davidson for the `nroots` lowest eigenpairs of a symmetric
    operator, on device tensors.

    matmat(V (n, m)) -> A V (n, m), FP64 device tensors.  diag (n,) device.
    x0 (n, m0) guess columns.  precond(R, U, theta, cycle) -> corrections; the
    default is the Jacobi-Davidson one on `diag` (one fused device pass).
    When the space would exceed `max_space` it collapses onto the current
    Ritz vectors.  `on_stall` is called when the largest unconverged residual
    has not fallen by 10x over `stall_cycles` cycles (the learned
    preconditioner's fallback hook).

    The trial space V and its image AV are the two column halves of one
    device buffer, so a product with [V | AV] or [V | T] is one pass over
    the pair index (dvd.hip).  The Rayleigh matrix H = V^T A V (kept whole:
    a lower-precision operator is not exactly symmetric) and G = (AV)^T AV
    grow block by block on the device.  Per cycle the host receives H and G
    once (Ritz pairs by LAPACK, and the residual norms
    |A x - theta x|^2 = y^T G y - 2 theta y^T H y + theta^2 for x = V y) and
    the new block's projections once (its Gram eigensolve); everything else
    stays on the device.  The corrections are projected off V twice
    (classical Gram-Schmidt) and orthonormalised among themselves by the
    eigendecomposition of their Gram matrix; columns left below `drop` of
    their norm are discarded.  The returned residual norms are recomputed
    explicitly from A X - X theta.
    `phases`, a dict, receives the seconds of each part of the cycle (the
    device is synchronised at every boundary, so only for profiling).
    `keep`: Ritz vectors kept at a collapse (thick restart; default nroots).
    `pick(w, v, nroots, envs) -> (w, v, order)` selects the Ritz pairs to
    converge instead of the lowest (root following, `overlap_pick`): envs
    holds 'ref_coeff', V^T span (m, r), when the pick carries a `span`
    (n, r), else 'xs', the trial vectors as host arrays.
    precond_roots: call precond(R, U, theta, cycle, roots, rnorm) with the
    tracked roots' indices of the columns (ascending) and their residual
    norms from the Ritz step (a host array).
    ritz_min: the Rayleigh matrices from this size on (up to 192, with at
    most 64 Ritz pairs wanted and no `pick`) are solved on the device
    (ritz.hip, the lowest pairs only); smaller ones by host LAPACK, and so
    are all once `early` has been called: its check then keeps the device
    busy, and the one-workgroup solver would wait for it while the host
    solve runs beside it.
    `early(X, w)` is called once, at the first cycle whose Ritz step finds
    the leading `early_roots` pairs converged while a later tracked pair is
    not, with those Ritz vectors (n, early_roots) on the device and their
    values: a caller can start checking them while the iteration goes on.
    """
    last = [time.perf_counter()]

    def mark(name):
        if phases is not None:
            torch.cuda.synchronize()
            now = time.perf_counter()
            phases[name] = phases.get(name, 0.0) + now - last[0]
            last[0] = now
    dev = diag.device
    n = diag.numel()
    nroots = max(1, min(int(nroots), n))
    if max_space is None:
        max_space = max(10 * nroots + 30, 60)
    max_space = int(min(max(max_space, 2 * nroots), n))
    lindep = lindep_for(tol) if lindep is None else lindep
    drop = float(np.sqrt(lindep))
    use_jd = precond is None
    t0 = time.time()
    T0 = _orthonormal_new(x0, None, 1e-12)
    cap = max_space + max(T0.shape[1], 2 * nroots)
    VA = torch.empty((n, 2 * cap), dtype=torch.float64, device=dev)    # [V | AV]
    HG = torch.zeros((cap, 2 * cap), dtype=torch.float64, device=dev)  # [H | G]
    m = T0.shape[1]
    VA[:, :m] = T0
    VA[:, cap:cap + m] = matmat(T0)
    P = _tnr(Ranges(VA, (0, m), (cap, m)), Ranges(VA, (cap, m)))
    HG[:m, :m] = P[:m]
    HG[:m, cap:cap + m] = P[m:]
    vectors = m
    span = getattr(pick, 'span', None)
    if span is not None:
        span = torch.as_tensor(np.asarray(span, float).reshape(n, -1), device=dev)

    def ritz_select(evals, evecs, k):
        """(w, y, order): the k Ritz pairs the iteration converges."""
        if pick is None:
            return evals[:k], evecs[:, :k], np.arange(k)
        if span is not None:
            envs = {'ref_coeff': _tn(VA[:, :m], span).cpu().numpy()}
        else:
            envs = {'xs': list(VA[:, :m].T.cpu().numpy())}
        return pick(evals, evecs, k, envs)

    trace = []
    conv = np.zeros(nroots, bool)
    w = rn = None
    hist = []
    stalled = exhausted = False
    side_busy = False
    cycles = 0
    y = None
    y_rows = m
    mark('setup')
    for cycle in range(max_cycle):
        cycles = cycle + 1
        hg = HG[:m].cpu().numpy()
        Hns, G = hg[:, :m], hg[:, cap:cap + m]
        mark('rr_transfer')
        G = 0.5 * (G + G.T)
        k = min(nroots, m)
        kk = min(m, max(k, keep or k))
        if (pick is None and VA.is_cuda and ritz_min <= m <= ritz_max_m()
                and kk <= 64 and not side_busy):
            # large subspaces: the Ritz pairs from the device eigensolver
            Yd, wq = ritz_eigh(HG[:m, :m], HG[:m, cap:cap + m], kk)
            wq = wq.cpu().numpy()
            evals, evecs = wq[:kk], Yd.cpu().numpy()
            w, y, order = evals[:k], evecs[:, :k], np.arange(k)
            sq, xax = wq[kk:kk + k], wq[2 * kk:2 * kk + k]
        else:
            H = 0.5 * (Hns + Hns.T)
            with _host_lapack():
                evals, evecs = np.linalg.eigh(H)
            w, y, order = ritz_select(evals, evecs, k)
            sq = ((G @ y) * y).sum(0)
            xax = ((Hns @ y) * y).sum(0)
        y_rows = m
        # |A x - w x|^2 = |A x|^2 - 2 w x^T A x + w^2 (|x| = 1), exact for a
        # non-symmetric A as well
        rn = np.sqrt(np.clip(sq - 2.0 * w * xax + w ** 2, 0.0, None))
        # y^T G y - theta^2 cancels: below ~sqrt(1e3 eps) |theta| it is noise,
        # so residuals that small (and a tolerance that asks for them) are
        # recomputed explicitly from A x - theta x.
        floor = np.sqrt(1e3 * np.finfo(float).eps * np.maximum(sq, w ** 2))
        fuzzy = np.flatnonzero(rn < 10 * floor)
        if fuzzy.size and tol < 10 * floor.max():
            yf = torch.as_tensor(np.ascontiguousarray(y[:, fuzzy]), device=dev)
            wf = torch.as_tensor(np.ascontiguousarray(w[fuzzy]), device=dev)
            rn[fuzzy] = _ritz(VA, m, cap, yf, wf)[1].norm(dim=0).cpu().numpy()
        conv = rn <= tol
        mark('rr_eigh')
        trace.append({'cycle': cycle, 'space': int(m), 'max_residual': float(rn.max()),
                      'converged': int(conv.sum())})
        if conv.all():
            break
        if early is not None and 0 < early_roots <= k and conv[:early_roots].all():
            ye = torch.as_tensor(np.ascontiguousarray(y[:, :early_roots]), device=dev)
            Xe = torch.empty((n, early_roots), dtype=torch.float64, device=dev)
            _nn(Ranges(VA, (0, m)), ye, Xe, 0)
            early(Xe, np.array(w[:early_roots], dtype=float))
            early = None
            side_busy = True
        worst = float(rn[~conv].max())
        hist.append(worst)
        if stall_cycles and on_stall is not None and len(hist) > stall_cycles:
            if worst > 0.1 * hist[-1 - stall_cycles]:
                stalled = True
                new_pre = on_stall(cycle)
                if new_pre is not None:
                    precond, use_jd = new_pre, False
                hist = []
        idx = np.flatnonzero(~conv)
        ku = len(idx)
        # the unconverged Ritz vectors and their residuals, one pass over V, AV
        Yt = np.empty((m + 1, ku))
        Yt[:m], Yt[m] = y[:, idx], w[idx]
        Yd = torch.as_tensor(Yt, device=dev)
        wu = Yd[m]
        Xu, Ru = _ritz(VA, m, cap, Yd[:m], wu)
        mark('residuals')
        if observe is not None:
            observe(cycle, VA[:, :m], Ru, Xu, wu)

        def correction(pre, jd):
            if jd:
                return _jd(Xu, Ru, diag, wu)
            if precond_roots:
                return pre(Ru, Xu, wu, cycle, idx, rn[idx]).contiguous()
            return pre(Ru, Xu, wu, cycle).contiguous()
        Tc = correction(precond, use_jd)
        mark('precond')
        p = Tc.shape[1]
        if m + p > max_space:
            # collapse onto the selected and then the lowest Ritz vectors
            kk = min(m, max(k, keep or k))
            cols = list(order) + [j for j in range(m) if j not in set(order)]
            yk = np.ascontiguousarray(evecs[:, cols[:kk]])
            ykd = torch.as_tensor(yk, device=dev)
            Xc = torch.empty((n, kk), dtype=torch.float64, device=dev)
            AXc = torch.empty_like(Xc)
            _nn(Ranges(VA, (0, m)), ykd, Xc, 0)
            _nn(Ranges(VA, (cap, m)), ykd, AXc, 0)
            VA[:, :kk] = Xc
            VA[:, cap:cap + kk] = AXc
            # the host copies first: on a CPU tensor `hg` views HG itself
            Hns, Gc = yk.T @ Hns @ yk, yk.T @ G @ yk
            HG[:kk, :kk] = torch.as_tensor(Hns, device=dev)
            HG[:kk, cap:cap + kk] = torch.as_tensor(Gc, device=dev)
            m = kk
        mark('collapse')
        q, C, Hrow = _orthonormalize_block(VA, m, cap, Tc, Hns, drop, matmat)
        if q == 0 and on_stall is not None and not stalled:
            # the correction added nothing new: the same fallback as a stall
            stalled = True
            new_pre = on_stall(cycle)
            if new_pre is not None:
                precond, use_jd = new_pre, False
            Tc = correction(precond, use_jd)
            q, C, Hrow = _orthonormalize_block(VA, m, cap, Tc, Hns, drop, matmat)
        # the operator's pass on the new block ran inside (its image is in
        # AV's next columns)
        mark('orthonormal_sigma')
        if q == 0:
            exhausted = True
            break
        vectors += q
        P = _tnr(Ranges(VA, (0, m + q), (cap, m + q)), Ranges(VA, (cap + m, q)))
        HG[:m + q, m:m + q] = P[:m + q]
        HG[:m + q, cap + m:cap + m + q] = P[m + q:]
        HG[m:m + q, :m] = Hrow
        HG[m:m + q, cap:cap + m] = P[m + q:m + q + m].T
        m += q
        mark('append')
    if y_rows != m:
        # the space grew after the last Rayleigh-Ritz step (cycle budget ran out)
        H = HG[:m, :m].cpu().numpy()
        evals, evecs = np.linalg.eigh(0.5 * (H + H.T))
        k = min(nroots, m)
        w, y, _ = ritz_select(evals, evecs, k)
    yd = torch.as_tensor(np.ascontiguousarray(y), device=dev)
    wt = torch.as_tensor(np.ascontiguousarray(w), device=dev)
    X, R = _ritz(VA, m, cap, yd, wt)
    rn = R.norm(dim=0).cpu().numpy()
    conv = rn <= tol
    return DavidsonResult(e=np.asarray(w), X=X, conv=conv, rnorm=rn, cycles=cycles,
                          vectors=vectors, seconds=time.time() - t0, trace=trace,
                          stalled=stalled, exhausted=exhausted or m >= n)


def _orthonormalize_block(VA, m, cap, Tc, Hns, drop, matmat=None):
    """Append the corrections Tc (n, p) orthonormalised against the m columns
    of V in VA (n, 2 cap) as V's next columns.  Returns (q, C, Hrow): the
    number of columns kept, the device (p, q) map from Tc's projection to
    them, and the device block (q, m) = T^T AV of the rows they add to H.

    Two classical Gram-Schmidt passes, each one product [V | T]^T T over the
    pair index; the first also gives AV^T T.  The projected block's Gram
    matrix G2 = T1^T T1 - S2^T S2 (V orthonormal) goes to the host with the
    projections in one transfer: a column whose norm falls below `drop` of
    its norm before the projection is dropped, the rest are orthonormalised
    by G2's eigendecomposition (T2 C).  T^T AV follows from the algebra
    AV^T T2 = AV^T T - H^T (S1 + S2), H = V^T AV the whole Rayleigh matrix
    (the host copy `Hns`).

    With `matmat`, the operator is applied to the projected block as soon
    as the second pass has formed it, while the host works out C: the image
    of the kept columns is (A T2) C (A is linear), written into AV's next
    columns beside them.  The block goes to the operator with its columns
    scaled to unit norm (the dropped ones zeroed; the scale and the choice
    made on the device travel with the projections), so a lower-precision
    operator's rounding, relative to its input, is not magnified by C's
    1 / |column| factors."""
    dev = VA.device
    n, p = Tc.shape
    VA[:, m:m + p] = Tc
    # pass 1: [V | T | AV]^T T -> S1 = V^T T, G0 = T^T T, Pav = AV^T T
    stage = torch.empty((3 * m + 2 * p + 2, p), dtype=torch.float64, device=dev)
    P1 = stage[:2 * m + p]
    _tnr(Ranges(VA, (0, m + p), (cap, m)), Ranges(VA, (m, p)), out=P1)
    _nn(Ranges(VA, (0, m)), P1[:m], VA, m, alpha=-1.0, beta=1.0)
    # pass 2: [V | T1]^T T1 -> S2 = V^T T1, G1 = T1^T T1
    P2 = stage[2 * m + p:3 * m + 2 * p]
    _tnr(Ranges(VA, (0, m + p)), Ranges(VA, (m, p)), out=P2)
    _nn(Ranges(VA, (0, m)), P2[:m], VA, m, alpha=-1.0, beta=1.0)
    if matmat is not None and VA.is_cuda:
        block_scale(stage, m, p, drop, VA, m)
    elif matmat is not None:
        # |T2 columns| = diag(G1 - S2^T S2); drop as below, scale the rest
        g1 = torch.diagonal(P2[m:])
        after = (g1 - (P2[:m] * P2[:m]).sum(0)).clamp_min(0.0).sqrt()
        before = torch.diagonal(P1[m:m + p]).clamp_min(0.0).sqrt()
        kept = after > drop * before.clamp_min(1e-300)
        stage[-2] = after
        stage[-1] = kept.to(torch.float64)
        VA[:, m:m + p] *= torch.where(kept, 1.0 / after.clamp_min(1e-300),
                                      torch.zeros_like(after))[None, :]
    if VA.is_cuda:
        # the projections leave first; the operator's pass on T2 is queued
        # behind the copy and runs while the host solves the Gram problem
        host = torch.empty(stage.shape, dtype=torch.float64, pin_memory=True)
        host.copy_(stage, non_blocking=True)
        landed = torch.cuda.Event()
        landed.record()
    else:
        host = stage
    if matmat is not None:
        VA[:, cap + m:cap + m + p] = matmat(VA[:, m:m + p])
    if VA.is_cuda:
        landed.synchronize()
    h = host.numpy()
    S1, G0, Pav = h[:m], h[m:m + p], h[m + p:2 * m + p]
    S2, G1 = h[2 * m + p:3 * m + p], h[3 * m + p:3 * m + 2 * p]
    G2 = G1 - S2.T @ S2
    G2 = 0.5 * (G2 + G2.T)
    if matmat is not None:
        # the device's choice and scale (the block it sent to the operator)
        scale, keep = h[-2], np.flatnonzero(h[-1] > 0.5)
    else:
        before = np.sqrt(np.clip(np.diag(G0), 0.0, None))
        after = np.sqrt(np.clip(np.diag(G2), 0.0, None))
        keep = np.flatnonzero(after > drop * np.maximum(before, 1e-300))
        scale = np.ones(p)
    if keep.size == 0:
        return 0, None, None
    with _host_lapack():
        s, U = np.linalg.eigh(G2[np.ix_(keep, keep)])
    good = s > 1e-10 * s.max()
    q = int(good.sum())
    C = np.zeros((p, q))
    C[keep] = U[:, good] / np.sqrt(s[good])
    Hrow = ((Pav - Hns.T @ (S1 + S2)) @ C).T
    # T2 C = (T2 / scale) (scale C): the block in VA holds T2 / scale
    up = np.concatenate([(scale[:, None] * C).ravel(), Hrow.ravel()])
    if VA.is_cuda:
        # queued behind the operator's pass without holding the host
        pinned = torch.empty(up.size, dtype=torch.float64, pin_memory=True)
        pinned.numpy()[:] = up
        up = pinned.to(dev, non_blocking=True)
    else:
        up = torch.as_tensor(up)
    Cd, Hd = up[:p * q].reshape(p, q), up[p * q:].reshape(q, m)
    Tn = torch.empty((n, q), dtype=torch.float64, device=dev)
    _nn(Ranges(VA, (m, p)), Cd, Tn, 0)
    if matmat is not None:
        ATn = torch.empty((n, q), dtype=torch.float64, device=dev)
        _nn(Ranges(VA, (cap + m, p)), Cd, ATn, 0)
        VA[:, cap + m:cap + m + q] = ATn
    VA[:, m:m + q] = Tn
    return q, Cd, Hd


class overlap_pick:
    """A root-following `pick` (Solvers/davidson.py:80-101): at every step the
    `nroots` Ritz vectors with the largest overlap onto the fixed reference
    `ref` -- the norm of their projection onto the span of ref's columns --
    rather than the lowest, reported ascending in energy.  Overlaps within
    DAVIDSON_FOLLOW_OVERLAP_TIE of each other tie, and ties go to the lower
    energy, so roots the reference does not reach are taken in a fixed order.
    Called with envs['ref_coeff'] (V^T span, as the device iteration passes
    it) or envs['xs'] (the trial vectors, pyscf's form)."""

    def __init__(self, ref):
        ref = np.asarray(ref, float)
        self.span = np.linalg.qr(ref)[0] if ref.ndim == 2 else (ref / np.linalg.norm(ref))[:, None]

    def __call__(self, w, v, nroots, envs):
        w = np.asarray(w, float)
        if 'ref_coeff' in envs:
            coeff = np.asarray(envs['ref_coeff'], float)              # (m, r)
        else:
            coeff = np.array([self.span.T @ np.asarray(x, float).ravel() for x in envs['xs']])
        overlap = np.linalg.norm(coeff.T @ v, axis=0)
        tier = np.round(overlap / DAVIDSON_FOLLOW_OVERLAP_TIE)
        chosen = np.lexsort((w, -tier))[:nroots]
        order = chosen[np.argsort(w[chosen], kind='stable')]
        return w[order], v[:, order], order


def _off(V, X):
    """Columns of V projected off the orthonormal X twice and normalised;
    those left below DAVIDSON_NULL_DIRECTION are dropped."""
    for _ in range(2):
        V = V - X @ _tn(X, V)
    norms = V.norm(dim=0)
    keep = norms > DAVIDSON_NULL_DIRECTION
    return V[:, keep] / norms[keep]


def _deflated(block, X, lift):
    """The operator P A P + lift X X^T, P = 1 - X X^T, on (n, m) device
    blocks: A restricted to the complement of X, with X moved to `lift`."""
    def mm(V):
        C = _tn(X, V)
        AV = block(V - X @ C)
        return AV - X @ _tn(X, AV) + lift * (X @ C)
    return mm


def _complete_lowest(block, diag_t, e, X, warn, opts, label):
    """Certify that the converged (e, X) are the lowest len(e) eigenpairs
    (Solvers/davidson.py:362-404): the lowest root of the operator on the
    complement of X, with X lifted above the top root, replaces the top root
    when it lies below it by more than the tolerance; repeated until a
    complement solve finds nothing lower, at most k + 1 rounds, one root per
    round.  A complement solve that does not converge, or no certificate
    after k + 1 rounds, reports every root unconverged."""
    n, k = X.shape
    tol = opts['tol']
    rng = np.random.default_rng(DAVIDSON_COMPLETION_SEED)
    nseed = min(n, 2 * k + 4)
    seeds = torch.zeros((n, nseed), dtype=torch.float64, device=X.device)
    low = torch.argsort(diag_t)[:nseed]
    seeds[low, torch.arange(nseed, device=X.device)] = 1.0
    reason = None
    for _ in range(k + 1):
        if n - k < 1:
            return e, X, np.ones(k, bool)
        lift = e[-1] + abs(e[-1]) + (e[-1] - e[0]) + 1.0
        rand = torch.as_tensor(rng.standard_normal((n, k)), device=X.device)
        start = _off(torch.cat([seeds, rand], dim=1), X)
        res = davidson(_deflated(block, X, lift), diag_t, 1, start, **opts)
        if not res.conv.all():
            reason = 'the solve on their complement did not converge'
            break
        mu = float(res.e[0])
        if mu >= e[-1] - tol:
            return e, X, np.ones(k, bool)
        y = res.X[:, 0]
        y = y - X @ (X.T @ y)
        y = y / y.norm()
        w_all = np.append(e, mu)
        X_all = torch.cat([X, y[:, None]], dim=1)
        order = np.argsort(w_all, kind='stable')[:k]
        e, X = w_all[order], X_all[:, torch.as_tensor(order, device=X.device)]
    else:
        reason = f'{k + 1} rounds of merging did not end in a certificate'
    if warn:
        warnings.warn(f'{label}: the {k} converged roots are not certified to be the '
                      f'lowest ({reason}); all reported unconverged.', RuntimeWarning,
                      stacklevel=3)
    return e, X, np.zeros(k, bool)


def solve_symmetric(matvec, diag, nroots=1, x0=None, pick=None,
                    tol_residual=DAVIDSON_TOL_RESIDUAL, tol_eig=None,
                    max_cycle=DAVIDSON_MAX_CYCLE, max_subspace=None, precond=None,
                    lindep=None, warn_unconverged=True, label='Davidson',
                    complete_lowest=True, **unused):
    """(e, X, conv) of the `nroots` lowest eigenpairs, or with `pick` the ones
    it selects (Solvers/davidson.py:179-326).

    matvec: an object with `.matmat` (n, m) -> (n, m) on device FP64 tensors,
    or a callable on such blocks.  diag (n,).  x0 None seeds the 2 nroots + 4
    lowest-diagonal unit vectors.  pick: e.g. `overlap_pick(ref)`.
    complete_lowest (without pick): the converged set is certified to be the
    lowest by solves on its complement (`_complete_lowest`).  Returns NumPy
    e and conv and a device X, ascending.
    """
    block = getattr(matvec, 'matmat', matvec)
    dev = torch.device('cuda', torch.cuda.current_device())
    diag_t = torch.as_tensor(np.asarray(diag, float).ravel(), device=dev)
    n = diag_t.numel()
    nroots = max(1, min(int(nroots), n))
    if x0 is None:
        cols = np.argsort(diag_t.cpu().numpy())[:min(n, 2 * nroots + 4)]
        x0 = torch.zeros((n, len(cols)), dtype=torch.float64, device=dev)
        x0[torch.as_tensor(cols, device=dev), torch.arange(len(cols), device=dev)] = 1.0
    elif not isinstance(x0, torch.Tensor):
        x0 = torch.as_tensor(np.asarray(x0, float).reshape(n, -1), device=dev)
    pre = None
    if precond is not None:
        def pre(R, U, theta, cycle):
            return precond(R, U, theta)
    opts = dict(tol=tol_residual, max_cycle=max_cycle, max_space=max_subspace,
                precond=pre, lindep=lindep)
    res = davidson(block, diag_t, nroots, x0, pick=pick, **opts)
    e, X, conv = np.asarray(res.e), res.X, np.asarray(res.conv)
    order = np.argsort(e, kind='stable')
    e, X, conv = e[order], X[:, torch.as_tensor(order, device=dev)], conv[order]
    if complete_lowest and pick is None and conv.all() and e.size < n:
        e, X, conv = _complete_lowest(block, diag_t, e, X, warn_unconverged, opts,
                                      label + ' (complement)')
        if not conv.all():
            return e, X, conv
    if warn_unconverged and not conv.all():
        warnings.warn(f'{label} left {int((~conv).sum())} of {len(conv)} '
                      f'roots unconverged after {max_cycle} cycles',
                      RuntimeWarning, stacklevel=2)
    return e, X, conv


def dense_casida(apply_AB, diag_d, nroots, conv_tol, batch=64, tda=False):
    """(omega, X, Y, converged) of the `nroots` lowest roots from A (and B)
    built whole out of a device block action, one unit vector per pair
    (LR/davidson.py:1296-1330, `_dense_casida`): the completion of a
    Davidson whose trial space ran out.  Tamm-Dancoff: eigh(A), Y = 0.  Full:
    with A - B = L L^T (a Cholesky failure raises CasidaInstability),
    L^T (A + B) L z = omega^2 z, X + Y = L z / omega^1/2 and
    X - Y = L^-T z omega^1/2, so that X^T X - Y^T Y = 1.  FP64 on the
    device; `converged` is each root's residual against conv_tol.  X, Y
    are device (n_pair, nroots)."""
    op = _operator_of(apply_AB)
    no, nv = np.shape(diag_d)
    n = no * nv
    if n > BSE_DENSE_MAX_NOV:
        raise ValueError(f'a dense solve of {n} pairs exceeds BSE_DENSE_MAX_NOV '
                         f'({BSE_DENSE_MAX_NOV})')
    dev = op.device
    A = torch.empty((n, n), dtype=torch.float64, device=dev)
    B = None if tda else torch.empty_like(A)
    for c0 in range(0, n, batch):
        c1 = min(c0 + batch, n)
        unit = torch.zeros((c1 - c0, n), dtype=torch.float64, device=dev)
        unit[torch.arange(c1 - c0, device=dev), torch.arange(c0, c1, device=dev)] = 1.0
        z = unit.reshape(c1 - c0, no, nv)
        if tda:
            A[c0:c1] = op.apply(z).reshape(c1 - c0, n)
        else:
            Az, Bz = op.apply_ab(z)
            A[c0:c1], B[c0:c1] = Az.reshape(c1 - c0, n), Bz.reshape(c1 - c0, n)
    A = 0.5 * (A + A.T)
    k = min(nroots, n)
    if tda:
        w, V = torch.linalg.eigh(A)
        X = V[:, :k].contiguous()
        w = w[:k]
        r = (A @ X - X * w[None, :]).norm(dim=0)
        Y = torch.zeros_like(X)
    else:
        B = 0.5 * (B + B.T)
        L, info = torch.linalg.cholesky_ex(A - B)
        if int(info) != 0:
            raise CasidaInstability('(A - B) is not positive definite on the pair space')
        H = L.T @ (A + B) @ L
        w2, Z = torch.linalg.eigh(0.5 * (H + H.T))
        if not float(w2[0]) > 0:
            raise CasidaInstability('omega^2 is not positive on the pair space')
        w = torch.sqrt(w2[:k])
        Z = Z[:, :k]
        plus = (L @ Z) / torch.sqrt(w)[None, :]
        minus = torch.linalg.solve_triangular(L.T, Z, upper=True) * torch.sqrt(w)[None, :]
        X, Y = 0.5 * (plus + minus), 0.5 * (plus - minus)
        r = torch.sqrt(((A @ X + B @ Y - X * w[None, :]) ** 2).sum(0) +
                       ((B @ X + A @ Y + Y * w[None, :]) ** 2).sum(0))
    rn = r.cpu().numpy()
    return w.cpu().numpy(), X, Y, rn <= conv_tol, rn


# ---------------------------------------------------------------------------
# the TDA driver (LR/davidson.py)
# ---------------------------------------------------------------------------

def _operator_of(apply_AB):
    op = getattr(apply_AB, 'operator', None)
    if op is None:
        raise ValueError('run_davidson on the GPU needs a device block action '
                         '(mbptcode_amd.block_actions.isdf_block_action)')
    return op


def run_davidson(apply_AB, diag_d, nroots, conv_tol, max_cycle, x_sym=None,
                 guess_factor=DAVIDSON_GUESS_FACTOR, stats=None, comm=None,
                 timings=None, refuse_unconverged=False, max_memory=None,
                 preconditioner=DAVIDSON_PRECONDITIONER, read_roots=None,
                 read_tol=None, guess=None, tda=False, tracked=None,
                 extra_roots=None, max_space=None, refine_apply=None,
                 learned=None, stall_cycles=3, learned_until=1000.0,
                 learned_mode='augment', phases=None, keep=None, learned_gate='root',
                 learned_roots='requested'):
    """(omega, X, Y): the TDA roots from a GPU block action
    (LR/davidson.py:631 with tda=True -> `_tda_davidson_root` :462).

    preconditioner: 'bare' (d), 'screened' (d - (ii|W|aa)) or 'learned' (the
    MLP correction of `learned`, a `LearnedPreconditioner`, with the screened
    diagonal as its fallback whenever the residual stalls for `stall_cycles`
    cycles).  The learned correction is applied to the roots whose residual
    is at least learned_until * conv_tol (`learned_gate='root'`), or to all
    while the largest one is (`'max'`); below, Jacobi-Davidson alone.  With
    learned_roots='requested' only the nroots requested roots get it (the
    extra tracked roots, which only guard against a skipped root, keep the
    Jacobi-Davidson correction); 'all': every tracked root.  refine_apply: an FP64 block action of the same operator; the
    roots are then re-checked and polished in FP64 (the FP64 final residual
    check).  Returns NumPy omega (nroots,), X (n_pair, nroots), Y = 0.
    """
    if not tda:
        return _casida_root(apply_AB, diag_d, nroots, conv_tol, max_cycle,
                            guess_factor=guess_factor, stats=stats, comm=comm,
                            timings=timings, refuse_unconverged=refuse_unconverged,
                            preconditioner=preconditioner, guess=guess,
                            tracked=tracked, extra_roots=extra_roots,
                            max_space=max_space, refine_apply=refine_apply,
                            x_sym=x_sym)
    if x_sym is not None:
        raise ValueError('the Tamm-Dancoff Davidson takes no pair symmetry')
    if comm is not None and comm.Get_size() > 1:
        raise NotImplementedError('one GPU: no pair-row split')
    check_preconditioner(preconditioner)
    op = _operator_of(apply_AB)
    no, nv = np.shape(diag_d)
    npair = no * nv
    dev = op.device
    t_all = time.time()
    counters = {}
    t0 = time.time()
    if preconditioner == 'bare':
        pdiag = torch.as_tensor(np.asarray(diag_d, float).ravel(), device=dev)
    else:
        pdiag = op.screened_diagonal_device().reshape(-1)
    t_precond = time.time() - t0
    order = guess_indices(diag_d, nroots, guess_factor, counters=counters)
    track = tracked_roots(nroots, npair, extra_roots)
    x0sym = None
    if x_sym is not None:
        x_sym = np.asarray(x_sym).ravel()
        if x_sym.size != npair:
            raise ValueError(f'x_sym has {x_sym.size} labels for {npair} pairs')
        if guess is not None:
            raise ValueError('a warm-started Davidson takes no pair symmetry: the guess '
                             'vectors carry no irrep label')
        x0sym = x_sym[np.asarray(order)]
    x0 = torch.zeros((npair, len(order)), dtype=torch.float64, device=dev)
    x0[torch.as_tensor(order, device=dev), torch.arange(len(order), device=dev)] = 1.0
    if guess is not None:
        gx = torch.as_tensor(np.asarray(guess[0], float).reshape(npair, -1), device=dev)
        x0 = torch.cat([gx, x0], dim=1)

    def matmat_of(o):
        def mm(V):
            k = V.shape[1]
            Z = V.T.contiguous().reshape(k, no, nv)
            return o.apply(Z).reshape(k, npair).T
        return mm

    matmat = matmat_of(op)
    precond = None
    on_stall = None
    if preconditioner == 'learned':
        if learned is None:
            raise ValueError("preconditioner='learned' needs `learned` "
                             '(a LearnedPreconditioner)')
        state = {'fallback': False}

        def precond(R, U, theta, cycle, roots=None, rnorm=None):
            # near convergence the useful (out-of-space) part of a correction
            # is small next to the network's BF16 output noise: the
            # Jacobi-Davidson correction takes over below learned_until * tol,
            # root by root
            jd = _jd(U, R, pdiag, theta)
            if state['fallback']:
                return jd
            if learned_gate == 'root' and learned_until is not None:
                # the gate reads the Ritz step's residual norms on the host
                # when the iteration passes them (no device round trip)
                rsel = (np.asarray(rnorm) if rnorm is not None
                        else R.norm(dim=0).cpu().numpy())
                sel = rsel >= learned_until * conv_tol
                if learned_roots == 'requested' and roots is not None:
                    sel &= np.asarray(roots) < nroots
                big = np.flatnonzero(sel)
                if big.size == 0:
                    return jd
                big = torch.as_tensor(big, device=R.device)
                if big.numel() < R.shape[1]:
                    t = learned.correction(R[:, big], U[:, big], theta[big], pdiag,
                                           jd=jd[:, big])
                    if learned_mode == 'augment':
                        return torch.cat([jd, t], dim=1)
                    jd = jd.clone()
                    jd[:, big] = t
                    return jd
            elif learned_until is not None:
                worst = (float(np.max(rnorm)) if rnorm is not None
                         else float(R.norm(dim=0).max()))
                if worst < learned_until * conv_tol:
                    return jd
            t = learned.correction(R, U, theta, pdiag, jd=jd)
            # 'augment': both directions enter the space (the diagonal one is
            # never lost); 'replace': the learned one alone
            return torch.cat([jd, t], dim=1) if learned_mode == 'augment' else t

        def on_stall(cycle):
            state['fallback'] = True
            counters['learned_fallback_cycle'] = cycle
            return None
    early, checked = None, {}
    if refine_apply is not None:
        op64 = _operator_of(refine_apply)
        mm64 = matmat_of(op64)
        side = _side_stream(dev) if dev.type == 'cuda' else None

        def fp64_check(Xr):
            """(theta, |A x - theta x|) of the columns of Xr with FP64 sigma."""
            AX = mm64(Xr)
            theta = (Xr * AX).sum(0)
            return theta, (AX - Xr * theta[None, :]).norm(dim=0)

        if side is not None:
            def early(Xr, w):
                # the requested roots have converged while tracked extras
                # have not: check them in FP64 on a side stream while the
                # iteration finishes the extras on this one
                main = torch.cuda.current_stream(dev)
                side.wait_stream(main)
                if hasattr(op, 'profile_pause'):
                    op.profile_pause()     # this operator's applies now share the chip
                with torch.cuda.stream(side):
                    theta, rn_e = fp64_check(Xr)
                Xr.record_stream(side)
                for t in (theta, rn_e):
                    t.record_stream(main)
                checked.update(X=Xr, theta=theta, rn=rn_e)
    res = davidson(matmat, pdiag, track, x0, tol=conv_tol, max_cycle=max_cycle,
                   max_space=max_space, precond=precond,
                   stall_cycles=stall_cycles if preconditioner == 'learned' else None,
                   on_stall=on_stall, phases=phases, keep=keep, early=early,
                   early_roots=nroots if track > nroots else 0,
                   precond_roots=precond is not None)
    cycles, vectors = res.cycles, res.vectors
    e, X, conv, rn = res.e, res.X, res.conv, res.rnorm
    exhausted = res.exhausted
    refine = {}
    if refine_apply is not None:
        # the FP64 check covers the requested roots; the extra tracked ones
        # only bound the unconfirmed-root test and keep their residuals
        Xr = X[:, :nroots].contiguous()
        if checked:
            torch.cuda.current_stream(dev).wait_stream(side)
            if hasattr(op, 'profile_resume'):
                op.profile_resume()
            rn64 = checked['rn'].cpu().numpy()
            refine['early'] = bool((rn64 <= conv_tol).all())
            if refine['early']:
                # the roots as they were when they converged, checked in
                # FP64 while the extras finished
                Xr, theta = checked['X'], checked['theta']
                X = torch.cat([Xr, X[:, nroots:]], dim=1)
            else:
                checked.clear()
        if not checked:
            theta, rn_t = fp64_check(Xr)
            rn64 = rn_t.cpu().numpy()
        refine['rnorm_before'] = rn64.copy()
        if not (rn64 <= conv_tol).all():
            res2 = davidson(mm64, pdiag, track, X, tol=conv_tol, max_cycle=max_cycle,
                            max_space=max_space, phases=phases)
            e, X, conv, rn = res2.e, res2.X, res2.conv, res2.rnorm
            exhausted = res2.exhausted
            refine['cycles'] = res2.cycles
            refine['vectors'] = res2.vectors
            cycles += res2.cycles
            vectors += res2.vectors
            rn64 = rn[:nroots]
        else:
            e = np.array(e, dtype=float)
            e[:nroots] = theta.cpu().numpy()
            rn = np.array(rn, dtype=float)
            rn[:nroots] = rn64
            conv = rn <= conv_tol
        refine['rnorm_after'] = np.asarray(rn64).copy()
    dense = False
    if not np.all(conv[:nroots]) and exhausted and npair <= BSE_DENSE_MAX_NOV:
        # the trial space ran out short of the tolerance: the pair space is
        # small enough to finish densely (LR/davidson.py:1147-1193)
        act = refine_apply if refine_apply is not None else apply_AB
        e, X, _, conv, rn = dense_casida(act, diag_d, track, conv_tol, tda=True)
        dense = True
        cycles += 1
        vectors += npair
    unconfirmed = np.zeros(nroots, bool)
    extra = np.flatnonzero(~conv[nroots:]) + nroots
    if extra.size:
        lower = np.min(e[extra] - rn[extra])
        unconfirmed = e[:nroots] > lower
    seconds = time.time() - t_all
    if stats is not None:
        stats.update({'davidson_vind_calls': cycles, 'davidson_block_actions': vectors,
                      'davidson_dense_completion': dense})
    if timings is not None:
        timings.update({'davidson_s': seconds, 'davidson_iterations': cycles,
                        'davidson_vectors': vectors, 'davidson_precond_s': t_precond,
                        'davidson_converged': conv[:nroots].tolist(),
                        'davidson_residuals': np.asarray(rn[:nroots]).tolist(),
                        'davidson_unconfirmed': unconfirmed.tolist(),
                        'davidson_trace': res.trace,
                        'davidson_refine': {k: (v.tolist() if hasattr(v, 'tolist') else v)
                                            for k, v in refine.items()},
                        'davidson_flags': counters})
    if tracked is not None:
        tracked.update(omega=np.asarray(e).copy(), X=X.cpu().numpy(),
                       Y=np.zeros((npair, len(e))))
    bad = ~conv[:nroots]
    if bad.any():
        msg = (f'Tamm-Dancoff Davidson left {int(bad.sum())} of {nroots} roots '
               f'above |r| <= {conv_tol:g} after {cycles} cycles')
        if refuse_unconverged:
            raise DavidsonNotConverged(msg)
        warnings.warn(msg, RuntimeWarning, stacklevel=2)
    Xn = _to_host(X[:, :nroots])
    return np.asarray(e[:nroots]), Xn, np.zeros_like(Xn)


def _casida_root(apply_AB, diag_d, nroots, conv_tol, max_cycle,
                 guess_factor=DAVIDSON_GUESS_FACTOR, stats=None, comm=None, timings=None,
                 refuse_unconverged=False, preconditioner=DAVIDSON_PRECONDITIONER,
                 guess=None, tracked=None, extra_roots=None, max_space=None,
                 refine_apply=None, x_sym=None):
    """(omega, X, Y): the full Casida roots of a GPU block action, the
    reference's `_davidson_root` (LR/davidson.py:998) with its paired loop
    (`casida.casida_davidson`): the same guess pairs and tracked roots as the
    Tamm-Dancoff solve, the unit guesses with Y = 0 and a warm `guess`
    (X, Y) ahead of them.  refine_apply: an FP64 action of the same operator; the
    requested roots' residuals are recomputed on it, and the solve continues
    on it from those roots where one is above the tolerance.  x_sym: the
    symmetry label of every pair (`orbital_spaces.pair_symmetry`); the guess
    pairs carry their pair's label and the loop keeps the sectors apart
    (LR/davidson.py:1127-1130)."""
    from .casida import casida_davidson
    if comm is not None and comm.Get_size() > 1:
        raise NotImplementedError('one GPU: no pair-row split')
    check_preconditioner(preconditioner)
    op = _operator_of(apply_AB)
    no, nv = np.shape(diag_d)
    npair = no * nv
    dev = op.device
    t_all = time.time()
    counters = {}
    if preconditioner == 'bare':
        pdiag = torch.as_tensor(np.asarray(diag_d, float).ravel(), device=dev)
    elif preconditioner == 'learned':
        raise NotImplementedError('the learned correction is built for the '
                                  'Tamm-Dancoff Davidson only')
    else:
        pdiag = op.screened_diagonal_device().reshape(-1)
    order = guess_indices(diag_d, nroots, guess_factor, counters=counters)
    track = tracked_roots(nroots, npair, extra_roots)
    x0sym = None
    if x_sym is not None:
        x_sym = np.asarray(x_sym).ravel()
        if x_sym.size != npair:
            raise ValueError(f'x_sym has {x_sym.size} labels for {npair} pairs')
        if guess is not None:
            raise ValueError('a warm-started Davidson takes no pair symmetry: the guess '
                             'vectors carry no irrep label')
        x0sym = x_sym[np.asarray(order)]
    x0 = torch.zeros((npair, len(order)), dtype=torch.float64, device=dev)
    x0[torch.as_tensor(order, device=dev), torch.arange(len(order), device=dev)] = 1.0
    y0 = torch.zeros_like(x0)
    if guess is not None:
        gx = torch.as_tensor(np.asarray(guess[0], float).reshape(npair, -1), device=dev)
        gy = (torch.zeros_like(gx) if guess[1] is None else
              torch.as_tensor(np.asarray(guess[1], float).reshape(npair, -1), device=dev))
        x0, y0 = torch.cat([gx, x0], dim=1), torch.cat([gy, y0], dim=1)

    def ab_of(o):
        def ab(Z):
            k = Z.shape[0]
            Az, Bz = o.apply_ab(Z.reshape(k, no, nv))
            return Az.reshape(k, npair), Bz.reshape(k, npair)
        return ab

    probe_on = apply_AB if refine_apply is None else refine_apply

    def complete_densely(why):
        """The dense finish of a pair space the trial space ran out on
        (LR/davidson.py:1147-1193); an unstable reference refuses there."""
        try:
            return dense_casida(probe_on, diag_d, track, conv_tol, tda=False)
        except CasidaInstability as exc:
            raise UnstableReference(f'completed densely ({why}): {exc}', None) from exc

    dense = None
    try:
        res = casida_davidson(ab_of(op), pdiag, track, x0, tol=conv_tol,
                              max_cycle=max_cycle, max_space=max_space, y0=y0,
                              x0sym=x0sym)
    except np.linalg.LinAlgError as exc:
        if not (getattr(exc, 'space_exhausted', False) and npair <= BSE_DENSE_MAX_NOV):
            raise casida_breakdown(probe_on, diag_d, conv_tol) from exc
        dense = complete_densely('the projected (A - B) broke down')
        res = None
    if res is None:
        cycles, vectors = 1, npair
        e, X, Y, conv, rn = dense
        trace = []
    else:
        cycles, vectors, trace = res.cycles, res.vectors, res.trace
        e, X, Y, conv, rn = np.asarray(res.e, float), res.X, res.Y, res.conv, res.rnorm
    refine = {}
    if refine_apply is not None and dense is None:
        op64 = _operator_of(refine_apply)
        ab64 = ab_of(op64)
        Xr, Yr = X[:, :nroots].contiguous(), Y[:, :nroots].contiguous()
        A2, B2 = ab64(torch.cat([Xr, Yr], dim=1).T.contiguous())   # one apply of both
        AX, AY, BX, BY = A2[:nroots], A2[nroots:], B2[:nroots], B2[nroots:]
        w = torch.as_tensor(e[:nroots], device=dev)
        Rx = AX.T + BY.T - Xr * w
        Ry = BX.T + AY.T + Yr * w
        rn64 = torch.sqrt((Rx * Rx).sum(0) + (Ry * Ry).sum(0)).cpu().numpy()
        refine['rnorm_before'] = rn64.copy()
        if not (rn64 <= conv_tol).all():
            try:
                res2 = casida_davidson(ab64, pdiag, track, X.contiguous(), tol=conv_tol,
                                       max_cycle=max_cycle, max_space=max_space,
                                       y0=Y.contiguous(), x0sym=res.labels)
            except np.linalg.LinAlgError as exc:
                raise casida_breakdown(refine_apply, diag_d, conv_tol) from exc
            e, X, Y, conv, rn = np.asarray(res2.e, float), res2.X, res2.Y, res2.conv, res2.rnorm
            refine['cycles'] = res2.cycles
            refine['vectors'] = res2.vectors
            cycles += res2.cycles
            vectors += res2.vectors
            rn64 = rn[:nroots]
        else:
            rn = np.array(rn, dtype=float)
            rn[:nroots] = rn64
            conv = rn <= conv_tol
        refine['rnorm_after'] = np.asarray(rn64).copy()
    conv = np.asarray(conv, bool)
    if (dense is None and res is not None and not conv[:nroots].all() and res.exhausted
            and npair <= BSE_DENSE_MAX_NOV):
        dense = complete_densely('the trial space ran out')
        e, X, Y, conv, rn = dense
        cycles += 1
        vectors += npair
    if len(e) < nroots:
        raise RuntimeError(f'Casida Davidson found {len(e)} of {nroots} roots')
    unconfirmed = np.zeros(nroots, bool)
    extra = np.flatnonzero(~conv[nroots:]) + nroots
    if extra.size:
        lower = np.min(e[extra] - rn[extra])
        unconfirmed = e[:nroots] > lower
    if stats is not None:
        stats.update({'davidson_vind_calls': cycles, 'davidson_block_actions': vectors,
                      'davidson_dense_completion': dense is not None})
    if timings is not None:
        timings.update({'davidson_s': time.time() - t_all, 'davidson_iterations': cycles,
                        'davidson_vectors': vectors,
                        'davidson_converged': conv[:nroots].tolist(),
                        'davidson_residuals': np.asarray(rn[:nroots]).tolist(),
                        'davidson_unconfirmed': unconfirmed.tolist(),
                        'davidson_trace': trace,
                        'davidson_refine': {k: (v.tolist() if hasattr(v, 'tolist') else v)
                                            for k, v in refine.items()},
                        'davidson_flags': counters})
    Xn = X.cpu().numpy()
    Yn = Y.cpu().numpy()
    if tracked is not None:
        tracked.update(omega=e.copy(), X=Xn.copy(), Y=Yn.copy())
    bad = ~conv[:nroots]
    if bad.any():
        msg = (f'Casida Davidson left {int(bad.sum())} of {nroots} roots above '
               f'|r| <= {conv_tol:g} after {cycles} cycles')
        if refuse_unconverged:
            raise DavidsonNotConverged(msg)
        warnings.warn(msg, RuntimeWarning, stacklevel=3)
    return e[:nroots], Xn[:, :nroots], Yn[:, :nroots]


def solve_casida_davidson(lr_solver, nocc, numerics=None, polarizability='RPA',
                          W_aux=None, orbsym=None, isdf_factors=None, stats=None,
                          spin='singlet', comm=None, timings=None,
                          refuse_unconverged=False, max_memory=None,
                          read_roots=None, read_tol=None, guess=None, tda=False,
                          one_body=None, tracked=None, probe=False, nroots=5,
                          conv_tol=1e-8, max_cycle=150, guess_factor=2,
                          preconditioner=DAVIDSON_PRECONDITIONER,
                          precision='fp64', refine=True, learned=None):
    """The matrix-free Davidson entry (LR/davidson.py:252) on the GPU: the
    Tamm-Dancoff solve (tda=True) or the full Casida one (tda=False, the
    paired loop of `casida.casida_davidson`).

    `numerics` may carry a `bse` section with nroots, conv_tol, max_cycle,
    guess_factor and preconditioner (numerics.py:241-286); otherwise the
    keyword values (the reference defaults) are used.  polarizability 'RPA'
    (the reference's default), 'TDHF' or 'BSE'.  precision: the block
    action's, 'fp64' by default as the reference computes; 'fp32' or 'bf16x3'
    are an explicit opt-in, and with refine (default) such a solve is
    re-checked and polished on an FP64 action.  probe (full solve only): True or 'sign' runs
    the (A - B) probe after the solve on the FP64 action (`probe_after_solve`),
    puts min eig(A - B) in stats['min_eig_amb'] and refuses the roots of an
    unstable reference (UnstableReference).  Returns (omega, X, Y).
    """
    from .block_actions import block_action
    bse = getattr(numerics, 'bse', None)
    if bse is not None:
        nroots, conv_tol, max_cycle = bse.nroots, bse.conv_tol, bse.max_cycle
        guess_factor, preconditioner = bse.guess_factor, bse.preconditioner
    check_preconditioner(preconditioner)
    if orbsym is not None and tda:
        raise ValueError('the Tamm-Dancoff Davidson takes no pair symmetry: its '
                         'trial space is one Hermitian block')
    if one_body is not None:
        raise NotImplementedError('one_body coupling is not built')
    from .orbital_spaces import bse_pair_diagonal
    check_residual_floor(conv_tol, bse_pair_diagonal(lr_solver.eps, nocc))
    apply_AB, diag_d = block_action(lr_solver, nocc, polarizability, W_aux,
                                    isdf_factors, spin=spin, comm=comm, tda=tda,
                                    precision=precision)
    refine_apply = None
    if refine and precision != 'fp64':
        refine_apply = block_action(lr_solver, nocc, polarizability, W_aux,
                                    isdf_factors, spin=spin, comm=comm, tda=tda,
                                    precision='fp64')[0]
    from .orbital_spaces import get_occ_virt_indices, pair_symmetry
    occ, virt = get_occ_virt_indices(lr_solver.eps, nocc)
    omega, X, Y = run_davidson(apply_AB, diag_d, nroots, conv_tol, max_cycle,
                               pair_symmetry(orbsym, occ, virt),
                               guess_factor=guess_factor, stats=stats, comm=comm,
                               timings=timings, refuse_unconverged=refuse_unconverged,
                               preconditioner=preconditioner, guess=guess, tda=tda,
                               tracked=tracked, refine_apply=refine_apply,
                               learned=learned)
    if probe and not tda:
        amb = probe_after_solve(apply_AB if refine_apply is None else refine_apply,
                                diag_d, omega, X, Y, probe, stats=stats, comm=comm)
        if stats is not None:
            stats['min_eig_amb'] = amb
    return omega, X, Y
