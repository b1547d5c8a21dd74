"""The full (non-Tamm-Dancoff) Casida Davidson on the GPU.

The roots of [[A, B], [B, A]] [X; Y] = omega [[1, 0], [0, -1]] [X; Y] from a
paired trial space, as the reference's `casida_rows`
(LR/trial_space.py:454-786, pyscf's `real_eig` on pair rows) holds it:

* Trial pairs (v, w) enter with their partners (w, v); the holders keep V, W
  and the action's U1 = A V + B W, U2 = A W + B V (two block actions per
  batch, on V and on W: LR/davidson.py:1071-1088).
* The projected blocks are a = sym(V^T U1) + sym(W^T U2),
  b = sym(V^T U2) + sym(W^T U1), sigma = V^T V - W^T W and
  pi = V^T W - W^T V (`projection_addend` :216-234).  The new rows of all
  four come from one device reduction of the new columns against the whole
  holders.
* The small problem [[a, b], [b, a]] [x; y] = omega [[sigma, pi], [-pi,
  -sigma]] [x; y] (pyscf's `TDDFT_subspace_eigen_solver`, not vendored) is
  solved here by its Hermitian form, in FP64 on the host.  With
  u = x + y, v = x - y, T = sigma + pi and a - b = L L^T:
  (a + b) u = omega^2 T^T (a - b)^-1 T u; C = L^-1 T turns that into the
  symmetric C^-T (a + b) C^-1 w = omega^2 w, w = C u, and
  v = omega L^-T w.  Scaling |w|^2 = 1/omega gives X^T X - Y^T Y = 1.
* Ritz pairs are X = V x + W y, Y = W x + V y.  The residuals are
  R_x = U1 x + U2 y - omega X and R_y = U2 x + U1 y + omega Y
  (`ritz_rows` :237-253).
* Corrections are R_x / (d - omega) and R_y / (d + omega): the diagonal
  approximation of (H - omega S) t = r with H = [[A, B], [B, A]] and
  S = diag(1, -1), whose Y row is A + omega.  The reference's `precond_rows`
  (LR/davidson.py:1093-1120) divides R_y by -d - omega, pyscf's hdiag for the
  form [[A, B], [-B, -A]] whose residual's Y row has the opposite sign; with
  R_y as held here that flips the correction's Y part, and the solve takes
  about twice the cycles (11 against 6 at 4500 pairs, tol 1e-5).  The
  denominators are floored as in the Tamm-Dancoff loop.  They are projected twice off the held space and its
  partner space, then orthonormalised together with their partners
  (`gram_schmidt_rows` :340-386, `symmetric_coefficients` :296-337).
* A restart keeps the tracked Ritz pairs, their products combined from the
  held ones (:574-588).

Simplifications against the reference:

* Tracked extra pairs must converge; the reference also settles them by their
  residual interval clearing the highest requested root.
* Pair-symmetry sectors (`x0sym`) label the trial pairs; the warm start
  takes none (as the reference refuses one with them).
* A pair space the trial space exhausts is finished densely by the caller
  (`davidson._casida_root` through `davidson.dense_casida`).
* A projected a - b that is not positive definite raises LinAlgError here;
  `davidson.run_davidson` turns it into CasidaBreakdown or
  UnstableReference by the (A - B) probe (`probe.py`).
* Corrections keep their raw length.
"""
import time
import warnings

import numpy as np
import scipy.linalg as sla
import torch

from .constants import DAVIDSON_DENOMINATOR_FLOOR, DAVIDSON_LINDEP_FLOOR
from .davidson import _host_lapack


def _tn(A, B):
    if A.is_cuda:
        from .subspace import tn
        return tn(A, B)
    return A.T @ B


def subspace_solve(a, b, sigma, pi, k):
    """(omega, x, y): the k lowest positive roots of the projected Casida
    problem [[a, b], [b, a]] [x; y] = omega [[sigma, pi], [-pi, -sigma]] [x; y]
    (host FP64, m x m each), by the Hermitian form described above.  Raises
    LinAlgError where a - b is not positive definite on the space."""
    P, Q, T = a + b, a - b, sigma + pi
    L = np.linalg.cholesky(0.5 * (Q + Q.T))
    # C = L^-1 T, so C^-1 = T^-1 L: one LU of T serves H and the vectors
    U = sla.lu_solve(sla.lu_factor(T, check_finite=False), L, check_finite=False)
    H = U.T @ P @ U                                         # C^-T P C^-1
    H = 0.5 * (H + H.T)
    m = H.shape[0]
    top = min(m, 2 * k)
    w2, Wv = sla.eigh(H, subset_by_index=[0, top - 1], driver='evr')
    if top < m and np.count_nonzero(w2 > 0) < k:
        w2, Wv = np.linalg.eigh(H)
    keep = np.flatnonzero(w2 > 0)[:k]
    w = np.sqrt(w2[keep])
    Wv = Wv[:, keep] / np.sqrt(w)                           # |w|^2 = 1 / omega
    u = U @ Wv                                              # C^-1 Wv
    v = sla.solve_triangular(L.T, Wv, lower=False) * w
    return w, 0.5 * (u + v), 0.5 * (u - v)


def symmetric_coefficients(s11, s21, lindep):
    """(c_a, c_b): the combinations x' = X c_a + Y c_b, y' = Y c_a + X c_b
    that make k directions [X; Y] and their partners [Y; X] one orthonormal
    set, from s11 = X^T X + Y^T Y and s21 = X^T Y + Y^T X; None when no
    direction survives `lindep`.

    Whiten s11 (eigenvalues above lindep), then diagonalise the whitened
    partner overlap s21.  A direction d of unit length with partner overlap
    w becomes (a d + b d') with a, b from [1 w; w 1]^(-1/2) = [a b; b a]:
    a = ((1+w)^-1/2 + (1-w)^-1/2) / 2, b = ((1+w)^-1/2 - (1-w)^-1/2) / 2.
    Directions almost their own partner (|w| > 1 - sqrt(lindep)) are dropped,
    and the whole is done twice so round-off from a nearly dependent set does
    not survive."""
    cx = None
    for _ in range(2):
        S11 = s11 if cx is None else cx.T @ s11 @ cx
        e, c = np.linalg.eigh(0.5 * (S11 + S11.T))
        mask = e > lindep
        if not mask.any():
            return None
        c = c[:, mask] / np.sqrt(e[mask])
        cx = c if cx is None else cx @ c
        csc = cx.T @ s21 @ cx
        w, u = np.linalg.eigh(0.5 * (csc + csc.T))
        mask = 1.0 - np.abs(w) > np.sqrt(lindep)
        if not mask.any():
            return None
        cx = cx @ u[:, mask]
        w = w[mask]
    a1, a2 = (1.0 + w) ** -0.5, (1.0 - w) ** -0.5
    return cx * (0.5 * (a1 + a2)), cx * (0.5 * (a1 - a2))


class CasidaResult:
    __slots__ = ('e', 'X', 'Y', 'conv', 'rnorm', 'cycles', 'vectors', 'seconds',
                 'trace', 'exhausted', 'labels')

    def __init__(self, **kw):
        for k in self.__slots__:
            setattr(self, k, kw.get(k))


def sector_solve(a, b, sigma, pi, labels, k):
    """(omega, x, y, labels): `subspace_solve` sector by sector when the
    trial pairs carry symmetry labels (`labels`, one per trial pair): the
    projected problem of each sector is solved apart, its coefficients
    embedded in the whole space (zero on the other sectors), and the roots of
    all sectors merged in ascending order; the k lowest are kept with the
    labels of their sectors (trial_space.py:624-648).  labels None: one
    solve."""
    if labels is None:
        w, x, y = subspace_solve(a, b, sigma, pi, k)
        return w, x, y, None
    m = a.shape[0]
    ws, xs, ys, ls = [], [], [], []
    for lab in np.unique(labels):
        idx = np.flatnonzero(labels == lab)
        sub = np.ix_(idx, idx)
        w, x, y = subspace_solve(a[sub], b[sub], sigma[sub], pi[sub], idx.size)
        xe = np.zeros((m, w.size))
        ye = np.zeros((m, w.size))
        xe[idx], ye[idx] = x, y
        ws.append(w)
        xs.append(xe)
        ys.append(ye)
        ls.append(np.full(w.size, lab))
    w, x, y, lab = (np.concatenate(ws), np.concatenate(xs, 1), np.concatenate(ys, 1),
                    np.concatenate(ls))
    order = np.argsort(w, kind='stable')[:k]
    return w[order], x[:, order], y[:, order], lab[order]


def casida_davidson(apply_ab, diag, nroots, x0, tol, max_cycle=100, max_space=None,
                    lindep=None, y0=None, phases=None, x0sym=None):
    """Paired Davidson for the `nroots` lowest Casida roots, on device (or
    CPU) tensors.

    apply_ab(Z (p, n)) -> (A Z, B Z), each (p, n), FP64.  diag (n,) the
    preconditioner diagonal.  x0 (n, p0) guess columns for X, y0 those for Y
    (default 0; a warm start passes converged (X, Y) pairs).  Every
    returned root satisfies |[R_x; R_y]| <= tol where `conv` says so.
    `phases`, a dict, receives the seconds of each part of the cycle (the
    device is synchronised at every boundary, so only for profiling).
    x0sym: a symmetry label for each guess column (the irrep of its pair);
    the trial pairs then keep labels, the projected problem is solved sector
    by sector (`sector_solve`), a root's correction takes the root's label
    and the corrections are orthonormalised sector by sector (against the
    whole held space), as the reference's symmetry-blocked loop does
    (trial_space.py:424-452, :514-648); `labels` of the result are the
    roots' sectors.

    The holders V, W, U1, U2 are four column blocks of one row-major buffer
    (n, 4 cap), so [V | W] or [U1 | U2] is one operand of the two-range
    reductions and products (dvd.hip).  A cycle's corrections are written
    straight into V's and W's next columns and orthonormalised there: one
    reduction and two products project them off the held pairs, one more
    reduction gives the second projection and their Gram matrices together,
    and one host transfer per cycle carries that (plus one for the new rows
    of the projected blocks)."""
    last = [time.perf_counter()]

    def mark(name):
        if phases is not None:
            if diag.is_cuda:
                torch.cuda.synchronize()
            now = time.perf_counter()
            phases[name] = phases.get(name, 0.0) + now - last[0]
            last[0] = now
    from .davidson import _nn, _tnr
    from .subspace import Ranges
    dev = diag.device
    n = diag.numel()
    nroots = max(1, min(int(nroots), n))
    if max_space is None:
        max_space = max(10 * nroots + 30, 60)
    max_space = int(min(max(max_space, 2 * nroots), n))
    if lindep is None:
        lindep = float(max(DAVIDSON_LINDEP_FLOOR, min(1e-14, 1e-2 * tol * tol)))
    floor = DAVIDSON_DENOMINATOR_FLOOR
    t0 = time.time()
    cap = max_space + max(x0.shape[1], nroots)
    oV, oW, oU1, oU2 = 0, cap, 2 * cap, 3 * cap
    H4 = torch.zeros((n, 4 * cap), dtype=torch.float64, device=dev)   # [V | W | U1 | U2]
    a = np.zeros((cap, cap))
    b, sg, pi = np.zeros_like(a), np.zeros_like(a), np.zeros_like(a)
    m = 0
    vectors = 0
    exhausted = False
    lab = None if x0sym is None else np.zeros(cap, dtype=np.int64)

    def pair(o1, o2, c0, c):
        """The columns c0 .. c0 + c of the blocks at o1 and o2, one operand."""
        return Ranges(H4, (o1 + c0, c), (o2 + c0, c))

    def act(c0, q, cold):
        """U1 = A V + B W, U2 = A W + B V of the new columns into their slots
        (two block actions per batch, LR/davidson.py:1071-1088)."""
        nonlocal vectors
        V, W = H4[:, oV + c0:oV + c0 + q], H4[:, oW + c0:oW + c0 + q]
        if cold:
            # a cold start's pairs have no Y half: U1 = A V, U2 = B V from
            # one apply of V alone
            AZ, BZ = apply_ab(V.T.contiguous())
            vectors += q
            H4[:, oU1 + c0:oU1 + c0 + q] = AZ.T
            H4[:, oU2 + c0:oU2 + c0 + q] = BZ.T
            return
        AZ, BZ = apply_ab(torch.cat([V, W], dim=1).T.contiguous())
        vectors += 2 * q
        H4[:, oU1 + c0:oU1 + c0 + q] = (AZ[:q] + BZ[q:]).T
        H4[:, oU2 + c0:oU2 + c0 + q] = (AZ[q:] + BZ[:q]).T

    def append(m0, q):
        """Fill the rows of a, b, sigma and pi of the q columns held at m0."""
        nonlocal m
        m1 = m0 + q
        stage = torch.empty((3, 2 * m1, 2 * q), dtype=torch.float64, device=dev)
        _tnr(pair(oV, oW, 0, m1), pair(oV, oW, m0, q), out=stage[0])     # [V;W]^T [Vn|Wn]
        _tnr(pair(oV, oW, 0, m1), pair(oU1, oU2, m0, q), out=stage[1])   # [V;W]^T [U1n|U2n]
        _tnr(pair(oU1, oU2, 0, m1), pair(oV, oW, m0, q), out=stage[2])   # [U1;U2]^T [Vn|Wn]
        G = stage.cpu().numpy()

        def blk(s_, r, c):
            """(new, old) block: the transpose of stage s_ rows r, columns c."""
            return G[s_, r * m1:(r + 1) * m1, c * q:(c + 1) * q].T
        # rows of new^T old: V_n^T V = blk(0,0,0), W_n^T W = blk(0,1,1),
        # V_n^T W = blk(0,1,0), W_n^T V = blk(0,0,1); U1_n^T V = blk(1,0,0),
        # U2_n^T V = blk(1,0,1), U1_n^T W = blk(1,1,0), U2_n^T W = blk(1,1,1);
        # V_n^T U1 = blk(2,0,0), W_n^T U1 = blk(2,0,1), V_n^T U2 = blk(2,1,0),
        # W_n^T U2 = blk(2,1,1)
        a_rows = 0.5 * (blk(2, 0, 0) + blk(1, 0, 0)) + 0.5 * (blk(2, 1, 1) + blk(1, 1, 1))
        b_rows = 0.5 * (blk(2, 1, 0) + blk(1, 0, 1)) + 0.5 * (blk(2, 0, 1) + blk(1, 1, 0))
        s_rows = blk(0, 0, 0) - blk(0, 1, 1)
        p_rows = blk(0, 1, 0) - blk(0, 0, 1)
        for M_, rows_, sign in ((a, a_rows, 1), (b, b_rows, 1), (sg, s_rows, 1),
                                (pi, p_rows, -1)):
            M_[m0:m1, :m1] = rows_
            M_[:m0, m0:m1] = sign * rows_[:, :m0].T
        m = m1

    def orthonormal_new(p, tlab=None):
        """The p corrections held in V's and W's columns m .. m + p,
        projected twice off the held pairs and their partners and
        orthonormalised with their own partners, in place (sector by sector
        when `tlab` labels them); returns how many directions remain (in
        columns m .. m + q, their labels in lab[m:m + q])."""
        hx = lambda M_: torch.as_tensor(np.ascontiguousarray(M_), device=dev)
        if m:
            O = _tnr(pair(oV, oW, 0, m), pair(oV, oW, m, p))           # (2m, 2p)
            ox = O[:m, :p] + O[m:, p:]                                 # V^T tx + W^T ty
            oy = O[:m, p:] + O[m:, :p]                                 # V^T ty + W^T tx
            _nn(pair(oV, oW, 0, m), torch.cat([ox, oy]).contiguous(), H4, oV + m, -1.0, 1.0)
            _nn(pair(oV, oW, 0, m), torch.cat([oy, ox]).contiguous(), H4, oW + m, -1.0, 1.0)
        # second projection and the Gram matrices in one reduction:
        # [V tx | W ty]^T [tx | ty]
        G = _tnr(Ranges(H4, (oV, m + p), (oW, m + p)), pair(oV, oW, m, p)).cpu().numpy()
        r = m + p
        Vt, Tx, Wt, Ty = G[:m], G[m:r], G[r:r + m], G[r + m:]
        ox2 = Vt[:, :p] + Wt[:, p:]
        oy2 = Vt[:, p:] + Wt[:, :p]
        s11 = Tx[:, :p] + Ty[:, p:] - ox2.T @ ox2 - oy2.T @ oy2
        s21 = Tx[:, p:] + Ty[:, :p] - ox2.T @ oy2 - oy2.T @ ox2
        groups = ([np.arange(p)] if tlab is None else
                  [np.flatnonzero(tlab == g) for g in np.unique(tlab)])
        blocks = []
        with _host_lapack():
            for gi in groups:
                coef = symmetric_coefficients(s11[np.ix_(gi, gi)], s21[np.ix_(gi, gi)],
                                              lindep)
                if coef is not None:
                    blocks.append((gi, coef))
        if not blocks:
            return 0
        q = sum(c[0].shape[1] for _, c in blocks)
        ca, cb = np.zeros((p, q)), np.zeros((p, q))
        j = 0
        for gi, (cga, cgb) in blocks:
            ca[gi, j:j + cga.shape[1]] = cga
            cb[gi, j:j + cga.shape[1]] = cgb
            if tlab is not None:
                lab[m + j:m + j + cga.shape[1]] = tlab[gi[0]]
            j += cga.shape[1]
        # t_x' = (t_x - V ox2 - W oy2) ca + (t_y - W ox2 - V oy2) cb, and its
        # partner, each one product over [V t_x | W t_y]
        Yx = np.concatenate([-(ox2 @ ca + oy2 @ cb), ca, -(oy2 @ ca + ox2 @ cb), cb])
        Yy = np.concatenate([-(oy2 @ ca + ox2 @ cb), cb, -(ox2 @ ca + oy2 @ cb), ca])
        out = torch.empty((n, 2 * q), dtype=torch.float64, device=dev)
        A_ = Ranges(H4, (oV, m + p), (oW, m + p))
        _nn(A_, hx(Yx), out, 0)
        _nn(A_, hx(Yy), out, q)
        H4[:, oV + m:oV + m + q] = out[:, :q]
        H4[:, oW + m:oW + m + q] = out[:, q:]
        return q

    x0 = x0.to(torch.float64)
    p0 = x0.shape[1]
    H4[:, oV:oV + p0] = x0
    # a cold start's pairs have no Y half (y0 absent or zero): their W
    # columns stay zero through the orthonormalisation, so one apply serves
    cold = y0 is None or not bool(torch.any(y0))
    if not cold:
        H4[:, oW:oW + p0] = y0.to(device=x0.device, dtype=torch.float64)
    q = orthonormal_new(p0, None if x0sym is None else np.asarray(x0sym, dtype=np.int64))
    act(0, q, cold)
    append(0, q)
    trace = []
    conv = np.zeros(nroots, bool)
    w = x = y = rn = None
    X = Y = None
    wlab = None
    cycles = 0
    mark('setup')
    for cycle in range(max_cycle):
        cycles = cycle + 1
        k = min(nroots, m)
        try:
            with _host_lapack():
                w, x, y, wlab = sector_solve(a[:m, :m], b[:m, :m], sg[:m, :m],
                                             pi[:m, :m], None if lab is None else lab[:m], k)
        except np.linalg.LinAlgError as exc:
            # whether the trial space had spanned the pair space when the
            # projected (A - B) block broke down (the caller may then solve
            # densely)
            exc.space_exhausted = m >= n or exhausted
            raise
        mark('subspace')
        kk = len(w)
        xy = torch.as_tensor(np.ascontiguousarray(np.concatenate([x, y])), device=dev)
        yx = torch.as_tensor(np.ascontiguousarray(np.concatenate([y, x])), device=dev)
        wt = torch.as_tensor(w, device=dev)
        # X = V x + W y, Y = W x + V y; R_x = U1 x + U2 y - w X,
        # R_y = U2 x + U1 y + w Y (`ritz_rows` :237-253)
        XY = torch.empty((n, 4 * kk), dtype=torch.float64, device=dev)
        _nn(pair(oV, oW, 0, m), xy, XY, 0)
        _nn(pair(oV, oW, 0, m), yx, XY, kk)
        _nn(pair(oU1, oU2, 0, m), xy, XY, 2 * kk)
        _nn(pair(oU1, oU2, 0, m), yx, XY, 3 * kk)
        X, Y = XY[:, :kk], XY[:, kk:2 * kk]
        Rx = XY[:, 2 * kk:3 * kk] - X * wt
        Ry = XY[:, 3 * kk:] + Y * wt
        rn = torch.sqrt((Rx * Rx).sum(0) + (Ry * Ry).sum(0)).cpu().numpy()
        conv = rn <= tol
        trace.append({'cycle': cycle, 'space': int(m), 'max_residual': float(rn.max()),
                      'converged': int(conv.sum())})
        mark('residuals')
        if conv.all() or kk < k:
            break
        idx = np.flatnonzero(~conv)
        p = len(idx)
        it = torch.as_tensor(idx, device=dev)
        wi = wt[it]
        den_x = diag[:, None] - wi[None, :]
        den_y = diag[:, None] + wi[None, :]
        den_x = torch.where(den_x.abs() < floor, torch.full_like(den_x, floor), den_x)
        den_y = torch.where(den_y.abs() < floor, torch.full_like(den_y, floor), den_y)
        tx, ty = Rx[:, it] / den_x, Ry[:, it] / den_y
        nrm = torch.sqrt((tx * tx).sum(0) + (ty * ty).sum(0)).clamp_min(1e-300)
        H4[:, oV + m:oV + m + p] = tx / nrm
        H4[:, oW + m:oW + m + p] = ty / nrm
        mark('correction')
        q = orthonormal_new(p, None if wlab is None else wlab[idx])
        mark('orthonormal')
        if q == 0:
            exhausted = True
            break
        if m + q > max_space:
            # restart on the lowest Ritz pairs, their products combined from
            # the held ones (the new directions stay orthogonal to them)
            keep = H4[:, [oV + m + j for j in range(q)] + [oW + m + j for j in range(q)]]
            keep_lab = None if lab is None else lab[m:m + q].copy()
            U12 = torch.empty((n, 2 * kk), dtype=torch.float64, device=dev)
            _nn(pair(oU1, oU2, 0, m), xy, U12, 0)
            _nn(pair(oU1, oU2, 0, m), yx, U12, kk)
            H4[:, oV:oV + kk] = X
            H4[:, oW:oW + kk] = Y
            H4[:, oU1:oU1 + kk] = U12[:, :kk]
            H4[:, oU2:oU2 + kk] = U12[:, kk:]
            if lab is not None:
                lab[:kk] = wlab
            append(0, kk)
            H4[:, oV + m:oV + m + q] = keep[:, :q]
            H4[:, oW + m:oW + m + q] = keep[:, q:]
            if lab is not None:
                lab[m:m + q] = keep_lab
        mark('collapse')
        act(m, q, False)
        mark('sigma')
        append(m, q)
        mark('append')
    if len(w) < nroots:
        warnings.warn(f'Casida Davidson: only {len(w)} roots in a space of {m}',
                      RuntimeWarning, stacklevel=2)
    return CasidaResult(e=np.asarray(w), X=X, Y=Y, conv=conv, rnorm=rn, cycles=cycles,
                        vectors=vectors, seconds=time.time() - t0, trace=trace,
                        exhausted=exhausted or m >= n, labels=wlab)
