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
* There are no pair-symmetry sectors.
* There is no dense completion for a pair space the trial space exhausts.
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
                 'trace', 'exhausted')

    def __init__(self, **kw):
        for k in self.__slots__:
            setattr(self, k, kw.get(k))


def casida_davidson(apply_ab, diag, nroots, x0, tol, max_cycle=100, max_space=None,
                    lindep=None, y0=None, phases=None):
    """Paired Davidson for the `nroots` lowest Casida roots, on device (or
    CPU) tensors.

    apply_ab(Z (p, n)) -> (A Z, B Z), each (p, n), FP64.  diag (n,) the
    preconditioner diagonal.  x0 (n, p0) guess columns for X, y0 those for Y
    (default 0; a warm start passes converged (X, Y) pairs).  Every
    returned root satisfies |[R_x; R_y]| <= tol where `conv` says so.
    `phases`, a dict, receives the seconds of each part of the cycle (the
    device is synchronised at every boundary, so only for profiling)."""
    last = [time.perf_counter()]

    def mark(name):
        if phases is not None:
            if diag.is_cuda:
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
    if lindep is None:
        lindep = float(max(DAVIDSON_LINDEP_FLOOR, min(1e-14, 1e-2 * tol * tol)))
    floor = DAVIDSON_DENOMINATOR_FLOOR
    t0 = time.time()
    cap = max_space + x0.shape[1] + nroots
    Vh = torch.zeros((n, cap), dtype=torch.float64, device=dev)
    Wh, U1h, U2h = torch.zeros_like(Vh), torch.zeros_like(Vh), torch.zeros_like(Vh)
    a = np.zeros((cap, cap))
    b, sg, pi = np.zeros_like(a), np.zeros_like(a), np.zeros_like(a)
    m = 0
    vectors = 0
    exhausted = False

    def act(V, W):
        nonlocal vectors
        p = V.shape[1]
        if not bool(torch.any(W)):
            # a cold start's pairs have no Y half: U1 = A V, U2 = B V from
            # one apply of V alone
            AZ, BZ = apply_ab(V.T.contiguous())
            vectors += p
            return AZ.T, BZ.T
        AB = apply_ab(torch.cat([V, W], dim=1).T.contiguous())
        AZ, BZ = AB[0].T, AB[1].T
        vectors += 2 * p
        return AZ[:, :p] + BZ[:, p:], AZ[:, p:] + BZ[:, :p]   # U1, U2

    def append(V, W, U1, U2):
        """Hold the new columns and fill their rows of a, b, sigma, pi."""
        nonlocal m
        p = V.shape[1]
        m0, m1 = m, m + p
        Vh[:, m0:m1], Wh[:, m0:m1], U1h[:, m0:m1], U2h[:, m0:m1] = V, W, U1, U2
        new = torch.cat([Vh[:, m0:m1], Wh[:, m0:m1], U1h[:, m0:m1], U2h[:, m0:m1]], dim=1)
        old = torch.cat([Vh[:, :m1], Wh[:, :m1], U1h[:, :m1], U2h[:, :m1]], dim=1)
        G = _tn(new, old).cpu().numpy()                          # (4p, 4 m1)
        blk = lambda i, j: G[i * p:(i + 1) * p, j * m1:(j + 1) * m1]
        nV, nW, nU1, nU2 = 0, 1, 2, 3
        # new rows of V^T U1 and of U1^T V (the new columns of U1^T V are the
        # transposes of the U1_new^T V rows taken against the old V rows)
        a_rows = 0.5 * (blk(nV, nU1) + blk(nU1, nV)) + 0.5 * (blk(nW, nU2) + blk(nU2, nW))
        b_rows = 0.5 * (blk(nV, nU2) + blk(nU2, nV)) + 0.5 * (blk(nW, nU1) + blk(nU1, nW))
        s_rows = blk(nV, nV) - blk(nW, nW)
        p_rows = blk(nV, nW) - blk(nW, nV)
        for M_, rows_, sign in ((a, a_rows, 1), (b, b_rows, 1), (sg, s_rows, 1), (pi, p_rows, -1)):
            M_[m0:m1, :m1] = rows_
            M_[:m0, m0:m1] = sign * rows_[:, :m0].T
        m = m1

    def orthonormal_new(tx, ty):
        """The corrections projected twice off the held pairs and their
        partners, then orthonormalised with their own partners."""
        for _ in range(2):
            if m:
                ox = _tn(Vh[:, :m], tx) + _tn(Wh[:, :m], ty)
                oy = _tn(Vh[:, :m], ty) + _tn(Wh[:, :m], tx)
                tx = tx - Vh[:, :m] @ ox - Wh[:, :m] @ oy
                ty = ty - Wh[:, :m] @ ox - Vh[:, :m] @ oy
        s11 = (_tn(tx, tx) + _tn(ty, ty)).cpu().numpy()
        s21 = (_tn(tx, ty) + _tn(ty, tx)).cpu().numpy()
        with _host_lapack():
            coef = symmetric_coefficients(s11, s21, lindep)
        if coef is None:
            return tx[:, :0], ty[:, :0]
        ca = torch.as_tensor(coef[0], device=dev)
        cb = torch.as_tensor(coef[1], device=dev)
        return tx @ ca + ty @ cb, ty @ ca + tx @ cb

    x0 = x0.to(torch.float64)
    y0 = torch.zeros_like(x0) if y0 is None else y0.to(device=x0.device, dtype=torch.float64)
    V0, W0 = orthonormal_new(x0, y0)
    append(V0, W0, *act(V0, W0))
    trace = []
    conv = np.zeros(nroots, bool)
    w = x = y = rn = None
    cycles = 0
    mark('setup')
    for cycle in range(max_cycle):
        cycles = cycle + 1
        k = min(nroots, m)
        try:
            with _host_lapack():
                w, x, y = subspace_solve(a[:m, :m], b[:m, :m], sg[:m, :m], pi[:m, :m], k)
        except np.linalg.LinAlgError as exc:
            # whether the trial space had spanned the pair space when the
            # projected (A - B) block broke down (the caller may then solve
            # densely)
            exc.space_exhausted = m >= n or exhausted
            raise
        mark('subspace')
        xt = torch.as_tensor(x, device=dev)
        yt = torch.as_tensor(y, device=dev)
        wt = torch.as_tensor(w, device=dev)
        X = Vh[:, :m] @ xt + Wh[:, :m] @ yt
        Y = Wh[:, :m] @ xt + Vh[:, :m] @ yt
        Rx = U1h[:, :m] @ xt + U2h[:, :m] @ yt - X * wt
        Ry = U2h[:, :m] @ xt + U1h[:, :m] @ yt + Y * wt
        rn = torch.sqrt((Rx * Rx).sum(0) + (Ry * Ry).sum(0)).cpu().numpy()
        conv = rn <= tol
        trace.append({'cycle': cycle, 'space': int(m), 'max_residual': float(rn.max()),
                      'converged': int(conv.sum())})
        mark('residuals')
        if conv.all() or len(w) < k:
            break
        idx = np.flatnonzero(~conv)
        wi = wt[idx]
        den_x = diag[:, None] - wi[None, :]
        den_y = diag[:, None] + wi[None, :]
        den_x = torch.where(den_x.abs() < floor, torch.full_like(den_x, floor), den_x)
        den_y = torch.where(den_y.abs() < floor, torch.full_like(den_y, floor), den_y)
        tx, ty = Rx[:, idx] / den_x, Ry[:, idx] / den_y
        nrm = torch.sqrt((tx * tx).sum(0) + (ty * ty).sum(0)).clamp_min(1e-300)
        mark('correction')
        tx, ty = orthonormal_new(tx / nrm, ty / nrm)
        mark('orthonormal')
        if tx.shape[1] == 0:
            exhausted = True
            break
        if m + tx.shape[1] > max_space:
            # restart on the lowest Ritz pairs, their products combined from
            # the held ones (the new directions stay orthogonal to them)
            kk = len(w)
            xk, yk = xt, yt
            nV, nW = X[:, :kk].clone(), Y[:, :kk].clone()
            nU1 = U1h[:, :m] @ xk + U2h[:, :m] @ yk
            nU2 = U2h[:, :m] @ xk + U1h[:, :m] @ yk
            m = 0
            append(nV, nW, nU1, nU2)
        mark('collapse')
        prods = act(tx, ty)
        mark('sigma')
        append(tx, ty, *prods)
        mark('append')
    if len(w) < nroots:
        warnings.warn(f'Casida Davidson: only {len(w)} roots in a space of {m}',
                      RuntimeWarning, stacklevel=2)
    return CasidaResult(e=np.asarray(w), X=X, Y=Y, conv=conv, rnorm=rn, cycles=cycles,
                        vectors=vectors, seconds=time.time() - t0, trace=trace,
                        exhausted=exhausted or m >= n)
