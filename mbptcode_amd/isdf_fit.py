"""The separable (ISDF) factorization's format and fit, on the device in FP64
(Base/factorization/separable_ri.py:1616-1653, :1656-1712, :1802-1930;
sliced_factors.py:333-336).

The fit of Duchemin and Blase's eqs 8-9: a test set of co-densities rho,
sampled on the interpolation points as D[k, rho] and fitted in the auxiliary
basis as F[beta, rho], gives the interpolation matrix

    M = (F Dt^T) (Dt Dt^T + lambda I)^-1 diag(d),   Dt = diag(d) D,

with the rows of D balanced to unit length (d their inverse norms) and the
Tikhonov floor lambda on the balanced Gram matrix.  The factors the
sigma-vector and chi0 read are then X[k, mu] (the collocation), Z = M^T V M
on the points and its factor D = M^T V^1/2 (the 'separable' gauge, V^1/2 the
metric's square root with its numerical null space dropped).

Building D and F needs the AO collocation and the three-centre integrals,
which come from pyscf and are inputs here; `SlicedFactors` (the grid rows per
rank) is a multi-rank format and not part of this one-GPU build.
"""
import numpy as np
import torch

from .constants import AUX_METRIC_ROOT_FLOOR, ISDF_FIT_REGULARIZATION


def _dev(a, device=None):
    dev = device or torch.device('cuda', torch.cuda.current_device())
    if isinstance(a, torch.Tensor):
        return a.to(device=dev, dtype=torch.float64)
    return torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev)


def unpack_factors(factors):
    """(X_mo, D, X_ao, coords) of a factor tuple; the last two None for a
    shorter tuple (sliced_factors.py:333-336)."""
    factors = tuple(factors)
    if not 2 <= len(factors) <= 4:
        raise ValueError('a factor tuple is (X_mo, D[, X_ao[, coords]])')
    return factors + (None,) * (4 - len(factors))


def balanced_gram(D, regularization=ISDF_FIT_REGULARIZATION):
    """(Dt, d, G): D's rows scaled to unit length (d the inverse norms, 1 for a
    zero row) and their Gram matrix with the Tikhonov floor on its diagonal
    (separable_ri.py:1616-1627).  D (nk, nrho) FP64, device."""
    D = _dev(D)
    scale = torch.sqrt((D * D).sum(1))
    scale = torch.where(scale == 0, torch.ones_like(scale), scale)
    d = 1.0 / scale
    Dt = D * d[:, None]
    G = Dt @ Dt.T
    G.diagonal().add_(float(regularization))
    return Dt, d, G


def fit_M(D, F, regularization=ISDF_FIT_REGULARIZATION):
    """M (naux, nk) by the explicit inverse of the balanced Gram matrix, the
    reference estimator (separable_ri.py:1630-1637)."""
    Dt, d, G = balanced_gram(D, regularization)
    return ((_dev(F) @ Dt.T) @ torch.linalg.inv(G)) * d[None, :]


def fit_M_stable(D, F, regularization=ISDF_FIT_REGULARIZATION):
    """The same estimator through a Cholesky solve of the balanced Gram matrix
    (separable_ri.py:1640-1653)."""
    Dt, d, G = balanced_gram(D, regularization)
    C, info = torch.linalg.cholesky_ex(G)
    if int(info) != 0:
        raise np.linalg.LinAlgError('the balanced Gram matrix is not positive definite')
    R = (_dev(F) @ Dt.T).T                       # (nk, naux)
    return torch.cholesky_solve(R, C).T * d[None, :]


def fit_M_streaming(blocks, regularization=ISDF_FIT_REGULARIZATION, timings=None):
    """M (naux, nk) without holding D or F whole: the pyscf-free core of the
    reference's streaming fit (separable_ri.py:664-934).  `blocks` yields
    (D_blk (nk, nb), F_blk (naux, nb)) column blocks of the test set (the
    caller's three-centre pass, with the auxiliary block D = the auxiliary
    functions on the points, F = I among them); only

        S = sum D_blk D_blk^T  (nk, nk),   FD = sum F_blk D_blk^T  (naux, nk)

    are kept.  The row norms of D are diag(S), so the balancing is
    G = d S d + lambda I and the solve is one Cholesky: M = ((FD d) G^-1) d,
    fit_M_stable's estimator up to the order of the sums.  `fit='rows'`
    (grid tiles over ranks) is a multi-rank layout and not built.  timings,
    a dict: 'fit_blocks', 'fit_gram' and 'fit_cholesky' seconds (device
    synchronised)."""
    import time
    S = FD = None
    nblk = 0
    t0 = time.perf_counter()
    for D_blk, F_blk in blocks:
        Db, Fb = _dev(D_blk), _dev(F_blk)
        if Db.shape[1] != Fb.shape[1]:
            raise ValueError(f'block {nblk}: D has {Db.shape[1]} columns, F {Fb.shape[1]}')
        if S is None:
            S = torch.zeros((Db.shape[0], Db.shape[0]), dtype=torch.float64, device=Db.device)
            FD = torch.zeros((Fb.shape[0], Db.shape[0]), dtype=torch.float64, device=Db.device)
        elif Db.shape[0] != S.shape[0] or Fb.shape[0] != FD.shape[0]:
            raise ValueError(f'block {nblk} changes nk or naux')
        S.addmm_(Db, Db.T)
        FD.addmm_(Fb, Db.T)
        nblk += 1
    if S is None:
        raise ValueError('no test-set blocks')
    if timings is not None:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        timings['fit_blocks'] = t1 - t0
        timings['fit_blocks_total'] = nblk
        t0 = t1
    scale = torch.sqrt(torch.clamp(torch.diagonal(S), min=0.0))
    scale = torch.where(scale == 0, torch.ones_like(scale), scale)
    d = 1.0 / scale
    G = S.mul_(d[:, None]).mul_(d[None, :])      # balanced in place
    G.diagonal().add_(float(regularization))
    if timings is not None:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        timings['fit_gram'] = t1 - t0
        t0 = t1
    C, info = torch.linalg.cholesky_ex(G)
    if int(info) != 0:
        raise np.linalg.LinAlgError(
            f'Cholesky of the balanced Gram matrix failed at leading minor {int(info)}: '
            'the points are degenerate')
    M = torch.cholesky_solve((FD * d[None, :]).T, C).T * d[None, :]
    if timings is not None:
        torch.cuda.synchronize()
        timings['fit_cholesky'] = time.perf_counter() - t0
    return M


def aux_metric_sqrt(V, floor=AUX_METRIC_ROOT_FLOOR):
    """V^1/2 of a symmetric metric, directions below `floor` x the largest
    eigenvalue dropped (its numerical null space; separable_ri.py:1802-1836,
    :1915-1929 for the bare metric)."""
    V = _dev(V)
    w, v = torch.linalg.eigh(0.5 * (V + V.T))
    keep = w > floor * w.max()
    vk = v[:, keep]
    return (vk * torch.sqrt(w[keep])[None, :]) @ vk.T


def separable_factor(M, V):
    """D = M^T V^1/2 (nk, naux): the separable gauge's factor, Z = D D^T =
    M^T V M on the points (separable_ri.py:1656-1712)."""
    return _dev(M).T @ aux_metric_sqrt(V)


def screened_points(M, V):
    """Z = M^T V M (nk, nk), the bare interaction on the interpolation points."""
    Md = _dev(M)
    return Md.T @ (_dev(V) @ Md)
