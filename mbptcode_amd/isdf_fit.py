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
