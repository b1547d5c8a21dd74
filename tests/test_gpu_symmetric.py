"""`solve_symmetric` (Solvers/davidson.py:179-326) on the device against a
dense eigensolve, on the reference tests' model (tests/test_davidson_symmetric.py
`_test_matrix`, restated here: a diagonally dominant symmetric matrix of 400
with a wide diagonal and an exactly degenerate pair among the lowest roots,
rows 1 and 2 decoupled):

1. the lowest 6 against eigvalsh, both members of the degenerate pair;
2. a starved solve reports conv False and warns;
3. tol_residual 1e-9 is reached;
4. overlap_pick follows an interior root (Solvers/davidson.py:80-101);
5. complete_lowest finds a degenerate copy that the start never reaches
   (Solvers/davidson.py:362-404).
"""
import warnings

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def model(n=400, offdiag=0.02, seed=0):
    rng = np.random.default_rng(seed)
    A = rng.normal(scale=offdiag, size=(n, n))
    A = 0.5 * (A + A.T)
    A[np.diag_indices(n)] = np.sort(rng.uniform(-1.0, 3.0, n))
    A[[1, 2], :] = 0.0
    A[:, [1, 2]] = 0.0
    A[1, 1] = A[2, 2] = A.diagonal().min() - 0.5
    return A


class Dense:
    def __init__(self, A):
        self.A = torch.as_tensor(A, device='cuda')

    def matmat(self, V):
        return self.A @ V


def test_lowest_roots_with_a_degenerate_pair():
    from mbptcode_amd.davidson import solve_symmetric
    A = model()
    w = np.linalg.eigvalsh(A)
    print('solve_symmetric lowest 6', flush=True)
    e, X, conv = solve_symmetric(Dense(A), np.diag(A), nroots=6, tol_residual=1e-8)
    Xh = X.cpu().numpy()
    assert conv.all() and np.abs(e - w[:6]).max() < 1e-10
    assert np.linalg.norm(A @ Xh - Xh * e, axis=0).max() < 1e-8
    pair = np.abs(e - w[0]) < 1e-9
    assert pair.sum() == 2 and abs(Xh[:, pair][:, 0] @ Xh[:, pair][:, 1]) < 1e-6


def test_a_starved_solve_says_so():
    from mbptcode_amd.davidson import solve_symmetric
    A = model()
    with warnings.catch_warnings(record=True) as rec:
        warnings.simplefilter('always')
        _, _, conv = solve_symmetric(Dense(A), np.diag(A), nroots=6, tol_residual=1e-10,
                                     max_cycle=2)
    assert not conv.all()
    assert any(issubclass(r.category, RuntimeWarning) for r in rec)


def test_tight_tolerance_is_reached():
    from mbptcode_amd.davidson import solve_symmetric
    A = model()
    e, X, conv = solve_symmetric(Dense(A), np.diag(A), nroots=4, tol_residual=1e-9)
    Xh = X.cpu().numpy()
    assert conv.all() and np.linalg.norm(A @ Xh - Xh * e, axis=0).max() < 1e-9


def test_overlap_pick_follows_an_interior_root():
    from mbptcode_amd.davidson import overlap_pick, solve_symmetric
    A = model(offdiag=0.002)
    w, v = np.linalg.eigh(A)
    row = 300
    target = int(np.argmax(np.abs(v[row])))
    ref = np.zeros(A.shape[0])
    ref[row] = 1.0
    print(f'overlap_pick: interior root {target}', flush=True)
    e, X, conv = solve_symmetric(Dense(A), np.diag(A), nroots=1, x0=ref,
                                 pick=overlap_pick(ref), tol_residual=1e-8)
    assert target > 10 and conv.all()
    assert abs(e[0] - w[target]) < 1e-9
    assert abs(X[:, 0].cpu().numpy() @ v[:, target]) > 0.99


def test_complete_lowest_finds_a_copy_the_start_misses():
    """Start from unit vectors that leave out row 2, the decoupled twin of
    row 1: no trial vector ever reaches it, so the plain solve converges on
    the next root up; the completion solve on the complement finds it."""
    from mbptcode_amd.davidson import solve_symmetric
    A = model()
    w = np.linalg.eigvalsh(A)
    d = np.diag(A)
    cols = [c for c in np.argsort(d) if c != 2][:10]
    x0 = np.zeros((A.shape[0], len(cols)))
    x0[cols, np.arange(len(cols))] = 1.0
    e0, _, conv0 = solve_symmetric(Dense(A), d, nroots=3, x0=x0, tol_residual=1e-8,
                                   complete_lowest=False)
    assert conv0.all() and np.abs(e0 - w[:3]).max() > 1e-3     # the copy is missed
    e, X, conv = solve_symmetric(Dense(A), d, nroots=3, x0=x0, tol_residual=1e-8)
    assert conv.all() and np.abs(e - w[:3]).max() < 1e-9
