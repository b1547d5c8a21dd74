"""The Davidson iteration's host logic (mbptcode_amd.davidson.davidson is
device-agnostic: run here on CPU tensors with dense operators), checked
against dense eigensolves.  CPU only."""
import numpy as np
import pytest
import torch

from mbptcode_amd.davidson import (davidson, guess_indices, jacobi_davidson_correction,
                                   lindep_for, tracked_roots)
from oracle import isdf_oracle as O


def model(n=400, seed=1, coupling=0.1, rank=30):
    rng = np.random.default_rng(seed)
    d = np.sort(rng.uniform(1, 10, n))
    d[1] = d[0]                                   # a degenerate pair on the diagonal
    Q = np.linalg.qr(rng.normal(size=(n, n)))[0]
    A = np.diag(d) + coupling * (Q[:, :rank] @ np.diag(rng.normal(size=rank)) @ Q[:, :rank].T)
    return A


def unit_guess(diag, count):
    n = len(diag)
    x0 = torch.zeros((n, count), dtype=torch.float64)
    x0[torch.as_tensor(np.argsort(diag)[:count]), torch.arange(count)] = 1.0
    return x0


@pytest.mark.parametrize('max_space', [None, 20])
def test_lowest_roots_of_a_model_matrix(max_space):
    A = model()
    At = torch.as_tensor(A)
    ref = np.linalg.eigvalsh(A)[:6]
    res = davidson(lambda V: At @ V, torch.as_tensor(np.diag(A).copy()), 6,
                   unit_guess(np.diag(A), 12), tol=1e-9, max_space=max_space)
    assert res.conv.all()
    assert np.abs(res.e - ref).max() < 1e-12
    assert res.rnorm.max() <= 1e-9
    X = res.X.numpy()
    assert np.abs(X.T @ X - np.eye(6)).max() < 1e-10


def test_inconsistent_operator_still_converges():
    """A lower-precision operator is neither exactly linear nor symmetric; the
    residual estimate must stay exact for it (x^T A x taken from the
    non-symmetric Rayleigh matrix), or the iteration stalls above tol."""
    A = model(seed=3)
    At = torch.as_tensor(A)
    gen = torch.Generator().manual_seed(0)

    def noisy(V):
        return At @ V + 1e-8 * torch.randn(V.shape, generator=gen, dtype=torch.float64)

    ref = np.linalg.eigvalsh(A)[:6]
    res = davidson(noisy, torch.as_tensor(np.diag(A).copy()), 6,
                   unit_guess(np.diag(A), 12), tol=1e-6)
    assert res.conv.all() and res.cycles < 30
    assert np.abs(res.e - ref).max() < 1e-7


def test_isdf_tda_roots_match_dense():
    from mbptcode_amd import synth
    inst = synth.config(0)
    D = O.grid_factors(inst.V)
    A = O.dense_tda_matrix(inst.eps, 5, inst.X, D, np.eye(120))
    d = np.diag(A).copy()
    track = tracked_roots(5, A.shape[0])
    idx = guess_indices(O.pair_diagonal(inst.eps, 5), 5)
    x0 = torch.zeros((A.shape[0], len(idx)), dtype=torch.float64)
    x0[torch.as_tensor(idx), torch.arange(len(idx))] = 1.0
    res = davidson(lambda V: torch.as_tensor(A) @ V, torch.as_tensor(d), track, x0,
                   tol=1e-8)
    assert res.conv.all()
    assert np.abs(res.e - np.linalg.eigvalsh(A)[:track]).max() < 1e-12


def test_jacobi_davidson_correction_is_orthogonal_to_the_ritz_vector():
    """u^T t = 0 for t = K^-1 r - alpha K^-1 u, alpha = u.K^-1 r / u.K^-1 u:
    the correction carries no component along u (Solvers/davidson.py:43-77)."""
    rng = np.random.default_rng(0)
    n = 50
    diag = torch.as_tensor(rng.uniform(1, 3, n))
    R = torch.as_tensor(rng.normal(size=(n, 3)))
    U = torch.as_tensor(rng.normal(size=(n, 3)))
    theta = torch.as_tensor([0.5, 1.7, 2.2])
    T = jacobi_davidson_correction(R, U, diag, theta)
    assert float((U * T).sum(0).abs().max()) < 1e-12 * float(T.abs().max())
    # and the projection did act (random u is not K-orthogonal to r)
    plain = R / (diag[:, None] - theta[None, :])
    assert float((T - plain).abs().max()) > 0


def test_lindep_follows_the_tolerance():
    assert lindep_for(1e-5) == pytest.approx(1e-12)
    assert lindep_for(1e-12) == 1e-22
    assert lindep_for(1.0) == 1e-14


def test_overlap_pick_ranks_by_overlap_and_breaks_ties_by_energy():
    """overlap_pick (Solvers/davidson.py:80-101) on pyscf's envs['xs'] form:
    the Ritz vector with the largest projection on the reference span wins;
    overlaps within 1e-6 tie and go to the lower energy; the chosen roots come
    back ascending."""
    from mbptcode_amd.davidson import overlap_pick
    n = 6
    xs = list(np.eye(n))                       # trial vectors = unit vectors
    w = np.array([-1.0, 0.5, 0.2, 0.9, 2.0, 3.0])
    v = np.eye(n)                              # Ritz vectors = trial vectors
    ref = np.zeros(n)
    ref[3] = 1.0
    we, ve, order = overlap_pick(ref)(w, v, 1, {'xs': xs})
    assert order.tolist() == [3] and we.tolist() == [0.9]
    # two roots: root 3 first by overlap, then the lowest of the zero-overlap tie
    we, ve, order = overlap_pick(ref)(w, v, 2, {'xs': xs})
    assert order.tolist() == [0, 3] and np.all(np.diff(we) > 0)
    # the device form (V^T span) gives the same choice
    we2, _, order2 = overlap_pick(ref)(w, v, 2, {'ref_coeff': np.eye(n)[:, [3]]})
    assert order2.tolist() == order.tolist()
    # a two-column reference follows the span of both
    ref2 = np.zeros((n, 2))
    ref2[4, 0] = ref2[5, 1] = 1.0
    _, _, order = overlap_pick(ref2)(w, v, 2, {'xs': xs})
    assert order.tolist() == [4, 5]
