"""Host logic of the batched Davidson on the CPU (dense operators, torch's
eigh in place of the device Jacobi kernel): every instance's roots against
LAPACK, through trial-space collapses, ragged pair counts and instances that
converge at different cycles; the batched orthonormalisation's contract."""
import numpy as np
import torch

from mbptcode_amd.batched import (batched_guess, davidson_batched, jd_correction_batched,
                                  orthonormalize_batched)
from mbptcode_amd.davidson import tracked_roots


def dense_batch(sizes, seed=0):
    rng = np.random.default_rng(seed)
    mats = []
    for n in sizes:
        G = rng.normal(size=(n, n)) / np.sqrt(n)
        mats.append(np.diag(np.sort(rng.uniform(0.1, 3.0, n))) + 0.05 * (G + G.T))
    return mats


def batch_apply(mats, P):
    As = [torch.as_tensor(A) for A in mats]

    def apply(T, live):
        out = torch.zeros_like(T)
        for b, A in enumerate(As):
            n = A.shape[0]
            out[b, :n] = A @ T[b, :n]
        mask = torch.arange(T.shape[2])[None, :] < live[:, None]
        return out * mask[:, None, :]
    return apply


def test_davidson_batched_matches_lapack():
    sizes = [30, 45, 60, 12]
    mats = dense_batch(sizes)
    P = max(sizes)
    d = np.full((len(sizes), P), 1e6)
    for b, A in enumerate(mats):
        d[b, :sizes[b]] = np.diag(A)
    nroots = 4
    track = max(tracked_roots(nroots, n) for n in sizes)
    x0, live0 = batched_guess(d, np.asarray(sizes), nroots, P, 'cpu')
    res = davidson_batched(batch_apply(mats, P), torch.as_tensor(d), track, x0, live0,
                           tol=1e-9, max_cycle=200, max_space=24)
    assert res.conv.all()
    assert any(t['space'] < res.trace[i - 1]['space'] for i, t in enumerate(res.trace) if i)
    for b, A in enumerate(mats):
        w = np.linalg.eigvalsh(A)[:nroots]
        assert np.abs(res.e[b, :nroots] - w).max() < 1e-12
        X = res.X[b, :sizes[b], :nroots].numpy()
        assert np.abs(A @ X - X * res.e[b, :nroots]).max() < 1e-8
        assert (res.X[b, sizes[b]:] == 0).all()


def test_orthonormalize_batched_drops_dependent_columns():
    rng = np.random.default_rng(1)
    B, P, m = 3, 40, 6
    V = torch.as_tensor(np.linalg.qr(rng.normal(size=(B, P, m)))[0])
    T = rng.normal(size=(B, P, 4))
    T[0, :, 2] = T[0, :, 0] + T[0, :, 1]              # dependent
    T[1, :, 3] = V[1, :, 2].numpy()                    # inside the space
    T[2] = 0.0                                         # nothing
    Tn, live = orthonormalize_batched(torch.as_tensor(T), V, 1e-7)
    assert live.tolist() == [3, 3, 0]
    for b in range(B):
        k = int(live[b])
        Q = Tn[b, :, :k].numpy()
        assert (Tn[b, :, k:] == 0).all()
        if k:
            assert np.abs(Q.T @ Q - np.eye(k)).max() < 1e-12
            assert np.abs(V[b].numpy().T @ Q).max() < 1e-12


def test_jd_correction_batched_is_orthogonal_to_ritz_vectors_in_k_metric():
    rng = np.random.default_rng(2)
    B, P, r = 2, 25, 3
    R = torch.as_tensor(rng.normal(size=(B, P, r)))
    U = torch.as_tensor(rng.normal(size=(B, P, r)))
    diag = torch.as_tensor(rng.uniform(1, 2, (B, P)))
    theta = torch.as_tensor(rng.uniform(0, 0.5, (B, r)))
    t = jd_correction_batched(R, U, diag, theta)
    # u^T t = 0 for every instance and root (the Jacobi-Davidson projection)
    assert torch.abs((U * t).sum(1)).max() < 1e-12
