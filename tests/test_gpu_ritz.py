"""The device Rayleigh-Ritz eigensolver (ritz.hip) against LAPACK: the
lowest k eigenpairs of the symmetric part of an m x m block of a wider
row-major buffer, on random, degenerate, diagonal, clustered and slightly
non-symmetric matrices, and the quadratic forms y^T G y, y^T H y."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def run(H, G, k, ld_pad=7):
    from mbptcode_amd.subspace import ritz_eigh
    m = H.shape[0]
    buf = np.zeros((m, 2 * (m + ld_pad)))
    buf[:, :m] = H
    buf[:, m + ld_pad:2 * m + ld_pad] = G
    d = torch.as_tensor(buf, device='cuda')
    Y, wq = ritz_eigh(d[:, :m], d[:, m + ld_pad:2 * m + ld_pad], k)
    torch.cuda.synchronize()
    return Y.cpu().numpy(), wq.cpu().numpy()


def check(H, k, G=None, tol=1e-12):
    m = H.shape[0]
    G = H @ H.T if G is None else G
    S = 0.5 * (H + H.T)
    Y, wq = run(H, G, k)
    w = wq[:k]
    ref = np.linalg.eigvalsh(S)[:k]
    scale = max(np.abs(S).max(), 1e-300) * m
    assert np.abs(w - ref).max() <= tol * scale, np.abs(w - ref).max()
    assert np.abs(S @ Y - Y * w).max() <= 100 * tol * scale
    assert np.abs(Y.T @ Y - np.eye(k)).max() <= 1e-9
    assert np.allclose(wq[k:2 * k], np.einsum('ij,ik,kj->j', Y, G, Y), rtol=1e-12,
                       atol=1e-12 * np.abs(G).max() * m)
    assert np.allclose(wq[2 * k:], np.einsum('ij,ik,kj->j', Y, H, Y), rtol=1e-12,
                       atol=1e-12 * np.abs(H).max() * m)


@pytest.mark.parametrize('m', [1, 2, 3, 16, 17, 40, 64, 65, 100, 129, 144, 145, 160, 161, 176,
                               177, 192])
def test_random_symmetric(m):
    rng = np.random.default_rng(m)
    A = rng.normal(size=(m, m))
    H = A + A.T
    print(f'ritz random m={m}', flush=True)
    for k in sorted({1, min(m, 30), min(m, 64)}):
        check(H, k)


def test_a_davidson_like_spectrum():
    """A Rayleigh matrix like the Davidson's: Ritz values in a narrow band
    with close pairs, a wide top, a slightly non-symmetric operator."""
    rng = np.random.default_rng(3)
    m = 161
    low = 0.32 + 0.02 * np.sort(rng.random(40))
    low[7] = low[6] + 1e-5
    low[21] = low[20] + 3e-9
    lam = np.concatenate([low, 0.5 + 2.5 * rng.random(m - 40)])
    Q = np.linalg.qr(rng.normal(size=(m, m)))[0]
    H = (Q * lam) @ Q.T
    H += 1e-9 * rng.normal(size=(m, m))
    print('ritz davidson-like m=161', flush=True)
    check(H, 30)
    check(H, 64)


def test_degenerate_and_diagonal():
    rng = np.random.default_rng(5)
    m = 50
    lam = np.repeat(np.arange(1.0, 11.0), 5)      # five-fold degenerate
    Q = np.linalg.qr(rng.normal(size=(m, m)))[0]
    print('ritz degenerate m=50', flush=True)
    check((Q * lam) @ Q.T, 20, tol=1e-11)
    D = np.diag(np.repeat([3.0, -1.0, 2.0, 0.5], [10, 10, 10, 10]))
    print('ritz diagonal m=40', flush=True)
    check(D, 25)
    print('ritz zero matrix m=9', flush=True)
    Y, wq = run(np.zeros((9, 9)), np.zeros((9, 9)), 4)
    assert np.abs(wq[:4]).max() < 1e-300 and np.abs(Y.T @ Y - np.eye(4)).max() < 1e-12


def test_tridiagonal_with_split_off_diagonals():
    """Zero and tiny couplings split the matrix into independent blocks."""
    m = 120
    rng = np.random.default_rng(9)
    d = rng.normal(size=m)
    e = rng.normal(size=m - 1)
    e[[10, 55, 90]] = 0.0
    e[[30, 70]] = 1e-200
    H = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    print('ritz split tridiagonal m=120', flush=True)
    check(H, 40)


def test_refuses_out_of_range():
    from mbptcode_amd.subspace import ritz_eigh, ritz_max_m
    assert ritz_max_m() == 192
    big = torch.zeros((193, 193), dtype=torch.float64, device='cuda')
    with pytest.raises(ValueError):
        ritz_eigh(big, big, 10)
    small = torch.zeros((8, 8), dtype=torch.float64, device='cuda')
    with pytest.raises(ValueError):
        ritz_eigh(small, small, 9)


@pytest.mark.parametrize('max_space', [None, 60])
def test_davidson_with_device_ritz_matches_host_ritz(max_space):
    """The Davidson iteration with every Ritz step on the device eigensolver
    (threshold 1) and with host LAPACK: the same roots, both converged."""
    from mbptcode_amd.davidson import davidson
    rng = np.random.default_rng(11)
    n = 1500
    d = np.sort(rng.uniform(1, 10, n))
    d[1] = d[0]                                   # a degenerate pair on the diagonal
    Q = np.linalg.qr(rng.normal(size=(n, 40)))[0]
    A = np.diag(d) + 0.1 * (Q @ np.diag(rng.normal(size=40)) @ Q.T)
    Ad = torch.as_tensor(A, device='cuda')
    diag = torch.as_tensor(np.diag(A).copy(), device='cuda')
    x0 = torch.zeros((n, 16), dtype=torch.float64, device='cuda')
    x0[torch.as_tensor(np.argsort(np.diag(A))[:16], device='cuda'),
       torch.arange(16, device='cuda')] = 1.0
    ref = np.linalg.eigvalsh(A)[:8]
    out = {}
    for name, rmin in (('device', 1), ('host', 10 ** 6)):
        print(f'davidson ritz={name} max_space={max_space}', flush=True)
        res = davidson(lambda V: Ad @ V, diag, 8, x0.clone(), tol=1e-9, max_space=max_space,
                       ritz_min=rmin)
        assert res.conv.all()
        assert np.abs(res.e - ref).max() < 1e-11
        X = res.X.cpu().numpy()
        assert np.abs(X.T @ X - np.eye(8)).max() < 1e-9
        out[name] = res
    assert np.abs(out['device'].e - out['host'].e).max() < 1e-12
