"""The Davidson products of dvd.hip against FP64 NumPy: the two-range A^T B
(ragged sizes, ranges of one buffer, a single row), the small-K update
C = beta C + alpha A Y into a column range, the Ritz vectors with their
residuals, and the Jacobi-Davidson correction (Solvers/davidson.py:43-77)
against the torch restatement `jacobi_davidson_correction`."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device='cuda')


@pytest.mark.parametrize('n', [1, 95, 129, 1953, 4097, 70000])
def test_tn_ranges(n):
    from mbptcode_amd.subspace import Ranges, tn_ranges
    rng = np.random.default_rng(n)
    buf = rng.normal(size=(n, 300))
    d = dev(buf)
    print(f'tn_ranges n={n}', flush=True)
    for (ra, rb) in ((((0, 161), (200, 30)), ((161, 30),)),
                     (((5, 1),), ((0, 3), (250, 17))),
                     (((0, 70), (100, 64)), ((0, 64), (200, 40)))):
        A, B = Ranges(d, *ra), Ranges(d, *rb)
        got = tn_ranges(A, B).cpu().numpy()
        Ah = np.concatenate([buf[:, a:a + c] for a, c in ra], axis=1)
        Bh = np.concatenate([buf[:, a:a + c] for a, c in rb], axis=1)
        ref = Ah.T @ Bh
        assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), (ra, rb)


@pytest.mark.parametrize('n', [7, 65, 1953, 70000])
def test_nn_ranges_and_ritz(n):
    from mbptcode_amd.subspace import Ranges, nn_ranges, ritz_residual
    rng = np.random.default_rng(n + 1)
    buf = rng.normal(size=(n, 260))
    d = dev(buf)
    print(f'nn_ranges / ritz n={n}', flush=True)
    for k, c in ((161, 30), (3, 1), (40, 64), (17, 17), (230, 64)):
        Y = rng.normal(size=(k, c))
        C = dev(rng.normal(size=(n, 70)))
        C0 = C.cpu().numpy()
        A = Ranges(d, (0, min(k, 100)), (130, k - min(k, 100)))
        nn_ranges(A, dev(Y), C, 3, alpha=-1.0, beta=1.0)
        Ah = np.concatenate([buf[:, :min(k, 100)], buf[:, 130:130 + k - min(k, 100)]], axis=1)
        ref = C0.copy()
        ref[:, 3:3 + c] = C0[:, 3:3 + c] - Ah @ Y
        got = C.cpu().numpy()
        assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), (k, c)
        nn_ranges(A, dev(Y), C, 0)
        assert np.abs(C.cpu().numpy()[:, :c] - Ah @ Y).max() <= 1e-12 * np.abs(Ah @ Y).max()
    m, k = 120, 30
    Y = rng.normal(size=(m, k))
    th = rng.normal(size=k)
    X = torch.empty((n, k), dtype=torch.float64, device='cuda')
    R = torch.empty_like(X)
    ritz_residual(d, m, 0, 130, dev(Y), dev(th), X, R)
    Xr = buf[:, :m] @ Y
    Rr = buf[:, 130:130 + m] @ Y - Xr * th
    assert np.abs(X.cpu().numpy() - Xr).max() <= 1e-12 * np.abs(Xr).max()
    assert np.abs(R.cpu().numpy() - Rr).max() <= 1e-12 * np.abs(Rr).max()


@pytest.mark.parametrize('n,k', [(5, 23), (1953, 23), (70000, 23), (300, 1), (1953, 70)])
def test_jd_correction(n, k):
    from mbptcode_amd.davidson import jacobi_davidson_correction
    from mbptcode_amd.subspace import jd_correction
    rng = np.random.default_rng(n + 2)
    U, R = dev(rng.normal(size=(n, k))), dev(rng.normal(size=(n, k)))
    diag = dev(np.sort(rng.uniform(0.1, 3.0, n)))
    theta = dev(rng.uniform(0.1, 1.0, k))
    # pin one denominator under the floor
    diag[0] = theta[0] + 1e-12
    ref = jacobi_davidson_correction(R, U, diag, theta)
    T = torch.full((n, k + 5), float('nan'), dtype=torch.float64, device='cuda')
    jd_correction(U, R, diag, theta, T, t_off=2)
    got = T[:, 2:2 + k]
    scale = float(ref.abs().max())
    assert float((got - ref).abs().max()) <= 1e-12 * scale
    assert torch.isnan(T[:, :2]).all() and torch.isnan(T[:, 2 + k:]).all()


@pytest.mark.parametrize('n,m,p', [(1, 0, 3), (1953, 40, 30), (70000, 161, 64)])
def test_block_scale_against_numpy(n, m, p):
    """The unit scaling of a projected correction block: norms after the
    projections from G1 - S2^T S2, keep flags against the norms before, and
    the block's columns scaled (kept) or zeroed (dropped)."""
    from mbptcode_amd.subspace import block_scale
    rng = np.random.default_rng(n + m + p)
    S1 = rng.normal(size=(m, p))
    T = rng.normal(size=(n, p))
    G0 = T.T @ T
    S2 = 1e-3 * rng.normal(size=(m, p))
    T1 = T.copy()
    T1[:, :2] *= 1e-9      # two columns that the projections all but removed
    G1 = T1.T @ T1 + S2.T @ S2
    stage = np.concatenate([S1, G0, rng.normal(size=(m, p)), S2, G1, np.zeros((2, p))])
    buf = rng.normal(size=(n, 3 * p + 5))
    off = p + 2
    buf[:, off:off + p] = T1
    drop = 1e-6
    sd, bd = dev(stage), dev(buf)
    print(f'block_scale n={n} m={m} p={p}', flush=True)
    block_scale(sd, m, p, drop, bd, off)
    got_s, got_b = sd.cpu().numpy(), bd.cpu().numpy()
    after = np.sqrt(np.clip(np.diag(G1) - (S2 * S2).sum(0), 0, None))
    before = np.sqrt(np.clip(np.diag(G0), 0, None))
    keep = after > drop * np.maximum(before, 1e-300)
    assert not keep[:2].any() and keep[2:].all()
    # a dropped column's norm is a difference far below the terms it comes
    # from: it matches to the rounding of those terms (sum order differs)
    assert np.allclose(got_s[-2][keep], after[keep], rtol=1e-12, atol=0)
    assert np.allclose(got_s[-2][~keep], after[~keep], rtol=0,
                       atol=1e-14 * np.sqrt(np.diag(G1).max()))
    assert np.array_equal(got_s[-1], keep.astype(float))
    assert np.array_equal(got_s[:-2], stage[:-2])
    ref = buf.copy()
    ref[:, off:off + p] = T1 * np.where(keep, 1.0 / np.maximum(after, 1e-300), 0.0)
    assert np.abs(got_b - ref).max() <= 1e-13 * np.abs(ref).max()
