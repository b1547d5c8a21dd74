"""The device A^T B kernel of the Davidson subspace against an FP64 product,
including column slices of a wider buffer and ragged sizes."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('n,a,b', [(1, 1, 1), (95, 7, 3), (1953, 80, 15),
                                   (24000, 130, 70), (70000, 24, 24)])
def test_tn_matches_fp64_product(n, a, b):
    from mbptcode_amd.subspace import tn
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, a))
    B = rng.normal(size=(n, b))
    got = tn(torch.as_tensor(A, device='cuda'), torch.as_tensor(B, device='cuda')).cpu().numpy()
    ref = A.T @ B
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max() * max(1, np.sqrt(n) / 10)


def test_tn_on_column_slices():
    from mbptcode_amd.subspace import tn
    rng = np.random.default_rng(0)
    buf = torch.as_tensor(rng.normal(size=(500, 40)), device='cuda')
    A, B = buf[:, 3:17], buf[:, 20:31]
    out = torch.full((14, 11), float('nan'), dtype=torch.float64, device='cuda')
    tn(A, B, out=out)
    ref = (A.cpu().numpy().T @ B.cpu().numpy())
    assert np.allclose(out.cpu().numpy(), ref, rtol=1e-12, atol=1e-12)
