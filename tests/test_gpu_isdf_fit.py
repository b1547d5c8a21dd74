"""The separable fit on the device (FP64 GEMMs, Cholesky and eigh through
torch on ROCm) against the FP64 oracle (separable_ri.py:1616-1712,
:1802-1930): relative 1e-10 on M, the metric root and the factor, and
Z = D D^T = M^T V M."""
import numpy as np
import pytest
import torch

from oracle import isdf_oracle as O

pytestmark = pytest.mark.gpu


def rel(got, ref):
    got = got.cpu().numpy() if isinstance(got, torch.Tensor) else got
    return float(np.abs(got - ref).max() / np.abs(ref).max())


@pytest.mark.parametrize('nk, nrho, naux', [(50, 300, 20), (700, 4000, 300)])
def test_fit_against_oracle(nk, nrho, naux):
    from mbptcode_amd import isdf_fit as F
    rng = np.random.default_rng(nk)
    D = rng.normal(size=(nk, nrho)) * rng.uniform(0.5, 2.0, size=(nk, 1))
    Ff = rng.normal(size=(naux, nrho))
    print(f'isdf fit {nk} {nrho} {naux}', flush=True)
    assert rel(F.fit_M_stable(D, Ff), O.fit_M_stable(D, Ff)) < 1e-10
    assert rel(F.fit_M(D, Ff), O.fit_M(D, Ff)) < 1e-9
    B = rng.normal(size=(naux, naux))
    V = B @ B.T / naux + 0.05 * np.eye(naux)
    assert rel(F.aux_metric_sqrt(V), O.aux_metric_sqrt(V)) < 1e-10
    M = O.fit_M_stable(D, Ff)
    Dsep = F.separable_factor(M, V)
    Z = F.screened_points(M, V)
    assert rel(Dsep, M.T @ O.aux_metric_sqrt(V)) < 1e-10
    assert rel(Dsep @ Dsep.T, Z.cpu().numpy()) < 1e-10


def test_planted_matrix_on_the_device():
    from mbptcode_amd import isdf_fit as F
    rng = np.random.default_rng(7)
    D = rng.normal(size=(120, 500))
    M = rng.normal(size=(40, 120))
    got = F.fit_M_stable(D, M @ D, regularization=0.0)
    assert rel(got, M) < 1e-9


@pytest.mark.parametrize('widths', [(600,), (128, 1000, 77, 495)])
def test_streaming_fit_against_the_whole_fit(widths):
    """fit_M_streaming over column blocks of (D, F), the last block the
    auxiliary one (F = I), against the oracle's fit_M_stable on the
    concatenated test set; a zero row (balanced by 1) included.  The test
    set is wider than the points, so the balanced Gram matrix is well
    conditioned and the two orders of summation agree to 1e-10 (with fewer
    columns than points it is conditioned by the 4e-7 floor alone, and the
    reordering's rounding grows by that factor)."""
    from mbptcode_amd import isdf_fit as F
    rng = np.random.default_rng(len(widths))
    nk, naux = 400, 90
    Ds = [rng.normal(size=(nk, w)) * rng.uniform(0.5, 2.0, size=(nk, 1)) for w in widths]
    Fs = [rng.normal(size=(naux, w)) for w in widths]
    Ds.append(0.3 * rng.normal(size=(nk, naux)))
    Fs.append(np.eye(naux))
    for Db in Ds:
        Db[17] = 0.0
    D, Ff = np.hstack(Ds), np.hstack(Fs)
    print(f'isdf streaming fit {widths}', flush=True)
    timings = {}
    got = F.fit_M_streaming(zip(Ds, Fs), timings=timings)
    assert rel(got, O.fit_M_stable(D, Ff)) < 1e-10
    assert timings['fit_blocks_total'] == len(Ds)
    with pytest.raises(ValueError):
        F.fit_M_streaming([(Ds[0], Fs[0][:, :5])])
    with pytest.raises(ValueError):
        F.fit_M_streaming([])
