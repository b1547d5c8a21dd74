"""The separable fit's oracle, pinned by what the estimator must do
(separable_ri.py:1616-1653, :1802-1930): a test set that the points
represent exactly gives back the planted interpolation matrix, balancing the
rows changes nothing, the inverse and Cholesky estimators agree, and the
metric root squares to V with its null space dropped.  Plus the factor
tuple's unpacking (sliced_factors.py:333-336)."""
import numpy as np
import pytest

from mbptcode_amd.isdf_fit import unpack_factors
from oracle import isdf_oracle as O


def planted(nk=40, nrho=160, naux=25, seed=0):
    rng = np.random.default_rng(seed)
    D = rng.normal(size=(nk, nrho)) * rng.uniform(0.1, 10.0, size=(nk, 1))
    M = rng.normal(size=(naux, nk))
    return D, M, M @ D


def test_exact_test_set_recovers_the_planted_matrix():
    D, M, F = planted()
    for fit in (O.fit_M, O.fit_M_stable):
        got = fit(D, F, regularization=0.0)
        assert np.abs(got - M).max() < 1e-9 * np.abs(M).max()
    # the Tikhonov floor biases it by about lambda / (smallest balanced singular value^2)
    got = O.fit_M_stable(D, F)
    assert 1e-12 < np.abs(got - M).max() / np.abs(M).max() < 1e-4


def test_row_balancing_is_invisible():
    """Scaling the test co-densities' rows (the points) rescales M's columns
    inversely and nothing else: the balanced Gram matrix does not see it."""
    D, M, F = planted(seed=1)
    s = np.random.default_rng(2).uniform(0.01, 100.0, size=D.shape[0])
    a = O.fit_M_stable(D, F)
    b = O.fit_M_stable(D * s[:, None], F)
    assert np.abs(b * s[None, :] - a).max() < 1e-10 * np.abs(a).max()


def test_inverse_and_cholesky_estimators_agree():
    rng = np.random.default_rng(3)
    D = rng.normal(size=(60, 200))
    F = rng.normal(size=(30, 200))
    a, b = O.fit_M(D, F), O.fit_M_stable(D, F)
    assert np.abs(a - b).max() < 1e-9 * np.abs(b).max()


def test_metric_root():
    rng = np.random.default_rng(4)
    B = rng.normal(size=(30, 30))
    V = B @ B.T + 0.1 * np.eye(30)
    R = O.aux_metric_sqrt(V)
    assert np.abs(R @ R - V).max() < 1e-11 * np.abs(V).max()
    assert np.abs(R - R.T).max() == 0.0 or np.abs(R - R.T).max() < 1e-13
    # a rank-deficient metric: the root spans the range only
    C = rng.normal(size=(30, 12))
    V2 = C @ C.T
    R2 = O.aux_metric_sqrt(V2)
    assert np.abs(R2 @ R2 - V2).max() < 1e-10 * np.abs(V2).max()
    assert np.linalg.matrix_rank(R2, tol=1e-8 * np.abs(R2).max()) == 12


def test_unpack_factors():
    X, D = np.ones((3, 2)), np.ones((3, 4))
    assert unpack_factors((X, D))[2:] == (None, None)
    xs = unpack_factors((X, D, 'ao', 'pts'))
    assert xs[2] == 'ao' and xs[3] == 'pts'
    with pytest.raises(ValueError):
        unpack_factors((X,))
