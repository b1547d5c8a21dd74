"""The MLP oracle's backward pass against central finite differences of its
own loss (every parameter block), and the feature/correction conventions.
CPU only."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


@pytest.fixture(scope='module')
def small():
    rng = np.random.default_rng(3)
    p = M.init_params(hidden=8, seed=1)
    p['b1'] = rng.normal(size=8) * 0.1
    p['b2'] = rng.normal(size=8) * 0.1
    p['b3'] = np.array(0.3)
    x = rng.normal(size=(13, 3))
    t = rng.normal(size=13)
    return p, x, t


@pytest.mark.parametrize('name', ['W1', 'b1', 'W2', 'b2', 'w3', 'b3'])
def test_backward_matches_finite_differences(small, name):
    p, x, t = small
    _, grads = M.backward(p, x, t)
    g = grads[name]
    flat = np.asarray(p[name], float).reshape(-1)
    h = 1e-6
    fd = np.empty(flat.size)
    for i in range(flat.size):
        q = {k: np.array(v, float, copy=True) for k, v in p.items()}
        qf = q[name].reshape(-1)
        qf[i] += h
        lp = M.loss(q, x, t)
        qf[i] -= 2 * h
        lm = M.loss(q, x, t)
        fd[i] = (lp - lm) / (2 * h)
    assert np.abs(np.asarray(g).reshape(-1) - fd).max() <= 1e-6 * max(1.0, np.abs(fd).max())


def test_silu_grad_is_the_derivative():
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (M.silu(z + h) - M.silu(z - h)) / (2 * h)
    assert np.abs(M.silu_grad(z) - fd).max() < 1e-8


def test_features_and_scale():
    rng = np.random.default_rng(0)
    r = rng.normal(size=(10, 3))
    d = rng.uniform(0.3, 2.0, 10)
    th = np.array([0.2, 0.5, 0.9])
    f, rms = M.features(r, d, th)
    assert f.shape == (30, 3)
    assert np.allclose(rms, np.sqrt((r ** 2).mean(0)))
    assert np.allclose(f[10:20, 0], r[:, 1] / rms[1])      # root-major rows
    assert np.allclose(f[10:20, 1], d - 0.5) and np.allclose(f[10:20, 2], 0.5)
    p = M.init_params(hidden=8, seed=0)
    t = M.correction(p, r, d, th)
    assert t.shape == r.shape
    # the correction is scale-equivariant in the residual
    assert np.allclose(M.correction(p, 7.0 * r, d, th), 7.0 * t)
