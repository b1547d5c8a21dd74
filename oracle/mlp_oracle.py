"""TEST INFRASTRUCTURE ONLY -- FP64 NumPy oracle of the learned Davidson
preconditioner (the MLP has no counterpart in the reference: it replaces the
diagonal correction r / (d - theta) of davidson_shared.py:157-163 and
Solvers/davidson.py:43-77).

Network: 3 -> H -> H -> 1, SiLU activations, mean-squared-error loss.

    z1 = x W1 + b1,  h1 = silu(z1)
    z2 = h1 W2 + b2, h2 = silu(z2)
    y  = h2 . w3 + b3
    L  = mean_n (y_n - t_n)^2

Pinning: `backward` is checked against central finite differences of
`loss` in tests/test_mlp_oracle.py (a closed-form check of every gradient
component), so the oracle is pinned without reference to the device code.
"""
import numpy as np


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def init_params(hidden=256, seed=0):
    """He-style initialisation, FP64: dict W1 (3, H), b1 (H,), W2 (H, H),
    b2 (H,), w3 (H,), b3 ()."""
    rng = np.random.default_rng(seed)
    return {'W1': rng.normal(0, np.sqrt(2.0 / 3), (3, hidden)),
            'b1': np.zeros(hidden),
            'W2': rng.normal(0, np.sqrt(2.0 / hidden), (hidden, hidden)),
            'b2': np.zeros(hidden),
            'w3': rng.normal(0, np.sqrt(1.0 / hidden), hidden),
            'b3': np.zeros(())}


def forward(p, x):
    z1 = x @ p['W1'] + p['b1']
    h1 = silu(z1)
    z2 = h1 @ p['W2'] + p['b2']
    h2 = silu(z2)
    return h2 @ p['w3'] + p['b3'], (z1, h1, z2, h2)


def loss(p, x, t):
    y, _ = forward(p, x)
    return float(np.mean((y - t) ** 2))


def backward(p, x, t):
    """(loss, grads) with grads keyed like the parameters."""
    y, (z1, h1, z2, h2) = forward(p, x)
    n = len(x)
    g = 2.0 * (y - t) / n                       # dL/dy
    grads = {'w3': h2.T @ g, 'b3': np.array(g.sum())}
    dz2 = np.outer(g, p['w3']) * silu_grad(z2)
    grads['W2'] = h1.T @ dz2
    grads['b2'] = dz2.sum(0)
    dz1 = (dz2 @ p['W2'].T) * silu_grad(z1)
    grads['W1'] = x.T @ dz1
    grads['b1'] = dz1.sum(0)
    return float(np.mean((y - t) ** 2)), grads


def features(r, d, theta):
    """Per-pair inputs for a block of residuals r (n, k): columns
    [r_ia / rms_i(r), d_ia - theta, theta] flattened root-major, and the
    per-root scale rms(r) the output is multiplied back by."""
    rms = np.sqrt(np.mean(r ** 2, axis=0)) + 1e-300
    n, k = r.shape
    f = np.empty((k, n, 3))
    f[:, :, 0] = (r / rms).T
    f[:, :, 1] = d[None, :] - theta[:, None]
    f[:, :, 2] = theta[:, None]
    return f.reshape(k * n, 3), rms


def correction(p, r, d, theta):
    """The learned correction t (n, k) = rms(r) * net(features)."""
    f, rms = features(r, d, theta)
    y, _ = forward(p, f)
    return y.reshape(r.shape[1], r.shape[0]).T * rms[None, :]
