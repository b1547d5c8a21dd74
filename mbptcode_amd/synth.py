"""Synthetic ISDF instances of the five benchmark workloads.

Conventions (fixed here, since the workload descriptions fix seeds but not the
generator calls):

* ``numpy.random.default_rng(seed)``, draws in the order eps_occ, eps_vir, X,
  points;
* eps_occ ~ U[-1.0, -0.3] Ha and eps_vir ~ U[0.05, 3.0] Ha, each sorted
  ascending;
* collocation X[mu, p] ~ N(0, 1/M), shape (M, n_occ + n_vir);
* interpolation points r_mu ~ U[0, 10]^3 Bohr;
* V_mu,nu = erf(r/0.5)/r with the r -> 0 limit 2/(0.5 sqrt(pi)) on the
  diagonal (a Gaussian-smeared Coulomb kernel, positive definite);
* the batched workload draws its sizes from ``default_rng(1)``:
  n_occ = integers(5, 41, 256), n_vir = integers(20, 201, 256), M = 8 (n_occ +
  n_vir); instance i then draws from ``default_rng([1, i])``.

Everything is FP64 NumPy on the host.
"""
from dataclasses import dataclass, field

import numpy as np
from scipy.special import erf

SMEAR = 0.5
BOX = 10.0


@dataclass
class Instance:
    """One synthetic ISDF BSE/RPA problem."""
    nocc: int
    nvir: int
    M: int
    eps: np.ndarray          # (nocc + nvir,) ascending within each block
    X: np.ndarray            # (M, nocc + nvir) collocation
    points: np.ndarray       # (M, 3) Bohr
    V: np.ndarray            # (M, M) smeared Coulomb on the points
    nroots: int = 5
    ntau: int = 8
    tol: float = 1e-5
    name: str = ''
    extra: dict = field(default_factory=dict)

    @property
    def npair(self):
        return self.nocc * self.nvir

    @property
    def X_o(self):
        return np.ascontiguousarray(self.X[:, :self.nocc])

    @property
    def X_v(self):
        return np.ascontiguousarray(self.X[:, self.nocc:])


def coulomb_kernel(points, smear=SMEAR):
    """erf(r/smear)/r between every pair of points, 2/(smear sqrt(pi)) at r=0."""
    diff = points[:, None, :] - points[None, :, :]
    r = np.sqrt((diff ** 2).sum(-1))
    V = np.empty_like(r)
    off = r > 0
    V[off] = erf(r[off] / smear) / r[off]
    V[~off] = 2.0 / (smear * np.sqrt(np.pi))
    return V


def make_instance(nocc, nvir, M, seed, nroots=5, ntau=8, tol=1e-5, name=''):
    """An instance drawn from `seed` (an int or a sequence for default_rng)."""
    rng = np.random.default_rng(seed)
    e_o = np.sort(rng.uniform(-1.0, -0.3, nocc))
    e_v = np.sort(rng.uniform(0.05, 3.0, nvir))
    X = rng.normal(0.0, 1.0 / np.sqrt(M), (M, nocc + nvir))
    pts = rng.uniform(0.0, BOX, (M, 3))
    return Instance(nocc, nvir, M, np.concatenate([e_o, e_v]), X, pts,
                    coulomb_kernel(pts), nroots=nroots, ntau=ntau, tol=tol,
                    name=name)


#: (nocc, nvir, M, nroots, ntau, tol) of the single-instance workloads.
CONFIGS = {
    0: dict(nocc=5, nvir=19, M=120, nroots=5, ntau=8, tol=1e-6,
            name='water/cc-pVDZ size'),
    1: dict(nocc=21, nvir=93, M=1200, nroots=10, ntau=16, tol=1e-5,
            name='benzene/def2-SVP size'),
    2: dict(nocc=60, nvir=400, M=4000, nroots=16, ntau=20, tol=1e-5,
            name='large single instance'),
    4: dict(nocc=100, nvir=700, M=6000, nroots=20, ntau=20, tol=1e-5,
            name='learned-preconditioner instance'),
}


def inversion_instance(nocc, nvir, M_half, seed, nroots=5, ntau=8, tol=1e-5):
    """An instance with an exact inversion symmetry, for the pair-symmetry
    sectors: the points come in pairs r, -r (about the box centre), each
    orbital is even or odd (X[-r] = +-X[r]), and V, hence W, is invariant
    under the swap.  A pair (i, a) is even when i and a have the same
    parity, and A and B couple only pairs of the same parity.
    `extra['orbsym']` holds D2h-style labels (0 even, 1 odd) per orbital."""
    rng = np.random.default_rng(seed)
    e_o = np.sort(rng.uniform(-1.0, -0.3, nocc))
    e_v = np.sort(rng.uniform(0.05, 3.0, nvir))
    M = 2 * M_half
    Xh = rng.normal(0.0, 1.0 / np.sqrt(M), (M_half, nocc + nvir))
    parity = rng.integers(0, 2, nocc + nvir)
    sign = 1.0 - 2.0 * parity
    half = rng.uniform(-0.5 * BOX, 0.5 * BOX, (M_half, 3))
    pts = np.vstack([half, -half]) + 0.5 * BOX
    X = np.vstack([Xh, Xh * sign[None, :]])
    return Instance(nocc, nvir, M, np.concatenate([e_o, e_v]), X, pts,
                    coulomb_kernel(pts), nroots=nroots, ntau=ntau, tol=tol,
                    name=f'inversion[{seed}]', extra={'orbsym': parity})


def config(cfg, seed=0):
    """The single-instance workload `cfg` (0, 1, 2 or 4) from `seed`."""
    c = dict(CONFIGS[cfg])
    return make_instance(c.pop('nocc'), c.pop('nvir'), c.pop('M'), seed, **c)


def batched_sizes(count=256, seed=1):
    """(nocc, nvir, M) of the batched workload's instances."""
    rng = np.random.default_rng(seed)
    no = rng.integers(5, 41, count)
    nv = rng.integers(20, 201, count)
    return [(int(a), int(b), int(8 * (a + b))) for a, b in zip(no, nv)]


def batched_instance(i, sizes=None, seed=1):
    """Instance i of the batched workload: 5 roots, tol 1e-5, 8 time points."""
    sizes = batched_sizes(seed=seed) if sizes is None else sizes
    no, nv, M = sizes[i]
    return make_instance(no, nv, M, [seed, i], nroots=5, ntau=8, tol=1e-5,
                         name=f'batched[{i}]')
