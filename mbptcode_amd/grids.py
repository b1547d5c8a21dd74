"""Imaginary-axis grids of the space-time chi0 path, host-side (NumPy FP64).

* minimax imaginary-time grids, a lookup in the GreenX tables
  (src/Base/utils/grids/grids.py:78-158);
* the tau -> omega cosine transform, a per-frequency regularised least-squares
  fit (src/Base/utils/grids/time_frequency.py:282-380);
* the split grid object (time_frequency.py:498-555) with the 1/(2 e_min)
  time scaling of the time-frequency pair;
* the Laplace test and the automatic point count (time_frequency.py:151-215);
* Gauss-Legendre frequencies on [0, inf) and the gap-scaled w0
  (grids.py:161-185);
* minimax imaginary-frequency grids (grids.py:107-126), the four cosine /
  sine transforms (time_frequency.py:257-380), the self-energy's fit ranges
  (:387-427) and the space-time GW axes (GW/space_time.py:559-618,
  GW/imaginary_time.py:95-138).

These are set-up computations of a few microseconds to milliseconds; they run
on the host in FP64 and hand plain arrays (tau points, transform weights) to
the device kernels.
"""
import json
import os
import warnings

import numpy as np

from .constants import (DEFAULT_NPADE, DEFAULT_TAU_TARGET, GL_FREQUENCY_SCALE,  # noqa: F401
                        TRANSFORM_FIT_RCOND)
from .orbital_spaces import particle_hole_window

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data')
_TABLE_FILES = {'tau': 'minimax_tau.json', 'omega': 'minimax_omega.json'}
_tables = {}

#: GreenX transformation-type codes (minimax_utils.F90, time_frequency.py:95):
#: cosine tau -> omega, cosine omega -> tau, sine tau -> omega, sine omega -> tau
COSINE_TW, COSINE_WT, SINE_TW, SINE_WT = 1, 2, 3, 4
#: the self-energy fit ranges' padding (time_frequency.py:230)
SELF_ENERGY_PAD = (0.3, 3.0)

#: Fits at or above this many input points get a Tikhonov term
#: (time_frequency.py:109-110).
REGULARIZATION_ABOVE = 20
DEFAULT_REGULARIZATION = 1e-8
#: A transform fitting worse than this warns (time_frequency.py:115).
FIT_ERROR_WARN = 1e-3


def _load_tables(axis='tau'):
    if axis not in _tables:
        with open(os.path.join(_DATA, _TABLE_FILES[axis])) as f:
            _tables[axis] = json.load(f)['tables']
    return _tables[axis]


def minimax_tau_supported_sizes():
    """Tabulated time-grid sizes (even, 6..34)."""
    return sorted(int(k) for k in _load_tables())


def minimax_supported_sizes():
    """Tabulated frequency-grid sizes (grids.py:42-44)."""
    return sorted(int(k) for k in _load_tables('omega'))


def _require_window(e_min, e_max, who):
    if not e_min > 0:
        raise ValueError(f'{who}: e_min = {e_min:.6g} is not positive; a T=0 '
                         'minimax grid is scaled by the gap and needs one.')
    if e_max is not None and not e_max > e_min:
        raise ValueError(f'{who}: e_max = {e_max:.6g} must exceed '
                         f'e_min = {e_min:.6g}')


def table_row(ntau, e_min, e_max, axis='tau'):
    """(row, e_ratio): the coefficient row for this size and range, and the
    narrow-range stretch factor (grids.py:78-104).

    The row is the first tabulated range strictly wider than e_max/e_min (the
    last row serves everything wider).  Below the narrowest range, grids of more
    than 20 points are stretched by e_ratio = R_0 / R (divided by 1.5 when that
    exceeds 1.5); time points then MULTIPLY by it.
    """
    entry = _load_tables(axis).get(str(int(ntau)))
    if entry is None:
        sizes = sorted(int(k) for k in _load_tables(axis))
        raise ValueError(f'{ntau} points have no tabulated minimax {axis} grid; '
                         f'supported: {sizes}')
    ranges = entry['ranges']
    ratio = e_max / e_min
    pick = len(ranges)
    for j, r in enumerate(ranges):
        if ratio < r:
            pick = j
            break
    stretch = 1.0
    if pick == 0 and ntau > 20:
        stretch = ranges[0] / ratio
        if stretch > 1.5:
            stretch /= 1.5
    return np.asarray(entry['rows'][pick], float), stretch


def minimax_frequency_grid(nfreq, e_min, e_max):
    """(omega, w): the minimax imaginary-frequency grid on [e_min, e_max]
    (grids.py:107-126): table row / e_ratio * e_min (a frequency carries
    energy, so the narrow-range stretch divides)."""
    _require_window(e_min, e_max, 'minimax_frequency_grid')
    row, stretch = table_row(nfreq, e_min, e_max, axis='omega')
    return row[:nfreq] / stretch * e_min, row[nfreq:] / stretch * e_min


def minimax_time_grid(ntau, e_min, e_max):
    """(tau, w) with 1/x ~ sum_k w_k exp(-x tau_k) on [e_min, e_max]
    (grids.py:129-158): table row times e_ratio / e_min."""
    _require_window(e_min, e_max, 'minimax_time_grid')
    row, stretch = table_row(ntau, e_min, e_max)
    return row[:ntau] * stretch / e_min, row[ntau:] * stretch / e_min


def laplace_test_error(ntau, R, nsample=4001, relative=False):
    """max over x in [1, R] of |1/x - sum_k w_k e^{-x tau_k}| (or the relative
    form) on a log-spaced sample (time_frequency.py:151-182)."""
    tau, w = minimax_time_grid(ntau, 1.0, float(R))
    x = np.exp(np.linspace(0.0, np.log(float(R)), nsample))
    approx = np.exp(-np.outer(x, tau)) @ w
    if relative:
        return float(np.abs(1.0 - x * approx).max())
    return float(np.abs(1.0 / x - approx).max())


def minimax_points_for_accuracy(e_min, e_max, target, npoints_max=34,
                                relative=False):
    """(ntau, error): the smallest tabulated size meeting `target`, else the
    best one (time_frequency.py:185-215)."""
    _require_window(float(e_min), float(e_max), 'minimax_points_for_accuracy')
    R = float(e_max) / float(e_min)
    sizes = set(minimax_tau_supported_sizes())
    best = (None, np.inf)
    for n in range(6, npoints_max + 1, 2):
        if n not in sizes:
            continue
        err = laplace_test_error(n, R, relative=relative)
        if err < best[1]:
            best = (n, err)
        if err <= target:
            return n, err
    if best[0] is None:
        raise ValueError(f'no tabulated minimax grid fits the range ratio {R:.6g}')
    return best


def gauss_legendre_grid(nfreq, w0):
    """Modified Gauss-Legendre frequencies on [0, inf): w = w0 (1+x)/(1-x),
    weight w0 * 2 w_GL / (1-x)^2 (grids.py:161-170)."""
    _require_window(w0, None, 'gauss_legendre_grid (w0)')
    x, w = np.polynomial.legendre.leggauss(nfreq)
    return w0 * (1.0 + x) / (1.0 - x), w0 * 2.0 * w / (1.0 - x) ** 2


def gap_scaled_w0(eps, nocc, scale=0.5):
    """scale * (eps_LUMO - eps_HOMO) (grids.py:173-185)."""
    w0 = scale * (eps[nocc] - eps[nocc - 1])
    _require_window(w0, None, 'gap_scaled_w0')
    return w0


def _fit_target(kind, tau, omega, i, x):
    """(psi, A): the fitted function on the node set x and the design matrix
    for output point i of transform `kind` (time_frequency.py:257-279):

    * COSINE_TW: sum_j c_j cos(w tau_j) e^{-x tau_j} = 2x / (x^2 + w^2);
    * COSINE_WT: sum_k c_k cos(t w_k) 2x / (x^2 + w_k^2) = e^{-x t};
    * SINE_TW: sum_j c_j sin(w tau_j) e^{-x tau_j} = 2w / (x^2 + w^2);
    * SINE_WT: sum_k c_k sin(t w_k) 2w_k / (x^2 + w_k^2) = e^{-x t}."""
    if kind in (COSINE_TW, SINE_TW):
        w = omega[i]
        trig = np.cos if kind == COSINE_TW else np.sin
        psi = 2 * (x if kind == COSINE_TW else w) / (x ** 2 + w ** 2)
        return psi, trig(w * tau)[None, :] * np.exp(-np.outer(x, tau))
    if kind in (COSINE_WT, SINE_WT):
        t = tau[i]
        lor = 1.0 / (x[:, None] ** 2 + omega[None, :] ** 2)
        if kind == COSINE_WT:
            return np.exp(-x * t), np.cos(t * omega)[None, :] * 2 * x[:, None] * lor
        return np.exp(-x * t), np.sin(t * omega)[None, :] * 2 * omega[None, :] * lor
    raise ValueError(f'unknown transform kind {kind}')


def transform_weights(kind, tau, omega, e_min, e_max, regularization=None,
                      nodes_factor=200, warn=True):
    """(T, max_fit_error): one transform matrix, fitted row by row on a
    log-spaced node set over [e_min, e_max] (time_frequency.py:282-380).
    Rows are the output axis (omega for *_TW, tau for *_WT); the cos / sin
    factor is included.

    Node count: max((floor(log10 R) + 1) * nodes_factor, n_in).  Each row is a
    filtered SVD solve, filter S / (reg^2 + S^2) on S > rcond * S_max, with
    reg = 0 below 20 input points and 1e-8 at or above."""
    tau = np.asarray(tau, float)
    omega = np.asarray(omega, float)
    to_omega = kind in (COSINE_TW, SINE_TW)
    n_out = len(omega) if to_omega else len(tau)
    n = len(tau) if to_omega else len(omega)
    if regularization is None:
        regularization = 0.0 if n < REGULARIZATION_ABOVE else DEFAULT_REGULARIZATION
    nx = max((int(np.log10(e_max / e_min)) + 1) * nodes_factor, n)
    x = e_min * (e_max / e_min) ** (np.arange(nx) / (nx - 1.0))
    T = np.zeros((n_out, n))
    worst = 0.0
    for i in range(n_out):
        psi, A = _fit_target(kind, tau, omega, i, x)
        U, S, Vt = np.linalg.svd(A, full_matrices=False)
        filt = np.zeros_like(S)
        np.divide(S, regularization ** 2 + S ** 2, out=filt,
                  where=S > TRANSFORM_FIT_RCOND * S[0])
        T[i] = Vt.T @ (filt * (U.T @ psi))
        err = float(np.abs(A @ T[i] - psi).max())
        if not (np.isfinite(err) and np.isfinite(T[i]).all()):
            raise FloatingPointError(f'transform kind {kind}: row {i} is not finite')
        worst = max(worst, err)
    if warn and not worst <= FIT_ERROR_WARN:
        warnings.warn(f'minimax transform (kind {kind}) fit reached only {worst:.2e} '
                      f'at {n} points over e_max/e_min = {e_max / e_min:.3g}',
                      RuntimeWarning, stacklevel=2)
    trig = np.cos if kind in (COSINE_TW, COSINE_WT) else np.sin
    phase = trig(np.outer(omega, tau))
    return (T * phase if to_omega else T * phase.T), worst


def cosine_transform_weights(tau, omega, e_min, e_max, regularization=None,
                             nodes_factor=200, warn=True):
    """(C, max_fit_error): the tau -> omega cosine weights (kind COSINE_TW),
    C[k, j] including the cos(omega_k tau_j) factor, so that
    sum_j C[k, j] e^{-x tau_j} ~ 2x / (x^2 + omega_k^2) on [e_min, e_max]."""
    return transform_weights(COSINE_TW, tau, omega, e_min, e_max, regularization,
                             nodes_factor, warn)


def self_energy_fit_ranges_from_window(eps, mu, w_lo, w_hi, pad=SELF_ENERGY_PAD):
    """((rW_lo, rW_hi), (rS_lo, rS_hi)) (time_frequency.py:387-398): the
    screening window [w_lo, w_hi] padded, and the self-energy's decay rates
    |eps - mu| + Omega, sums of both factors' rates, padded alike."""
    lo, hi = pad
    dG = np.abs(np.asarray(eps, float) - mu)
    return ((lo * w_lo, hi * w_hi),
            (lo * (float(dG.min()) + w_lo), hi * (float(dG.max()) + w_hi)))


def sigma_fit_ranges(eps, nocc, mu=None):
    """(rW, rS): the ranges the W(i omega) -> tau and Sigma(tau) -> i omega
    transforms are fitted on (time_frequency.py:401-427), mu at mid-gap by
    default."""
    eps = np.asarray(eps, float)
    occ, virt = eps[:nocc], eps[nocc:]
    if mu is None:
        mu = 0.5 * (occ.max() + virt.min())
    return self_energy_fit_ranges_from_window(eps, mu, virt.min() - occ.max(),
                                              virt.max() - occ.min())


def minimax_points_for_ranges(ratios, target=DEFAULT_TAU_TARGET, npoints_max=34):
    """(n, worst error): the smallest tabulated size resolving every range
    ratio; the widest binds (time_frequency.py:233-254)."""
    npoints, worst = 0, 0.0
    for R in ratios:
        n, err = minimax_points_for_accuracy(1.0, R, target=target,
                                             npoints_max=npoints_max)
        npoints, worst = max(npoints, n), max(worst, err)
    return npoints, worst


def minimax_points_for_gw(eps, nocc, mu=None, target=DEFAULT_TAU_TARGET, npoints_max=34):
    """(ntau, worst error) over the three ranges of the space-time route:
    chi0's [gap, span], rW and rS (GW/imaginary_time.py:108-138)."""
    gap, span, mid = particle_hole_window(eps, nocc)
    rW, rS = sigma_fit_ranges(eps, nocc, mid if mu is None else mu)
    return minimax_points_for_ranges((span / gap, rW[1] / rW[0], rS[1] / rS[0]),
                                     target=target, npoints_max=npoints_max)


def screening_frequency_grid(ntau, eps, nocc, mu=None):
    """(omega, w): the minimax frequencies W(i omega) is sampled on, over rW
    (GW/imaginary_time.py:95-105)."""
    rW, _ = sigma_fit_ranges(eps, nocc, mu=mu)
    return minimax_frequency_grid(ntau, *rW)


class TimeFrequencyGrid:
    """A minimax imaginary-time axis and the cosine transform onto a caller's
    frequency axis (time_frequency.py:398-555, the `minimax_split` case).

    tau_points, tau_weights : (ntau,)
    omega_points, omega_weights : (nfreq,)
    cosft_wt : (nfreq, ntau) even-in-tau -> omega weights
    """

    def __init__(self, tau_points, tau_weights, omega_points, omega_weights,
                 cosft_wt, fit_error, meta):
        self.tau_points = np.asarray(tau_points, float)
        self.tau_weights = np.asarray(tau_weights, float)
        self.omega_points = np.asarray(omega_points, float)
        self.omega_weights = np.asarray(omega_weights, float)
        self.cosft_wt = np.asarray(cosft_wt, float)
        self.fit_errors = {'cosft_wt': fit_error}
        self.meta = meta

    @property
    def ntau(self):
        return len(self.tau_points)

    @property
    def nfreq(self):
        return len(self.omega_points)

    @classmethod
    def minimax_split(cls, ntau, e_min, e_max, omega_points, omega_weights,
                      regularization=None):
        """Minimax time axis of `ntau` points on [e_min, e_max], halved (the
        1/(2 e_min) scaling of the time-frequency pair, time_frequency.py:
        524-536), and its cosine transform onto `omega_points`."""
        tau, w = minimax_time_grid(ntau, e_min, e_max)
        tau, w = 0.5 * tau, 0.5 * w
        C, err = cosine_transform_weights(tau, omega_points, e_min, e_max,
                                          regularization=regularization)
        return cls(tau, w, omega_points, omega_weights, C, err,
                   {'e_min': e_min, 'e_max': e_max, 'ntau': ntau,
                    'nfreq': len(np.atleast_1d(omega_points))})


def static_screening_grid(eps, nocc, ntau=None, tau_target=DEFAULT_TAU_TARGET):
    """The one-frequency (omega = 0) grid the static W is built on
    (LR/space_time.py:984-1000); ntau None or 'auto' resolves the count by the
    Laplace test (space_time.py:971-981)."""
    gap, span, _ = particle_hole_window(eps, nocc)
    if ntau is None or (isinstance(ntau, str) and ntau.lower() == 'auto'):
        ntau, _ = minimax_points_for_accuracy(gap, span, target=tau_target)
    return TimeFrequencyGrid.minimax_split(int(ntau), gap, span, [0.0], [1.0])


def space_time_gw_grids(eps, nocc, ntau=None, npade=DEFAULT_NPADE, w0=GL_FREQUENCY_SCALE,
                        tau_target=DEFAULT_TAU_TARGET):
    """The axes of one restricted space-time self-energy (GW/space_time.py:
    559-618) as a dict:

    * ntau: given, or the smallest size `minimax_points_for_gw` accepts;
    * grid: chi0's split grid, its minimax time axis on [gap, span] and the
      cosine transform onto the screening frequencies;
    * freq_points, freq_weights: the screening frequencies (minimax over rW);
    * tau_points: the self-energy's time axis, half the minimax points over rS;
    * pade_freq: `npade` Gauss-Legendre frequencies of scale w0, where the
      self-energy is returned;
    * ranges: (rW, rS); mu: mid-gap."""
    gap, span, mu = particle_hole_window(eps, nocc)
    if ntau is None or (isinstance(ntau, str) and ntau.lower() == 'auto'):
        ntau, _ = minimax_points_for_gw(eps, nocc, mu=mu, target=tau_target)
    ntau = int(ntau)
    freq_points, freq_weights = screening_frequency_grid(ntau, eps, nocc, mu=mu)
    ranges = sigma_fit_ranges(eps, nocc, mu=mu)
    return {'ntau': ntau, 'mu': mu,
            'grid': TimeFrequencyGrid.minimax_split(ntau, gap, span, freq_points,
                                                    freq_weights),
            'freq_points': freq_points, 'freq_weights': freq_weights,
            'tau_points': 0.5 * minimax_time_grid(ntau, *ranges[1])[0],
            'pade_freq': gauss_legendre_grid(int(npade), w0)[0],
            'ranges': ranges}


def rpa_energy_grid(eps, nocc, ntau, nfreq=40, w0_scale=0.5):
    """The dRPA energy's grid (gradients/rpa/ground_state.py:92-106): minimax
    time points on [gap, span], Gauss-Legendre frequencies with w0 = 0.5 gap."""
    gap, span, _ = particle_hole_window(eps, nocc)
    nu, wt = gauss_legendre_grid(nfreq, gap_scaled_w0(np.asarray(eps, float),
                                                      nocc, w0_scale))
    return TimeFrequencyGrid.minimax_split(int(ntau), gap, span, nu, wt)
