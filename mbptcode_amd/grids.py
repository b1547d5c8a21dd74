"""Imaginary-axis grids of the space-time chi0 path, host-side (NumPy FP64).

* minimax imaginary-time grids, a lookup in the GreenX tables
  (src/Base/utils/grids/grids.py:78-158);
* the tau -> omega cosine transform, a per-frequency regularised least-squares
  fit (src/Base/utils/grids/time_frequency.py:282-380);
* the split grid object (time_frequency.py:498-555) with the 1/(2 e_min)
  time scaling of the time-frequency pair;
* the Laplace test and the automatic point count (time_frequency.py:151-215);
* Gauss-Legendre frequencies on [0, inf) and the gap-scaled w0
  (grids.py:161-185).

These are set-up computations of a few microseconds to milliseconds; they run
on the host in FP64 and hand plain arrays (tau points, transform weights) to
the device kernels.
"""
import json
import os
import warnings

import numpy as np

from .constants import DEFAULT_TAU_TARGET, TRANSFORM_FIT_RCOND
from .orbital_spaces import particle_hole_window

_TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data',
                           'minimax_tau.json')
_tables = None

#: Fits at or above this many input points get a Tikhonov term
#: (time_frequency.py:109-110).
REGULARIZATION_ABOVE = 20
DEFAULT_REGULARIZATION = 1e-8
#: A transform fitting worse than this warns (time_frequency.py:115).
FIT_ERROR_WARN = 1e-3


def _load_tables():
    global _tables
    if _tables is None:
        with open(_TABLE_PATH) as f:
            _tables = json.load(f)['tables']
    return _tables


def minimax_tau_supported_sizes():
    """Tabulated time-grid sizes (even, 6..34)."""
    return sorted(int(k) for k in _load_tables())


def _require_window(e_min, e_max, who):
    if not e_min > 0:
        raise ValueError(f'{who}: e_min = {e_min:.6g} is not positive; a T=0 '
                         'minimax grid is scaled by the gap and needs one.')
    if e_max is not None and not e_max > e_min:
        raise ValueError(f'{who}: e_max = {e_max:.6g} must exceed '
                         f'e_min = {e_min:.6g}')


def table_row(ntau, e_min, e_max):
    """(row, e_ratio): the coefficient row for this size and range, and the
    narrow-range stretch factor (grids.py:78-104).

    The row is the first tabulated range strictly wider than e_max/e_min (the
    last row serves everything wider).  Below the narrowest range, grids of more
    than 20 points are stretched by e_ratio = R_0 / R (divided by 1.5 when that
    exceeds 1.5); time points then MULTIPLY by it.
    """
    entry = _load_tables().get(str(int(ntau)))
    if entry is None:
        raise ValueError(f'ntau={ntau} has no tabulated minimax time grid; '
                         f'supported: {minimax_tau_supported_sizes()}')
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


def cosine_transform_weights(tau, omega, e_min, e_max, regularization=None,
                             nodes_factor=200, warn=True):
    """(C, max_fit_error): the tau -> omega cosine weights, C[k, j] including the
    cos(omega_k tau_j) factor, fitted row by row so that
    sum_j C[k, j] e^{-x tau_j} = 2x / (x^2 + omega_k^2) on a log-spaced node set
    over [e_min, e_max] (time_frequency.py:282-380, kind COSINE_TW).

    Node count: max((floor(log10 R) + 1) * nodes_factor, ntau).  Each row is a
    filtered SVD solve, filter S / (reg^2 + S^2) on S > rcond * S_max, with
    reg = 0 below 20 input points and 1e-8 at or above.
    """
    tau = np.asarray(tau, float)
    omega = np.asarray(omega, float)
    n = len(tau)
    if regularization is None:
        regularization = 0.0 if n < REGULARIZATION_ABOVE else DEFAULT_REGULARIZATION
    nx = max((int(np.log10(e_max / e_min)) + 1) * nodes_factor, n)
    x = e_min * (e_max / e_min) ** (np.arange(nx) / (nx - 1.0))
    decay = np.exp(-np.outer(x, tau))                    # (nx, ntau)
    C = np.zeros((len(omega), n))
    worst = 0.0
    for k, w in enumerate(omega):
        psi = 2 * x / (x ** 2 + w ** 2)
        A = np.cos(w * tau)[None, :] * decay
        U, S, Vt = np.linalg.svd(A, full_matrices=False)
        filt = np.zeros_like(S)
        np.divide(S, regularization ** 2 + S ** 2, out=filt,
                  where=S > TRANSFORM_FIT_RCOND * S[0])
        C[k] = Vt.T @ (filt * (U.T @ psi))
        err = float(np.abs(A @ C[k] - psi).max())
        if not (np.isfinite(err) and np.isfinite(C[k]).all()):
            raise FloatingPointError(f'cosine transform row {k} is not finite')
        worst = max(worst, err)
    if warn and not worst <= FIT_ERROR_WARN:
        warnings.warn(f'minimax cosine transform fit reached only {worst:.2e} '
                      f'at {n} points over e_max/e_min = {e_max / e_min:.3g}',
                      RuntimeWarning, stacklevel=2)
    return C * np.cos(np.outer(omega, tau)), worst


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


def rpa_energy_grid(eps, nocc, ntau, nfreq=40, w0_scale=0.5):
    """The dRPA energy's grid (gradients/rpa/ground_state.py:92-106): minimax
    time points on [gap, span], Gauss-Legendre frequencies with w0 = 0.5 gap."""
    gap, span, _ = particle_hole_window(eps, nocc)
    nu, wt = gauss_legendre_grid(nfreq, gap_scaled_w0(np.asarray(eps, float),
                                                      nocc, w0_scale))
    return TimeFrequencyGrid.minimax_split(int(ntau), gap, span, nu, wt)
