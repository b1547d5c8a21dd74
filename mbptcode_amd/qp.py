"""The quasiparticle equation from Sigma^c on the imaginary axis: Thiele-Pade
continuation and the largest-weight root (GW/qp_solve.py:529-558, :575-711;
Base/utils/grids/analytic_continuation.py:15-89; Solvers/qp_equation.py:
1319-1389), host-side FP64: per state, an O(n^2) greedy node ordering, one
vectorised scan per window and a scalar bisection per bracket (about 1 ms).

    w = eps_p + <Sigma_x - v_xc>_pp + Re Sigma^c_pp(w)

The static term <Sigma_x - v_xc> needs the mean field's exchange and xc
potential (pyscf); here it is the caller's `xc_diagonal`, as the reference's
`finish_qp(xc_diagonal=...)` takes it.

The root search is a reduced form of the reference's `find_qp_root`: the
roots of f(w) = w - a - Re Sigma(w) (a = eps_p + xc_p) are bracketed on a
scan of half-width 0.5, then 1.0, then 2.0 Ha about the bare level eps_p
(the reference centres its windows there too), refined by
bisection, weighed by Z = 1/f'(w) (a central difference), and the root of
largest valid weight (0 < Z <= 1) is kept; a weight above 0.5 ends the
widening, and a scan with no valid weight keeps the root nearest a.  The
satellite bookkeeping, branch following and pole-sum bounds of the
reference are not restated.
"""
import numpy as np

#: root tolerance and scan points (constants.py:272-274), Z's derivative
#: step (:276), the quasiparticle weight threshold (:553) and the scan's
#: half-widths about the bare level, Ha
QP_GRAPHICAL_TOL = 1e-8
QP_GRAPHICAL_N_OMEGA = 150
QP_GRAPHICAL_MAX_BISECTION = 100
QP_Z_DERIV_STEP = 1e-3
QP_WINDOW_Z_MIN = 0.5
QP_WINDOWS = (0.5, 1.0, 2.0)


def thiele_coefficients(z, f):
    """Coefficients a_i of the Thiele continued fraction through (z_i, f_i),
    by the reciprocal-difference table (analytic_continuation.py:15-22)."""
    z = np.asarray(z, complex)
    n = len(z)
    g = np.zeros((n, n), complex)
    g[:, 0] = f
    for i in range(1, n):
        g[i:, i] = (g[i - 1, i - 1] - g[i:, i - 1]) / ((z[i:] - z[i - 1]) * g[i:, i - 1])
    return np.diagonal(g).copy()


def pade_eval(z_query, z, a):
    """The continued fraction a_0 / (1 + (w - z_0) a_1 / (1 + ...)) at w
    (analytic_continuation.py:25-31)."""
    zq = np.asarray(z_query, complex)
    out = a[-1] * np.ones_like(zq)
    for i in range(len(a) - 2, -1, -1):
        out = a[i] / (1.0 + (zq - z[i]) * out)
    return out


def greedy_pade_order(z, f):
    """Node order that keeps each next Thiele coefficient smallest, from the
    point nearest the origin (analytic_continuation.py:34-61).

    A candidate's coefficient if appended next is the next entry of its own
    row of the reciprocal-difference table, which depends only on the nodes
    already chosen; every candidate's row is advanced once per chosen node,
    so the ordering costs O(n^2) rather than a table per trial."""
    z = np.asarray(z, complex)
    f = np.asarray(f, complex)
    n = len(z)
    start = int(np.argmin(np.abs(z)))
    order = [start]
    left = np.ones(n, bool)
    left[start] = False
    row = f.copy()            # each candidate's current table entry g_x[c]
    a_prev, z_prev = f[start], z[start]
    with np.errstate(all='ignore'):
        while left.any():
            nxt = (a_prev - row) / ((z - z_prev) * row)
            score = np.where(left & np.isfinite(nxt), np.abs(nxt), np.inf)
            j = int(np.argmin(score))
            if not np.isfinite(score[j]):
                raise FloatingPointError(
                    f'greedy Pade ordering: every next coefficient of the '
                    f'{int(left.sum())} points left is not finite after {len(order)} nodes')
            order.append(j)
            left[j] = False
            row = nxt
            a_prev, z_prev = nxt[j], z[j]
    return np.array(order)


class PadeFit:
    """The Thiele-Pade fraction through (z_i, f_i), greedily ordered unless
    `greedy` is False, cut to the first `max_order` nodes of that order when
    given (analytic_continuation.py:64-89)."""

    def __init__(self, z, f, greedy=True, max_order=None):
        z = np.asarray(z, complex)
        f = np.asarray(f, complex)
        if greedy:
            order = greedy_pade_order(z, f)
            z, f = z[order], f[order]
        if max_order is not None and max_order < len(z):
            if max_order < 2:
                raise ValueError(f'max_order = {max_order}: a Pade fit needs at least two '
                                 'points')
            z, f = z[:max_order], f[:max_order]
        self.z, self.f = z, f
        self.coefficients = thiele_coefficients(z, f)

    def __call__(self, z_query):
        return pade_eval(z_query, self.z, self.coefficients)


def imaginary_axis_sample_points(freq_points, nocc, p_state, mu):
    """(z, i w): the line Re z = mu on the side of the gap the state sits on
    (qp_solve.py:698-703)."""
    sign = -1.0 if p_state < nocc else 1.0
    iw = sign * np.asarray(freq_points, float)
    return mu + 1j * iw, iw


def pole_strength(f, w, h=QP_Z_DERIV_STEP):
    """Z = 1 / f'(w) by a central difference (qp_equation.py:1336-1346)."""
    d = (f(w + h) - f(w - h)) / (2.0 * h)
    return np.nan if d == 0.0 or not np.isfinite(d) else 1.0 / d


def _bisect(f, lo, hi, flo, tol, steps):
    for _ in range(steps):
        mid = 0.5 * (lo + hi)
        fm = f(mid)
        if flo * fm <= 0.0:
            hi = mid
        else:
            lo, flo = mid, fm
        if hi - lo <= tol:
            break
    return 0.5 * (lo + hi)


def largest_weight_root(f, a, windows=QP_WINDOWS, n_omega=QP_GRAPHICAL_N_OMEGA,
                        tol=QP_GRAPHICAL_TOL, steps=QP_GRAPHICAL_MAX_BISECTION,
                        z_min=QP_WINDOW_Z_MIN, vectorized=False):
    """(w, Z) of f(w) = 0 as described in the module docstring; (nan, nan)
    when no window holds a root.  vectorized: f takes an array of points (the
    scan is then one call)."""
    best = None
    nearest = None
    for half in windows:
        grid = np.linspace(a - half, a + half, n_omega)
        vals = (np.asarray(f(grid), float) if vectorized
                else np.array([f(w) for w in grid]))
        for i in range(n_omega - 1):
            if not (np.isfinite(vals[i]) and np.isfinite(vals[i + 1])):
                continue
            if vals[i] == 0.0 or vals[i] * vals[i + 1] < 0.0:
                w = grid[i] if vals[i] == 0.0 else _bisect(f, grid[i], grid[i + 1], vals[i],
                                                           tol, steps)
                z = pole_strength(f, w)
                if nearest is None or abs(w - a) < abs(nearest[0] - a):
                    nearest = (w, z)
                if np.isfinite(z) and 0.0 < z <= 1.0 and (best is None or z > best[1]):
                    best = (w, z)
        if best is not None and best[1] > z_min:
            return best
    if best is not None:
        return best
    return nearest if nearest is not None else (np.nan, np.nan)


def solve_qp_from_imaginary_axis(eps_p, xc_p, z_fit, sigma_iw, greedy=True, with_z=False):
    """The root of w = eps_p + xc_p + Re Sigma_c(w), Sigma_c Pade-continued
    from (z_fit, sigma_iw) (qp_solve.py:529-558)."""
    fit = PadeFit(z_fit, sigma_iw, greedy=greedy)
    a = float(eps_p) + float(xc_p)

    def f(w):
        wa = np.asarray(w, float)
        out = wa - a - fit(wa.astype(complex)).real
        return out if out.ndim else float(out)
    w, z = largest_weight_root(f, float(eps_p), vectorized=True)
    return (w, z) if with_z else w


def quasiparticle_energies(sigma, eps, nocc, states, mu, pade_freq, xc_diagonal,
                           greedy=True, with_z=False):
    """Quasiparticle energies of `states` from Sigma^c_pp(i w) (nstates,
    npade), sampled on the positive imaginary axis at `pade_freq` above mu
    (an occupied state's values are conjugated onto its own branch,
    Sigma(-i w) = Sigma(i w)^*), and the caller's <Sigma_x - v_xc>_pp
    (qp_solve.py:575-696 with xc_diagonal given)."""
    eps = np.asarray(eps, float)
    states = np.atleast_1d(states)
    sig = np.atleast_2d(np.asarray(sigma, complex))
    xc = np.atleast_1d(np.asarray(xc_diagonal, float))
    if sig.shape[0] != len(states) or xc.shape[0] != len(states):
        raise ValueError('sigma, states and xc_diagonal differ in length')
    roots, zs = [], []
    for i, p in enumerate(states):
        z_fit, _ = imaginary_axis_sample_points(pade_freq, nocc, int(p), mu)
        s_p = np.conj(sig[i]) if p < nocc else sig[i]
        w, z = solve_qp_from_imaginary_axis(eps[p], xc[i], z_fit, s_p, greedy, with_z=True)
        roots.append(w)
        zs.append(z)
    roots, zs = np.array(roots), np.array(zs)
    return (roots, zs) if with_z else roots
