"""The space-time GW self-energy's host side and its oracle, pinned by
mathematics: the minimax frequency grid's quadrature, the four transform fits,
the fit ranges, and the oracle's Sigma^c(i w) against the exact GW@RPA
sum-over-poles self-energy (GW/space_time.py:559-849,
GW/imaginary_time.py:39-195, :536-578, time_frequency.py:257-449)."""
import numpy as np
import pytest

from mbptcode_amd import grids, synth
from oracle import isdf_oracle as O


@pytest.mark.parametrize('n', [10, 16, 20, 26])
def test_minimax_frequency_grid_quadrature(n):
    """The GreenX frequency grids integrate x^2 / (x^2 + w^2)^2 over [0, inf)
    (= pi / (4x)) for x across [e_min, e_max], better as n grows and worse as
    the range widens (bars about 3x the measured errors)."""
    bars = {10: (2e-8, 2e-4, 0.2), 16: (1e-9, 4e-8, 2e-3), 20: (3e-11, 2e-10, 6e-5),
            26: (5e-11, 5e-11, 3e-7)}[n]
    for (e_min, e_max), bar in zip(((0.2, 3.0), (1.0, 100.0), (0.05, 400.0)), bars):
        w, wt = grids.minimax_frequency_grid(n, e_min, e_max)
        assert np.all(w > 0) and np.all(wt > 0) and np.all(np.diff(w) > 0)
        x = np.exp(np.linspace(np.log(e_min), np.log(e_max), 1001))[:, None]
        q = (wt * x ** 3 / (x ** 2 + w ** 2) ** 2).sum(1) * 4 / np.pi
        assert np.abs(q - 1).max() < bar, (n, e_min, e_max, np.abs(q - 1).max())
    with pytest.raises(ValueError):
        grids.minimax_frequency_grid(7, 1.0, 10.0)
    with pytest.raises(ValueError):
        grids.minimax_frequency_grid(10, 0.0, 10.0)


@pytest.mark.parametrize('kind', [grids.COSINE_TW, grids.COSINE_WT, grids.SINE_TW,
                                  grids.SINE_WT])
def test_transform_weights_fit_their_model(kind):
    """Each transform reproduces its model pair on fresh x inside the range:
    cos tau->w: e^{-x tau} -> 2x/(x^2+w^2); sin tau->w: -> 2w/(x^2+w^2);
    the w->tau pair backwards."""
    e_min, e_max = 0.3, 30.0
    tau, _ = grids.minimax_time_grid(16, e_min, e_max)
    w, _ = grids.minimax_frequency_grid(16, e_min, e_max)
    T, err = grids.transform_weights(kind, tau, w, e_min, e_max)
    x = np.exp(np.linspace(np.log(e_min), np.log(e_max), 333))
    if kind in (grids.COSINE_TW, grids.SINE_TW):
        f_in = np.exp(-np.outer(x, tau))                        # (nx, ntau)
        num = x[:, None] if kind == grids.COSINE_TW else w[None, :]
        f_out = 2 * num / (x[:, None] ** 2 + w[None, :] ** 2)
    else:
        num = x[:, None] if kind == grids.COSINE_WT else w[None, :]
        f_in = 2 * num / (x[:, None] ** 2 + w[None, :] ** 2)  # (nx, nfreq)
        f_out = np.exp(-np.outer(x, tau))
    got = f_in @ T.T
    assert np.abs(got - f_out).max() < 5e-3
    assert err < 5e-3


def test_sigma_fit_ranges():
    """rW pads the transition window by (0.3, 3); rS adds |eps - mu|'s span
    to it (a product's decay rates are sums)."""
    inst = synth.config(0)
    eps, nocc = inst.eps, inst.nocc
    rW, rS = grids.sigma_fit_ranges(eps, nocc)
    occ, vir = eps[:nocc], eps[nocc:]
    mu = 0.5 * (occ.max() + vir.min())
    assert np.allclose(rW, (0.3 * (vir.min() - occ.max()), 3.0 * (vir.max() - occ.min())))
    dG = np.abs(eps - mu)
    assert np.isclose(rS[0], 0.3 * (dG.min() + vir.min() - occ.max()))
    assert np.isclose(rS[1], 3.0 * (dG.max() + vir.max() - occ.min()))
    n, err = grids.minimax_points_for_gw(eps, nocc)
    assert n in grids.minimax_tau_supported_sizes() and err <= 1e-10


def gw_inputs(cfg=0):
    inst = synth.config(cfg)
    return inst, O.grid_factors(inst.V)


def oracle_sigma(inst, D, states, ntau=None):
    ax = grids.space_time_gw_grids(inst.eps, inst.nocc, ntau=ntau)
    rW, rS = ax['ranges']
    tau = ax['tau_points']
    Ctw = grids.transform_weights(grids.COSINE_WT, tau, ax['freq_points'], *rW, warn=False)[0]
    C = grids.transform_weights(grids.COSINE_TW, tau, ax['pade_freq'], *rS, warn=False)[0]
    S = grids.transform_weights(grids.SINE_TW, tau, ax['pade_freq'], *rS, warn=False)[0]
    sig = O.gw_sigma_space_time(inst.X, D, inst.eps, inst.nocc, states, ax['grid'],
                                ax['freq_points'], tau, ax['pade_freq'], (Ctw, C, S),
                                ax['mu'])
    return sig, ax


def test_rpa_pole_expansion_of_the_screened_interaction():
    """W - I = chi0 (I - chi0)^-1 = -4 sum_S Omega_S/(Omega_S^2 + w^2) rho_S
    rho_S^T, with chi0 the closed form: pins the poles' normalisation."""
    inst, D = gw_inputs(0)
    om, rho = O.rpa_poles(inst.X, D, inst.eps, inst.nocc)
    for w in (0.0, 0.3, 2.0):
        c = O.chi0_omega_closed_form(inst.X, D, inst.eps, inst.nocc, w)
        Wc = np.linalg.inv(np.eye(len(c)) - c) - np.eye(len(c))
        sos = -4 * (rho * (om / (om ** 2 + w ** 2))) @ rho.T
        assert np.abs(Wc - sos).max() < 1e-12 * max(1.0, np.abs(Wc).max())


def test_space_time_sigma_matches_sum_over_poles():
    """The oracle's space-time chain (chi0 on the screening frequencies,
    Dyson, W - I to tau, the branch sums, cosine + sine back) against the
    exact GW@RPA pole sum at z = mu + i w: to the minimax fits' error, and
    closer on a denser grid."""
    inst, D = gw_inputs(0)
    states = [0, inst.nocc - 2, inst.nocc - 1, inst.nocc, inst.nocc + 1, inst.nocc + 5]
    errs = []
    for ntau in (None, 24):
        sig, ax = oracle_sigma(inst, D, states, ntau)
        sos = O.gw_sigma_sum_over_poles(inst.X, D, inst.eps, inst.nocc, states,
                                        ax['pade_freq'], ax['mu'])
        scale = np.abs(sos).max()
        errs.append(np.abs(sig - sos).max() / scale)
        assert scale > 1e-3 and np.abs(sos.imag).max() > 1e-3
    assert errs[0] < 2e-6 and errs[1] < 1e-7, errs


def test_planted_mistakes_are_caught():
    """Plausible slips in the chain miss the pole sum by far more than the
    fits' error: G at half the time (one operand's exponential only), the
    branch combination's sign swapped, and W taken without its - I."""
    inst, D = gw_inputs(0)
    states = [inst.nocc - 1, inst.nocc]
    ax = grids.space_time_gw_grids(inst.eps, inst.nocc)
    rW, rS = ax['ranges']
    tau = ax['tau_points']
    Ctw = grids.transform_weights(grids.COSINE_WT, tau, ax['freq_points'], *rW, warn=False)[0]
    C = grids.transform_weights(grids.COSINE_TW, tau, ax['pade_freq'], *rS, warn=False)[0]
    S = grids.transform_weights(grids.SINE_TW, tau, ax['pade_freq'], *rS, warn=False)[0]
    args = (inst.X, D, inst.eps, inst.nocc)
    chi = O.chi0_imaginary_frequency(*args, ax['grid'].tau_points, ax['grid'].cosft_wt,
                                     ax['mu'])
    Wt = O.screened_interaction_tau(chi, Ctw)
    sos = O.gw_sigma_sum_over_poles(*args, states, ax['pade_freq'], ax['mu'])
    good = O.self_energy_diagonal(O.self_energy_branch_sums(*args, states, Wt, tau, ax['mu']),
                                  C, S)
    half_g = O.self_energy_diagonal(
        O.self_energy_branch_sums(*args, states, Wt, 0.5 * tau, ax['mu']), C, S)
    sums = O.self_energy_branch_sums(*args, states, Wt, tau, ax['mu'])
    swapped = O.self_energy_diagonal(sums[::-1], C, S)
    eye = np.eye(chi.shape[-1])
    Wfull = np.tensordot(Ctw, np.stack([O.dyson_screening(c) for c in chi]), axes=(1, 0))
    no_minus_i = O.self_energy_diagonal(
        O.self_energy_branch_sums(*args, states, Wfull, tau, ax['mu']), C, S)
    scale = np.abs(sos).max()
    assert np.abs(good - sos).max() / scale < 2e-6
    for bad in (half_g, swapped, no_minus_i):
        assert np.abs(bad - sos).max() / scale > 1e-2
    assert np.allclose(Wfull - Wt, np.tensordot(Ctw, np.stack([eye] * len(chi)), axes=(1, 0)))


def test_branch_sums_against_the_quadratic_form():
    """One branch sum is x_p^T (Zt o G) x_p; the oracle's einsum path against
    an explicit double loop on a tiny grid."""
    rng = np.random.default_rng(3)
    M, no, nv = 7, 2, 3
    X = rng.normal(size=(M, no + nv))
    eps = np.array([-0.9, -0.5, 0.2, 0.6, 1.3])
    D = np.linalg.cholesky(np.eye(M) + 0.1 * np.ones((M, M)))
    Wt = rng.normal(size=(2, M, M))
    Wt = Wt + Wt.transpose(0, 2, 1)
    tau = np.array([0.3, 1.1])
    mu = 0.5 * (eps[no - 1] + eps[no])
    got = O.self_energy_branch_sums(X, D, eps, no, [1, 3], Wt, tau, mu)
    for k, t in enumerate(tau):
        Z = D @ Wt[k] @ D.T
        for j, p in enumerate([1, 3]):
            lt = gt = 0.0
            for P in range(M):
                for Q in range(M):
                    gl = sum(X[P, i] * X[Q, i] * np.exp((eps[i] - mu) * t) for i in range(no))
                    gg = sum(X[P, a] * X[Q, a] * np.exp(-(eps[a] - mu) * t)
                             for a in range(no, no + nv))
                    lt += X[P, p] * Z[P, Q] * gl * X[Q, p]
                    gt -= X[P, p] * Z[P, Q] * gg * X[Q, p]
            assert np.isclose(got[0, k, j], lt, rtol=1e-12)
            assert np.isclose(got[1, k, j], gt, rtol=1e-12)


# ---------------------------------------------------------------------------
# the quasiparticle equation (GW/qp_solve.py:529-711, analytic_continuation.py)
# ---------------------------------------------------------------------------

HA_MEV = 27211.386245988


def exact_qp(inst, D, states, xc):
    """Largest-weight roots of w = eps_p + xc_p + Re Sigma(w) with the exact
    pole-sum Sigma on the real axis."""
    from mbptcode_amd import qp
    out = []
    for i, p in enumerate(states):
        a = inst.eps[p] + xc[i]

        def f(w, p=p, a=a):
            return w - a - O.gw_sigma_sum_over_poles(inst.X, D, inst.eps, inst.nocc, [p],
                                                     None, None, z=[w])[0, 0].real
        out.append(qp.largest_weight_root(f, a))
    return np.array(out)


def test_pade_reproduces_a_rational_function():
    """Thiele's fraction through 2n+1 points of an [n/n] rational function is
    that function (greedy order or not)."""
    from mbptcode_amd.qp import PadeFit
    rng = np.random.default_rng(5)
    poles = rng.uniform(-2, 2, 3) - 0.3j
    res = rng.uniform(0.1, 1.0, 3)

    def g(z):
        return sum(r / (z - p) for r, p in zip(res, poles))
    z = 0.1 + 1j * np.linspace(0.05, 3.0, 9)
    zq = np.array([0.4 + 0.2j, -1.0 + 0.5j, 1.7 + 0.05j])
    for greedy in (True, False):
        fit = PadeFit(z, g(z), greedy=greedy)
        assert np.abs(fit(zq) - g(zq)).max() < 1e-9


def test_greedy_order_is_the_brute_force_order():
    """The incremental greedy ordering picks, at every step, the node whose
    Thiele coefficient is smallest when a whole table is rebuilt for each
    candidate (the definition, analytic_continuation.py:34-61)."""
    from mbptcode_amd.qp import greedy_pade_order, thiele_coefficients
    rng = np.random.default_rng(17)
    for _ in range(12):
        z = 0.05 + 1j * np.sort(rng.uniform(0.01, 15.0, 12)) * rng.choice([-1, 1])
        poles = rng.uniform(-1, 1, 4) - 1j * rng.uniform(0.02, 0.4, 4)
        amp = rng.uniform(0.05, 0.3, 4)
        f = (amp[None, :] / (z[:, None] - poles[None, :])).sum(1)
        order = [int(np.argmin(np.abs(z)))]
        rest = [j for j in range(len(z)) if j != order[0]]
        while rest:
            sc = [abs(thiele_coefficients(z[order + [j]], f[order + [j]])[-1]) for j in rest]
            order.append(rest.pop(int(np.argmin(sc))))
        assert greedy_pade_order(z, f).tolist() == order


def test_largest_weight_root_prefers_the_quasiparticle():
    """f(w) = w - a - s/(w - pole): two roots; the one of larger Z = 1/f' is
    kept, not the one nearer a."""
    from mbptcode_amd.qp import largest_weight_root
    a, s, pole = 0.0, 0.02, 0.1

    def f(w):
        return w - a - s / (w - pole)
    w, z = largest_weight_root(f, a)
    roots = np.roots([1.0, -(a + pole), a * pole - s])
    zs = [1.0 / (1.0 + s / (r - pole) ** 2) for r in roots]
    assert abs(w - roots[int(np.argmax(zs))]) < 1e-7 and abs(z - max(zs)) < 1e-5


def test_quasiparticle_energies_against_the_exact_roots():
    """The Pade continuation of the exact Sigma(i w) gives the exact roots to
    0.01 meV; the oracle's space-time Sigma to 1 meV."""
    from mbptcode_amd import qp
    inst, D = gw_inputs(0)
    no = inst.nocc
    states = [no - 2, no - 1, no, no + 1]
    xc = np.array([-0.05, -0.05, 0.02, 0.02])
    ax = grids.space_time_gw_grids(inst.eps, no)
    exact = exact_qp(inst, D, states, xc)
    assert np.all((exact[:, 1] > 0.9) & (exact[:, 1] <= 1.0))
    sos = O.gw_sigma_sum_over_poles(inst.X, D, inst.eps, no, states, ax['pade_freq'], ax['mu'])
    e1 = qp.quasiparticle_energies(sos, inst.eps, no, states, ax['mu'], ax['pade_freq'], xc)
    assert np.abs(e1 - exact[:, 0]).max() * HA_MEV < 0.01
    sig, ax = oracle_sigma(inst, D, states)
    e2 = qp.quasiparticle_energies(sig, inst.eps, no, states, ax['mu'], ax['pade_freq'], xc)
    assert np.abs(e2 - exact[:, 0]).max() * HA_MEV < 1.0


def test_pade_matches_the_reference_fixture():
    """Greedy node order, Thiele coefficients (greedy and plain) and values on
    the real axis, full and cut to 9 nodes, of the exact water-size
    self-energy of two states against the reference's own PadeFit
    (tests/golden/transform_weights.npz; analytic_continuation.py:15-89):
    the order exactly, the rest to 1e-10 relative."""
    import os
    from mbptcode_amd.qp import PadeFit
    g = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'transform_weights.npz'))

    def close(got, ref, tol=1e-10):
        return np.abs(np.asarray(got) - ref).max() <= tol * np.abs(ref).max()
    for i in range(2):
        p = f'pade_state{i}_'
        z, f, wq = g[p + 'z'], g[p + 'f'], g[p + 'query']
        fit = PadeFit(z, f)
        assert np.array_equal(fit.z, g[p + 'order_z'])
        assert close(fit.coefficients, g[p + 'coefficients'])
        assert close(fit(wq), g[p + 'values'])
        assert close(PadeFit(z, f, max_order=9)(wq), g[p + 'values_cut9'])
        assert close(PadeFit(z, f, greedy=False).coefficients, g[p + 'plain_coefficients'])


def test_quasiparticle_roots_match_the_reference_fixture():
    """Roots and weights of the quasiparticle equation on the Pade fits of the
    exact water-size self-energy, four states about the gap with a static
    term, against the reference's own root search on the residual its
    qp_solve.py:529-558 forms (tests/golden/transform_weights.npz;
    qp_equation.py:1353-1389): to 2e-8 Ha, twice the root tolerance (they
    agree to the bit), and Z to 1e-8."""
    import os
    from mbptcode_amd.qp import solve_qp_from_imaginary_axis
    g = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'transform_weights.npz'))
    for i, p in enumerate(g['qp_states']):
        w, z = solve_qp_from_imaginary_axis(g['qp_eps'][p], g['qp_xc'][i], g['qp_zfit'][i],
                                            g['qp_sigma'][i], with_z=True)
        assert abs(w - g['qp_roots'][i]) <= 2e-8 and abs(z - g['qp_z'][i]) <= 1e-8, p
