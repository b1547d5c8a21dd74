"""Host-side grids against the identities they exist to provide (the cases of
the reference's tests/test_minimax_tau_grid.py:42-146 and tests/test_grids.py).
CPU only."""
import numpy as np
import pytest

from mbptcode_amd import grids

IN_TABLE = [(6, 1.0, 3.0, 1e-6), (8, 1.0, 12.0, 1e-5), (14, 0.3, 30.0, 1e-6),
            (20, 0.3, 30.0, 1e-9), (30, 1.0, 4000.0, 1e-8)]
STRETCHED = [(22, 1.0, 100.0, 1e-9), (24, 1.0, 100.0, 1e-9), (28, 1.0, 100.0, 1e-9),
             (30, 1.0, 100.0, 1e-9), (34, 1.0, 100.0, 1e-9), (34, 0.3, 30.0, 1e-9),
             (30, 2.0, 60.0, 1e-9)]


def relative_laplace_error(ntau, e_min, e_max, nsample=2001):
    tau, w = grids.minimax_time_grid(ntau, e_min, e_max)
    assert tau.shape == (ntau,) and w.shape == (ntau,)
    x = np.geomspace(e_min, e_max, nsample)
    return float(np.abs(1.0 - x * (np.exp(-np.outer(x, tau)) @ w)).max())


@pytest.mark.parametrize('ntau,e_min,e_max,tol', IN_TABLE)
def test_laplace_identity_on_tabulated_ranges(ntau, e_min, e_max, tol):
    assert relative_laplace_error(ntau, e_min, e_max) < tol


@pytest.mark.parametrize('ntau,e_min,e_max,tol', STRETCHED)
def test_narrow_range_stretch_keeps_the_identity(ntau, e_min, e_max, tol):
    assert relative_laplace_error(ntau, e_min, e_max) < tol


def test_stretch_multiplies_time_points():
    """Dividing instead (the frequency grid's rule) collapses the fit."""
    row, stretch = grids.table_row(30, 1.0, 100.0)
    assert stretch > 1.5
    tau, _ = grids.minimax_time_grid(30, 1.0, 100.0)
    assert np.allclose(tau, row[:30] * stretch)


def test_more_points_never_worse():
    err = {n: relative_laplace_error(n, 1.0, 100.0)
           for n in grids.minimax_tau_supported_sizes()}
    assert max(err[n] for n in err if n > 20) < 1e-9


def test_gauss_legendre_integrates_lorentzian():
    for w0 in (0.2, 0.5, 2.0):
        x, w = grids.gauss_legendre_grid(40, w0)
        assert abs(np.sum(w / (1 + x ** 2)) - np.pi / 2) < 1e-8


def test_cosine_transform_fits_the_model_pair():
    """sum_j C[k, j] e^{-x tau_j} ~ 2x/(x^2 + w_k^2) over [e_min, e_max]."""
    e_min, e_max = 0.4, 40.0
    nu, wt = grids.gauss_legendre_grid(10, 0.2)
    for n, tol in ((14, 2e-3), (20, 1e-4)):
        g = grids.TimeFrequencyGrid.minimax_split(n, e_min, e_max, nu, wt)
        x = np.geomspace(e_min, e_max, 500)
        model = np.exp(-np.outer(x, g.tau_points)) @ g.cosft_wt.T
        ref = 2 * x[:, None] / (x[:, None] ** 2 + nu[None, :] ** 2)
        assert np.abs(model - ref).max() / np.abs(ref).max() < tol, n


def test_split_grid_halves_the_laplace_points():
    """The time-frequency pair takes the Laplace grid scaled by 1/(2 e_min)
    (time_frequency.py:524-536, GreenX minimax_grids.F90 `2 e_min`), and its
    omega = 0 cosine row is then a Laplace fit of 2/x on [e_min, e_max]."""
    e_min, e_max = 0.5, 50.0
    g = grids.TimeFrequencyGrid.minimax_split(14, e_min, e_max, [0.0], [1.0])
    tau, w = grids.minimax_time_grid(14, e_min, e_max)
    assert np.allclose(g.tau_points, 0.5 * tau) and np.allclose(g.tau_weights, 0.5 * w)
    x = np.geomspace(e_min, e_max, 1000)
    fit = np.exp(-np.outer(x, g.tau_points)) @ g.cosft_wt[0]
    assert np.abs(fit * x / 2 - 1).max() < 1e-3


def test_auto_point_count_meets_target():
    n, err = grids.minimax_points_for_accuracy(0.5, 5.0, 1e-8)
    assert err <= 1e-8 and n >= 6
    n2, _ = grids.minimax_points_for_accuracy(0.5, 50.0, 1e-8)
    assert n2 >= n


def test_refuses_bad_windows():
    with pytest.raises(ValueError):
        grids.minimax_time_grid(8, 0.0, 1.0)
    with pytest.raises(ValueError):
        grids.minimax_time_grid(7, 0.1, 1.0)
    with pytest.raises(ValueError):
        grids.gauss_legendre_grid(10, -1.0)


def test_transform_weights_match_the_reference_fixture():
    """The minimax split grid of the static W and the three self-energy
    transforms against the reference's own outputs on the same axes
    (tests/golden/transform_weights.npz, written by make_transform_weights.py
    from time_frequency.py:282-380, :498-555).  Tolerance: 1e-11 relative to
    the largest weight (the omega -> tau cosine weights reach 7e8 with
    cancelling signs; the others agree to the bit)."""
    import os
    import warnings
    path = os.path.join(os.path.dirname(__file__), 'golden', 'transform_weights.npz')
    g = np.load(path)
    kinds = {'cos_tw': grids.COSINE_TW, 'cos_wt': grids.COSINE_WT, 'sin_tw': grids.SINE_TW}
    for cfg in (0, 1):
        p = f'cfg{cfg}_static_'
        e_min, e_max = g[p + 'range']
        with warnings.catch_warnings():
            warnings.simplefilter('ignore', RuntimeWarning)
            ours = grids.TimeFrequencyGrid.minimax_split(int(g[p + 'ntau']), e_min, e_max,
                                                         [0.0], [1.0])
        for key, got in (('tau', ours.tau_points), ('tau_weights', ours.tau_weights),
                         ('cosft_wt', ours.cosft_wt)):
            ref = g[p + key]
            assert np.abs(np.asarray(got) - ref).max() <= 1e-11 * np.abs(ref).max(), (cfg, key)
        for name, kind in kinds.items():
            p = f'cfg{cfg}_gw_{name}_'
            W, err = grids.transform_weights(kind, g[p + 'tau'], g[p + 'omega'], *g[p + 'range'],
                                             warn=False)
            ref = g[p + 'weights']
            assert np.abs(W - ref).max() <= 1e-11 * np.abs(ref).max(), (cfg, name)
            assert abs(err - float(g[p + 'fit_error'])) <= 1e-8 * float(g[p + 'fit_error'])


def test_self_energy_axes_match_the_reference_fixture():
    """The self-energy's fit ranges, their minimax point count, the minimax
    frequency grid on rW and the Gauss-Legendre grid against the reference's
    own outputs (tests/golden/transform_weights.npz; time_frequency.py:
    233-254, :401-428, grids.py:107-126, :161-170), to 1e-14 relative."""
    import os
    path = os.path.join(os.path.dirname(__file__), 'golden', 'transform_weights.npz')
    g = np.load(path)

    def close(got, ref):
        got, ref = np.asarray(got, float), np.asarray(ref, float)
        return np.abs(got - ref).max() <= 1e-14 * max(np.abs(ref).max(), 1e-300)
    for cfg in (0, 1):
        p = f'cfg{cfg}_axes_'
        eps, nocc = g[p + 'eps'], int(g[p + 'nocc'])
        rW, rS = grids.sigma_fit_ranges(eps, nocc)
        assert close(rW, g[p + 'rW']) and close(rS, g[p + 'rS']), cfg
        n, err = grids.minimax_points_for_ranges([rW[1] / rW[0], rS[1] / rS[0]])
        assert n == int(g[p + 'npoints'][0]) and np.isclose(err, g[p + 'npoints'][1], rtol=1e-12)
        fp, fw = grids.minimax_frequency_grid(16, *rW)
        assert close(fp, g[p + 'freq16_points']) and close(fw, g[p + 'freq16_weights']), cfg
        gp, gw = grids.gauss_legendre_grid(16, 1.0)
        assert close(gp, g[p + 'gl16_points']) and close(gw, g[p + 'gl16_weights']), cfg


@pytest.mark.parametrize('cfg', [2, 4])
def test_regularized_twenty_point_grids_match_the_reference_bitwise(cfg):
    """The regularized branch of the transform fit (20 points and more,
    regularization 1e-8, time_frequency.py:325-331), which the static W and
    the dRPA energy of the large (60/400/4000) and learned-preconditioner
    (100/700/6000) instances run at 20 time points: time points, time weights
    and cosine weights equal, bit for bit, the reference's own on the same
    windows (tests/golden/transform_weights.npz, written by
    make_transform_weights.py), for the static grid (one frequency, 0) and the
    Gauss-Legendre 40-frequency energy grid (gradients/rpa/ground_state.py:
    92-106)."""
    import os
    import warnings
    from mbptcode_amd import synth
    g = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'transform_weights.npz'))
    inst = synth.config(cfg)
    assert int(g[f'cfg{cfg}_static_ntau']) == inst.ntau == 20
    with warnings.catch_warnings():
        warnings.simplefilter('ignore', RuntimeWarning)
        static = grids.static_screening_grid(inst.eps, inst.nocc, ntau=inst.ntau)
        energy = grids.rpa_energy_grid(inst.eps, inst.nocc, ntau=inst.ntau)
    for p, ours in ((f'cfg{cfg}_static_', static), (f'cfg{cfg}_rpa_', energy)):
        assert np.array_equal(np.asarray(ours.tau_points), g[p + 'tau']), p
        assert np.array_equal(np.asarray(ours.tau_weights), g[p + 'tau_weights']), p
        assert np.array_equal(np.asarray(ours.cosft_wt), g[p + 'cosft_wt']), p
    assert np.array_equal(np.asarray(energy.omega_points), g[f'cfg{cfg}_rpa_freq_points'])
    assert np.array_equal(np.asarray(energy.omega_weights), g[f'cfg{cfg}_rpa_freq_weights'])
