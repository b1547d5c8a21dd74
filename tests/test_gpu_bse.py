"""The BSE entry on the device (the core of BSE/solve.py:555-895): roots
against dense eigensolves of A (and the dense Casida roots) built from the
same W_aux at the same diagonal, for the mean-field diagonal, a caller's
array and the G0W0 energies from the space-time self-energy."""
import numpy as np
import pytest

from mbptcode_amd import synth
from oracle import isdf_oracle as O

pytestmark = pytest.mark.gpu

HA_MEV = 27211.386245988


def dense_roots(inst, D, eps_diag, W_aux, tda, n):
    A = O.dense_tda_matrix(eps_diag, inst.nocc, inst.X, D, W_aux)
    if tda:
        return np.linalg.eigvalsh(A)[:n]
    B = O.dense_casida_b_matrix(inst.nocc, inst.X, D, W_aux)
    return O.dense_casida_roots(A, B)[0][:n]


@pytest.mark.parametrize('tda', [True, False])
def test_mean_field_and_array_diagonals(tda):
    from mbptcode_amd.bse import solve_bse_isdf
    inst = synth.config(1)
    D = O.grid_factors(inst.V)
    shift = np.where(np.arange(len(inst.eps)) < inst.nocc, -0.05, 0.1)
    for qp in (None, inst.eps + shift):
        print(f'bse entry tda={tda} qp={"array" if qp is not None else None}', flush=True)
        w, X, Y, info = solve_bse_isdf(inst.eps, inst.nocc, (inst.X, None), V=inst.V, qp=qp,
                                       tda=tda, nroots=5, conv_tol=1e-6,
                                       probe=False if tda else 'sign')
        ref = dense_roots(inst, D, info['eps'], info['W_aux'], tda, 5)
        assert np.abs(w - ref).max() * HA_MEV < 1.0
        assert np.array_equal(info['eps_mf'], inst.eps)
        assert not info['davidson_flags']
        if not tda:
            assert info['min_eig_amb'] > 0


def test_g0w0_diagonal():
    from mbptcode_amd.bse import solve_bse_isdf
    inst = synth.config(0)
    D = O.grid_factors(inst.V)
    print('bse entry G0W0', flush=True)
    w, X, Y, info = solve_bse_isdf(inst.eps, inst.nocc, (inst.X, D), V=inst.V, qp='G0W0',
                                   xc_diagonal=np.zeros(len(inst.eps)), tda=True,
                                   nroots=5, conv_tol=1e-7)
    eq = info['eps']
    assert np.all(np.isfinite(eq)) and not np.array_equal(eq, inst.eps)
    assert np.all((info['qp_weights'] > 0.5) & (info['qp_weights'] <= 1.0))
    # the occupied levels move down and the virtual ones up is not guaranteed;
    # the roots are those of the dense A on the quasiparticle diagonal
    ref = dense_roots(inst, D, eq, info['W_aux'], True, 5)
    assert np.abs(w - ref).max() * HA_MEV < 1.0


def _factor_instance():
    inst = synth.make_instance(5, 19, 120, seed=13)
    rng = np.random.default_rng(13)
    D = O.grid_factors(inst.V)[:, :90] + 0.01 * rng.normal(size=(120, 90))
    return inst, D


def test_isdf_bse_factors_imaginary_time_and_df():
    """isdf_bse_factors (BSE/solve.py:1038-1142) on given factors: the
    imaginary-time W against the oracle's static_screening_matrix on the same
    grid, the 'df' W against [1 - chi0(0)]^-1 with the closed-form static
    chi0 -4 sum_ia B_ia B_ia^T / d_ia, and the two routes against each other
    to the transform's fit error at 20 time points."""
    from mbptcode_amd import grids
    from mbptcode_amd.bse import isdf_bse_factors
    from mbptcode_amd.numerics import Numerics
    inst, D = _factor_instance()
    num = Numerics(quadrature={'ntau_w': 20})
    print('isdf_bse_factors imaginary-time / df', flush=True)
    X, D2, W_it = isdf_bse_factors(inst.eps, inst.nocc, (inst.X, D), numerics=num)
    assert D2 is not None and W_it.shape == (90, 90)
    g = grids.static_screening_grid(inst.eps, inst.nocc, ntau=20)
    ref = O.static_screening_matrix(inst.X, D, inst.eps, inst.nocc, g.tau_points, g.cosft_wt[0])
    assert np.abs(W_it - ref).max() <= 1e-10 * np.abs(ref).max()
    _, _, W_df = isdf_bse_factors(inst.eps, inst.nocc, (inst.X, D), screening='df')
    chi = O.chi0_omega_closed_form(inst.X, D, inst.eps, inst.nocc, 0.0)
    exact = np.linalg.inv(np.eye(90) - chi)
    assert np.abs(W_df - exact).max() <= 1e-12 * np.abs(exact).max()
    assert np.abs(W_it - W_df).max() <= 1e-6 * np.abs(W_df).max()
    with pytest.raises(ValueError):
        isdf_bse_factors(inst.eps, inst.nocc, (inst.X, D), screening='casida')


def test_solve_bse_isdf_reads_numerics():
    """The Davidson settings come from numerics.bse (nroots, conv_tol) when the
    keywords are left None, and the static W from numerics.quadrature."""
    from mbptcode_amd.bse import solve_bse_isdf
    from mbptcode_amd.numerics import Numerics
    inst = synth.config(0)
    D = O.grid_factors(inst.V)
    num = Numerics(bse={'nroots': 3, 'conv_tol': 1e-7}, quadrature={'ntau_w': 8})
    print('solve_bse_isdf with numerics', flush=True)
    w, X, Y, info = solve_bse_isdf(inst.eps, inst.nocc, (inst.X, D), tda=True, numerics=num)
    assert len(w) == 3
    assert max(info['timings']['davidson_residuals']) <= 1e-7
    ref = dense_roots(inst, D, inst.eps, info['W_aux'], True, 3)
    assert np.abs(w - ref).max() * HA_MEV < 1e-3
