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
