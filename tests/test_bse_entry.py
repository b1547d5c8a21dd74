"""The BSE entry's refusals (BSE/solve.py:555-720), checked before any device
work: a W handed in beside qp='G0W0', an unknown qp, a G0W0 request without
the grid Coulomb kernel or the static term, factors without D and no V."""
import numpy as np
import pytest

from mbptcode_amd.bse import solve_bse_isdf


def test_refusals():
    rng = np.random.default_rng(0)
    X = rng.normal(size=(20, 6))
    eps = np.array([-0.8, -0.5, 0.1, 0.4, 0.9, 1.3])
    V = np.eye(20)
    with pytest.raises(ValueError):
        solve_bse_isdf(eps, 2, (X, None), V=V, qp='G0W0', xc_diagonal=np.zeros(6),
                       W_aux=np.eye(20))
    with pytest.raises(ValueError):
        solve_bse_isdf(eps, 2, (X, None), V=V, qp='evGW')
    with pytest.raises(ValueError):
        solve_bse_isdf(eps, 2, (X, None), V=V, qp='G0W0')
    with pytest.raises(ValueError):
        solve_bse_isdf(eps, 2, (X, None))
    with pytest.raises(ValueError):
        solve_bse_isdf(eps[:5], 2, (X, None), V=V)
