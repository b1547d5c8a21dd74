"""BSE from ISDF factors and orbital energies: the pyscf-free core of the
reference's entry point `solve_bse_isdf` (BSE/solve.py:555-895).

The reference builds the mean field, the factors and the static exchange
term from a pyscf molecule; here they are inputs.  What is kept is the
discipline of the calling sequence:

* one set of factors serves the GW stage, the static W and the BSE action;
* the static W is screened at the mean-field energies (the standard G0W0-BSE
  split, `screen_at='mean-field'`), whatever the diagonal is;
* the diagonal is the mean-field energies (`qp=None`), a caller's array
  (`qp=array`), or the G0W0 quasiparticle energies of every state from the
  space-time self-energy (`qp='G0W0'`, with the caller's static term
  <Sigma_x - v_xc>);
* a W handed in (`W_aux`) is refused beside `qp='G0W0'`, which builds its own.

Returns (omega, X, Y, info) as the reference does; the oscillator strengths
and exciton descriptors need the molecule's dipole integrals and are None.
"""
import time

import numpy as np

from .constants import BSE_DAVIDSON_CONV_TOL
from .isdf_fit import unpack_factors


class _Spectrum:
    """The two fields of the reference's solver object the block action reads."""

    def __init__(self, eps):
        self.eps = np.asarray(eps, float)
        self.spin_mode = 'restricted'


def solve_bse_isdf(eps, nocc, factors, V=None, qp=None, xc_diagonal=None, W_aux=None,
                   tda=False, spin='singlet', nroots=5, conv_tol=BSE_DAVIDSON_CONV_TOL,
                   max_cycle=150, preconditioner='screened', precision='fp64', probe=False,
                   guess=None,
                   refuse_unconverged=False, ntau=None):
    """(omega, X, Y, info) of the `nroots` lowest BSE excitations.

    eps (n_mo,), nocc: the mean-field spectrum.  factors: (X_mo, D[, X_ao,
    coords]) with Z = D D^T on the points; D None takes the grid gauge,
    D = chol(V).  V (M, M): the Coulomb kernel on the points, needed for
    qp='G0W0' (the space-time self-energy runs in the grid gauge) or D None.
    qp: None (BSE on eps), an array of n_mo energies, or 'G0W0' (then
    xc_diagonal, n_mo values of <Sigma_x - v_xc>_pp, is required).
    W_aux: a static W in D's gauge, built at eps when None.
    The remaining keywords are `solve_casida_davidson`'s (tda, spin, probe,
    guess, refuse_unconverged, the Davidson settings and the block action's
    precision)."""
    from .davidson import solve_casida_davidson
    from .grids import static_screening_grid
    from .space_time import static_screening_matrix
    eps = np.asarray(eps, float)
    X, D = unpack_factors(factors)[:2]
    X = np.asarray(X, float)
    if X.shape[1] != len(eps):
        raise ValueError(f'X has {X.shape[1]} orbitals, eps has {len(eps)}')
    if isinstance(qp, str) and qp.upper() != 'G0W0':
        raise ValueError(f"qp={qp!r}: None, an array of energies or 'G0W0'")
    g0w0 = isinstance(qp, str)
    if g0w0 and W_aux is not None:
        raise ValueError("W_aux is built by qp='G0W0' itself; hand it in with an "
                         'explicit qp array instead')
    if g0w0 and (V is None or xc_diagonal is None):
        raise ValueError("qp='G0W0' needs V (the grid Coulomb kernel) and xc_diagonal")
    if D is None:
        if V is None:
            raise ValueError('factors without D need V (the grid gauge, D = chol(V))')
        D = np.linalg.cholesky(np.asarray(V, float))
    D = np.asarray(D, float)
    t = {}
    t0 = time.perf_counter()
    if g0w0:
        from .gw import quasiparticle_energies_space_time
        xc = np.asarray(xc_diagonal, float)
        if xc.shape != eps.shape:
            raise ValueError('xc_diagonal must hold one value per orbital')
        states = np.arange(len(eps))
        eps_qp, z = quasiparticle_energies_space_time(X, V, eps, nocc, states, xc, ntau=ntau,
                                                      with_z=True)
        if not np.all(np.isfinite(eps_qp)):
            raise FloatingPointError('a quasiparticle equation has no root')
        t['qp'] = time.perf_counter() - t0
    elif qp is not None:
        eps_qp, z = np.asarray(qp, float), None
        if eps_qp.shape != eps.shape:
            raise ValueError('qp must hold one energy per orbital')
    else:
        eps_qp, z = eps, None
    t0 = time.perf_counter()
    if W_aux is None:
        W_aux = static_screening_matrix(X, D, eps, nocc, static_screening_grid(eps, nocc))
        t['static_w'] = time.perf_counter() - t0
    W_aux = np.asarray(W_aux, float)
    stats, tracked, timings = {}, {}, {}
    t0 = time.perf_counter()
    omega, Xv, Yv = solve_casida_davidson(
        _Spectrum(eps_qp), nocc, polarizability='BSE', W_aux=W_aux, isdf_factors=(X, D),
        stats=stats, spin=spin, timings=timings, refuse_unconverged=refuse_unconverged,
        guess=guess, tda=tda, tracked=tracked, probe=probe, nroots=nroots,
        conv_tol=conv_tol, max_cycle=max_cycle, preconditioner=preconditioner,
        precision=precision)
    t['davidson'] = time.perf_counter() - t0
    flags = {}
    conv = timings.get('davidson_converged')
    if conv is not None and not all(conv):
        flags['unconverged'] = [i for i, c in enumerate(conv) if not c]
    unconf = timings.get('davidson_unconfirmed')
    if unconf is not None and any(unconf):
        flags['unconfirmed'] = [i for i, c in enumerate(unconf) if c]
    info = {'eps': eps_qp, 'eps_mf': eps, 'qp_weights': z, 'factors': (X, D),
            'W_aux': W_aux, 'min_eig_amb': stats.get('min_eig_amb'),
            'oscillator_strength': None, 'transition_dipole': None,
            'exciton_descriptors': None, 'tracked': tracked, 'davidson_flags': flags,
            'stats': stats, 'timings': {**t, **timings}, 'nranks': 1}
    return omega, Xv, Yv, info
