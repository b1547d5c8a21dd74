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


def isdf_bse_factors(eps, nocc, factors, screening='imaginary-time', numerics=None,
                     precision='fp64'):
    """(X_mo, D, W_aux) for the ISDF Davidson BSE from one set of factors: the
    pyscf-free part of `isdf_bse_factors` (BSE/solve.py:1038-1142), with the
    factors given (the reference's `factors=` route) and the mean field's
    orbital energies as `eps`.

    W_aux = [1 - chi0(i0)]^-1 is built in D's auxiliary gauge from the
    factors themselves, so the block action and W cannot mix gauges:

    * screening='imaginary-time' (default): chi0(0) in imaginary time on
      `static_screening_grid(eps, nocc, ntau=quadrature.ntau_w,
      tau_target=quadrature.tau_target)` (`static_screening_matrix`), cubic
      and on the device kernels at `precision`;
    * screening='df': chi0(0) = -4 sum_ia B_ia B_ia^T / d_ia from the
      three-index factor the factors imply, B_A,ia = sum_k X_ki X_ka D_kA
      (`isdf_df_coefficients`, BSE/solve.py:396-407), exact for the factors
      but (naux, n_occ n_vir)-sized: small systems only.
    """
    from .numerics import numerics_or_default
    from .space_time import static_screening_matrix
    from .grids import static_screening_grid
    eps = np.asarray(eps, float)
    X, D = unpack_factors(factors)[:2]
    if D is None:
        raise ValueError('isdf_bse_factors needs the factor D of (X_mo, D)')
    X, D = np.asarray(X, float), np.asarray(D, float)
    if X.shape[1] != len(eps):
        raise ValueError(f'X has {X.shape[1]} orbitals, eps has {len(eps)}')
    if D.shape[0] != X.shape[0]:
        raise ValueError(f'X has {X.shape[0]} points, D has {D.shape[0]}')
    if screening == 'df':
        return X, D, _static_w_from_df(X, D, eps, nocc)
    if screening != 'imaginary-time':
        raise ValueError(f"screening={screening!r}; choose 'imaginary-time' or 'df'.")
    q = numerics_or_default(numerics).quadrature
    grid = static_screening_grid(eps, nocc, ntau=q.ntau_w, tau_target=q.tau_target)
    return X, D, static_screening_matrix(X, D, eps, nocc, grid, precision=precision)


#: Device bytes the 'df' static screening may take for its pair factors.
DF_SCREENING_BYTES = 8 * 2 ** 30


def _static_w_from_df(X, D, eps, nocc):
    """[1 - chi0(0)]^-1 with chi0(0) = 2 C f C^T, f = -2 / d (the static limit
    of imaginary_frequency.py:91-122 on the three-index factor of X and D),
    FP64 on the device."""
    import torch
    from .orbital_spaces import bse_pair_diagonal
    from .space_time import dyson_screening
    M, naux = D.shape
    npair = nocc * (X.shape[1] - nocc)
    if 8 * npair * (M + naux) > DF_SCREENING_BYTES:
        raise ValueError(f"screening='df' would hold {npair} x ({M} + {naux}) doubles; "
                         "use screening='imaginary-time' at this size")
    dev = torch.device('cuda', torch.cuda.current_device())
    Xd = torch.as_tensor(X, device=dev)
    Dd = torch.as_tensor(D, device=dev)
    C = (Xd[:, :nocc, None] * Xd[:, None, nocc:]).reshape(M, npair)
    B = Dd.T @ C
    d = torch.as_tensor(bse_pair_diagonal(eps, nocc).ravel(), device=dev)
    chi = -4.0 * (B / d[None, :]) @ B.T
    return dyson_screening(0.5 * (chi + chi.T)).cpu().numpy()


def solve_bse_isdf(eps, nocc, factors, V=None, qp=None, xc_diagonal=None, W_aux=None,
                   tda=False, spin='singlet', nroots=None, conv_tol=None,
                   max_cycle=None, preconditioner=None, precision='fp64', probe=False,
                   guess=None, refuse_unconverged=False, ntau=None, numerics=None):
    """(omega, X, Y, info) of the `nroots` lowest BSE excitations.

    eps (n_mo,), nocc: the mean-field spectrum.  factors: (X_mo, D[, X_ao,
    coords]) with Z = D D^T on the points; D None takes the grid gauge,
    D = chol(V).  V (M, M): the Coulomb kernel on the points, needed for
    qp='G0W0' (the space-time self-energy runs in the grid gauge) or D None.
    qp: None (BSE on eps), an array of n_mo energies, or 'G0W0' (then
    xc_diagonal, n_mo values of <Sigma_x - v_xc>_pp, is required).
    W_aux: a static W in D's gauge, built at eps when None (`isdf_bse_factors`,
    its time points `ntau`, else numerics.quadrature.ntau_w, else sized by the
    Laplace test).  numerics: a `Numerics`; the Davidson settings nroots,
    conv_tol, max_cycle and preconditioner left None come from its bse
    section (numerics.py:241-286; None: the defaults, 5 roots to 1e-8).
    The remaining keywords are `solve_casida_davidson`'s (tda, spin, probe,
    guess, refuse_unconverged and the block action's precision)."""
    from .numerics import numerics_or_default
    num = numerics_or_default(numerics)
    nroots = num.bse.nroots if nroots is None else nroots
    conv_tol = num.bse.conv_tol if conv_tol is None else conv_tol
    max_cycle = num.bse.max_cycle if max_cycle is None else max_cycle
    preconditioner = num.bse.preconditioner if preconditioner is None else preconditioner
    from .davidson import solve_casida_davidson
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
        if ntau is not None:
            num = num.derive('static W at the requested time points',
                             quadrature={'ntau_w': int(ntau)}) \
                if num.quadrature.ntau_w != int(ntau) else num
        W_aux = isdf_bse_factors(eps, nocc, (X, D), numerics=num)[2]
        t['static_w'] = time.perf_counter() - t0
    W_aux = np.asarray(W_aux, float)
    stats, tracked, timings = {}, {}, {}
    t0 = time.perf_counter()
    omega, Xv, Yv = solve_casida_davidson(
        _Spectrum(eps_qp), nocc, polarizability='BSE', W_aux=W_aux, isdf_factors=(X, D),
        stats=stats, spin=spin, timings=timings, refuse_unconverged=refuse_unconverged,
        guess=guess, tda=tda, tracked=tracked, probe=probe, nroots=nroots,
        conv_tol=conv_tol, max_cycle=max_cycle, preconditioner=preconditioner,
        guess_factor=num.bse.guess_factor, precision=precision)
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
