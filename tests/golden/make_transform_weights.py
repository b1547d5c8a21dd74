"""Writes tests/golden/transform_weights.npz: the reference's own minimax
transform weights (src/Base/utils/grids/time_frequency.py:282-380, the
`TimeFrequencyGrid.minimax_split` path at :498-555) on the axes the static W
and the space-time self-energy use for the water- and benzene-size instances,
the regularized 20-point static and dRPA-energy grids of the large and
learned-preconditioner instances,
with those axes' fit ranges (:401-428), point count (:233-254) and the
minimax and Gauss-Legendre frequency grids (grids.py:107-126, :161-170), and
the Thiele-Pade continuation of an exact self-energy
(analytic_continuation.py:15-89) with the quasiparticle roots the reference's
root search finds on it (Solvers/qp_equation.py:1353-1389).

The weights come from a regularized least-squares fit whose last bits depend
on the path through the SVD, so only the reference's code gives them; this
script imports that module from a reference checkout and records inputs and
outputs.  Run from the repository root, with the checkout at /root/reference:

    python tests/golden/make_transform_weights.py

The fixture is data only (axes, ranges, weights); nothing of the reference
is kept.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
REFERENCE = os.environ.get('MBPT_REFERENCE', '/root/reference')


def main():
    from mbptcode_amd import grids, synth
    sys.path.insert(0, REFERENCE)
    from src.Base.utils.grids import time_frequency as TF   # the reference
    from src.Base.utils.grids import grids as RG             # the reference
    out = {}
    kinds = {'cos_tw': grids.COSINE_TW, 'cos_wt': grids.COSINE_WT, 'sin_tw': grids.SINE_TW}
    for cfg in (0, 1):
        inst = synth.config(cfg)
        sg = grids.static_screening_grid(inst.eps, inst.nocc, ntau=inst.ntau)
        e_min, e_max = sg.meta['e_min'], sg.meta['e_max']
        ref = TF.TimeFrequencyGrid.minimax_split(inst.ntau, e_min, e_max, np.array([0.0]),
                                                 np.array([1.0]))
        p = f'cfg{cfg}_static_'
        out[p + 'range'] = np.array([e_min, e_max])
        out[p + 'ntau'] = np.array(inst.ntau)
        out[p + 'tau'] = np.asarray(ref.tau_points, float)
        out[p + 'tau_weights'] = np.asarray(ref.tau_weights, float)
        out[p + 'cosft_wt'] = np.asarray(ref.cosft_wt, float)
        ax = grids.space_time_gw_grids(inst.eps, inst.nocc, ntau=16)
        rW, rS = ax['ranges']
        # the self-energy's fit ranges, their point count and frequency grid
        p = f'cfg{cfg}_axes_'
        refW, refS = TF.sigma_fit_ranges(inst.eps, inst.nocc)
        out[p + 'eps'] = np.asarray(inst.eps, float)
        out[p + 'nocc'] = np.array(inst.nocc)
        out[p + 'rW'] = np.array(refW, float)
        out[p + 'rS'] = np.array(refS, float)
        n, err = TF.minimax_points_for_ranges([refW[1] / refW[0], refS[1] / refS[0]])
        out[p + 'npoints'] = np.array([n, err], float)
        fp, fw = RG.minimax_frequency_grid(16, *refW)
        out[p + 'freq16_points'] = np.asarray(fp, float)
        out[p + 'freq16_weights'] = np.asarray(fw, float)
        gp, gw = RG.gauss_legendre_grid(16, 1.0)
        out[p + 'gl16_points'] = np.asarray(gp, float)
        out[p + 'gl16_weights'] = np.asarray(gw, float)
        for name, kind in kinds.items():
            omega = ax['freq_points'] if name == 'cos_wt' else ax['pade_freq']
            rng = rW if name == 'cos_wt' else rS
            W, err = TF.minimax_transform_weights(kind, ax['tau_points'], omega, *rng, warn=False)
            p = f'cfg{cfg}_gw_{name}_'
            out[p + 'tau'] = np.asarray(ax['tau_points'], float)
            out[p + 'omega'] = np.asarray(omega, float)
            out[p + 'range'] = np.array(rng, float)
            out[p + 'weights'] = np.asarray(W, float)
            out[p + 'fit_error'] = np.array(err, float)
    # the regularized branch (20 time points and more, time_frequency.py:
    # 325-331) on the static W of the large and learned-preconditioner
    # instances and on their dRPA energy grids (Gauss-Legendre frequencies,
    # w0 = 0.5 gap; gradients/rpa/ground_state.py:92-106)
    for cfg in (2, 4):
        inst = synth.config(cfg)
        sg = grids.static_screening_grid(inst.eps, inst.nocc, ntau=inst.ntau)
        e_min, e_max = sg.meta['e_min'], sg.meta['e_max']
        ref = TF.TimeFrequencyGrid.minimax_split(inst.ntau, e_min, e_max, np.array([0.0]),
                                                 np.array([1.0]))
        p = f'cfg{cfg}_static_'
        out[p + 'range'] = np.array([e_min, e_max])
        out[p + 'ntau'] = np.array(inst.ntau)
        out[p + 'tau'] = np.asarray(ref.tau_points, float)
        out[p + 'tau_weights'] = np.asarray(ref.tau_weights, float)
        out[p + 'cosft_wt'] = np.asarray(ref.cosft_wt, float)
        nu, wt = RG.gauss_legendre_grid(40, RG.gap_scaled_w0(inst.eps, inst.nocc))
        ref = TF.TimeFrequencyGrid.minimax_split(inst.ntau, e_min, e_max, np.asarray(nu),
                                                 np.asarray(wt), with_sine=False,
                                                 with_inverse=False)
        p = f'cfg{cfg}_rpa_'
        out[p + 'range'] = np.array([e_min, e_max])
        out[p + 'ntau'] = np.array(inst.ntau)
        out[p + 'freq_points'] = np.asarray(nu, float)
        out[p + 'freq_weights'] = np.asarray(wt, float)
        out[p + 'tau'] = np.asarray(ref.tau_points, float)
        out[p + 'tau_weights'] = np.asarray(ref.tau_weights, float)
        out[p + 'cosft_wt'] = np.asarray(ref.cosft_wt, float)
    # Thiele-Pade continuation of the exact GW@RPA self-energy of the
    # water-size instance at the Pade frequencies (the oracle's pole sum as
    # input), and the fraction's values on the real axis
    from src.Base.utils.grids import analytic_continuation as AC   # the reference
    from oracle import isdf_oracle as O
    inst = synth.config(0)
    ax = grids.space_time_gw_grids(inst.eps, inst.nocc, ntau=16)
    states = [inst.nocc - 1, inst.nocc]
    z = ax['mu'] + 1j * np.asarray(ax['pade_freq'], float)
    f = O.gw_sigma_sum_over_poles(inst.X, O.grid_factors(inst.V), inst.eps, inst.nocc, states,
                                  ax['pade_freq'], ax['mu'])
    wq = np.linspace(inst.eps[inst.nocc - 1] - 0.6, inst.eps[inst.nocc] + 0.6, 41) + 0j
    for i, p_state in enumerate(states):
        fit = AC.PadeFit(z, f[i])
        cut = AC.PadeFit(z, f[i], max_order=9)
        p = f'pade_state{i}_'
        out[p + 'z'] = z
        out[p + 'f'] = np.asarray(f[i], complex)
        out[p + 'order_z'] = np.asarray(fit.z, complex)
        out[p + 'coefficients'] = np.asarray(fit.coefficients, complex)
        out[p + 'query'] = wq
        out[p + 'values'] = np.asarray(fit(wq), complex)
        out[p + 'values_cut9'] = np.asarray(cut(wq), complex)
        out[p + 'plain_coefficients'] = np.asarray(AC.PadeFit(z, f[i], greedy=False).coefficients,
                                                   complex)
    # the quasiparticle equation on those continuations: the reference's
    # root search (qp_equation.py:1353, `solve_qp_equation_pole_strength`) on
    # the residual its qp_solve.py:529-558 forms, four states around the gap
    from src.Solvers import qp_equation as QE                       # the reference
    from mbptcode_amd.qp import imaginary_axis_sample_points
    D = O.grid_factors(inst.V)
    no = inst.nocc
    qstates = [no - 2, no - 1, no, no + 1]
    xc = np.array([-0.05, -0.05, 0.02, 0.02])
    roots, zs, zfit, sfit = [], [], [], []
    for i, p_state in enumerate(qstates):
        zf, _ = imaginary_axis_sample_points(ax['pade_freq'], no, p_state, ax['mu'])
        sf = O.gw_sigma_sum_over_poles(inst.X, D, inst.eps, no, [p_state], None, None,
                                       z=zf)[0]
        fit = AC.PadeFit(zf, sf, greedy=True)

        def residual(w, fit=fit, p_state=p_state, x=xc[i]):
            return w - inst.eps[p_state] - x - fit(np.array([w], dtype=complex))[0].real
        root = QE.solve_qp_equation_pole_strength(residual, inst.eps[p_state])
        roots.append(float(root))
        zs.append(float(QE.pole_strength(residual, root)))
        zfit.append(zf)
        sfit.append(sf)
    out['qp_states'] = np.array(qstates)
    out['qp_eps'] = np.asarray(inst.eps, float)
    out['qp_xc'] = xc
    out['qp_zfit'] = np.array(zfit)
    out['qp_sigma'] = np.array(sfit)
    out['qp_roots'] = np.array(roots)
    out['qp_z'] = np.array(zs)
    path = os.path.join(HERE, 'transform_weights.npz')
    np.savez(path, **out)
    print(f'wrote {path}: {len(out)} arrays')


if __name__ == '__main__':
    main()
