"""The space-time GW self-energy on the GPU against the FP64 oracle and the
exact GW@RPA pole sum (GW/space_time.py:789-849, GW/imaginary_time.py:75-92,
:403-578).  The oracle chain is pinned to the pole sum in
tests/test_gw_oracle.py.

Tolerances (FP32 kernels, FP64 Dyson): the branch sums within 2e-5 of the
largest |sum| per time point; Sigma^c(i w) within 5e-5 of max |Sigma^c|."""
import numpy as np
import pytest
import torch

from mbptcode_amd import grids, synth
from oracle import isdf_oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(5, 19, 120), (13, 50, 333), (21, 93, 1200)]


def nerr(got, ref):
    return float(np.abs(got - ref).max() / np.abs(ref).max())


@pytest.mark.parametrize('shape', SHAPES)
def test_branch_sums_against_oracle(shape):
    """mbpt_chi0_sigma on a symmetric Zt at three time points, for a gap
    window and for 37 states (not a multiple of the 16-state groups)."""
    from mbptcode_amd.space_time import Chi0Builder
    no, nv, M = shape
    inst = synth.make_instance(no, nv, M, seed=[7, no])
    print(f'gw branch sums {shape}', flush=True)
    rng = np.random.default_rng(no)
    Z = rng.normal(size=(3, M, M))
    Z = (Z + Z.transpose(0, 2, 1)).astype(np.float32)
    tau = np.array([0.05, 0.6, 3.0])
    mu = 0.5 * (inst.eps[no - 1] + inst.eps[no])
    b = Chi0Builder(inst.X, inst.eps, no, mu=mu, precision='fp32')
    Zd = torch.as_tensor(Z, device='cuda')
    for states in ([no - 1, no], list(range(min(37, no + nv)))):
        got = b.sigma_sums(tau, Zd, states).cpu().numpy()
        ref = O.self_energy_branch_sums(inst.X, np.eye(M), inst.eps, no, states,
                                        Z.astype(np.float64), tau, mu)
        for br in range(2):
            for t in range(3):
                assert nerr(got[br, t], ref[br, t]) < 2e-5, (states[:3], br, t)


def test_self_energy_matches_the_pole_sum():
    """The whole device chain at the water-size instance against the exact
    GW@RPA self-energy at z = mu + i w."""
    from mbptcode_amd.gw import self_energy_diagonal_space_time
    inst = synth.config(0)
    D = O.grid_factors(inst.V)
    states = [0, inst.nocc - 2, inst.nocc - 1, inst.nocc, inst.nocc + 1, inst.nocc + 5]
    print('gw sigma cfg0 vs pole sum', flush=True)
    sig, ax = self_energy_diagonal_space_time(inst.X, inst.V, inst.eps, inst.nocc, states)
    sos = O.gw_sigma_sum_over_poles(inst.X, D, inst.eps, inst.nocc, states,
                                    ax['pade_freq'], ax['mu'])
    assert sig.shape == (len(states), len(ax['pade_freq']))
    assert nerr(sig, sos) < 5e-5, nerr(sig, sos)


def test_self_energy_against_the_oracle_chain():
    """The benzene-size instance: device chain against the FP64 oracle chain
    on the same axes and weights; the timings name every stage."""
    from mbptcode_amd.gw import self_energy_diagonal_space_time, self_energy_transforms
    inst = synth.config(1)
    D = O.grid_factors(inst.V)
    states = list(range(inst.nocc - 4, inst.nocc + 4))
    print('gw sigma cfg1 vs oracle', flush=True)
    timings = {}
    sig, ax = self_energy_diagonal_space_time(inst.X, inst.V, inst.eps, inst.nocc, states,
                                              timings=timings)
    ref = O.gw_sigma_space_time(inst.X, D, inst.eps, inst.nocc, states, ax['grid'],
                                ax['freq_points'], ax['tau_points'], ax['pade_freq'],
                                self_energy_transforms(ax), ax['mu'])
    assert nerr(sig, ref) < 5e-5, nerr(sig, ref)
    for key in ('transforms', 'screening', 'to_tau', 'sigma'):
        assert timings[key] >= 0
    k = timings['sigma_kernel']
    assert k['launches'] == 1 and k['flops'] > 0 and k['ms'] > 0


def test_reference_signature_with_an_auxiliary_factor():
    """self_energy_diagonal_imaginary_time: a non-square auxiliary factor D
    and W_omega = [I - chi0]^-1 on the screening frequencies, against the
    oracle's branch sums and transforms on the same inputs."""
    from mbptcode_amd.gw import self_energy_diagonal_imaginary_time
    inst = synth.make_instance(6, 20, 150, seed=11)
    rng = np.random.default_rng(11)
    D = 0.2 * rng.normal(size=(150, 40))
    ax = grids.space_time_gw_grids(inst.eps, inst.nocc, ntau=16)
    chi = O.chi0_imaginary_frequency(inst.X, D, inst.eps, inst.nocc, ax['grid'].tau_points,
                                     ax['grid'].cosft_wt, ax['mu'])
    W_omega = np.stack([O.dyson_screening(c) for c in chi])
    states = [4, 5, 6, 7]
    print('gw reference signature', flush=True)
    got = self_energy_diagonal_imaginary_time(inst.X, D, W_omega, inst.eps, inst.nocc, states,
                                              ax['tau_points'], ax['freq_points'],
                                              ax['pade_freq'])
    rW, rS = ax['ranges']
    Ctw = grids.transform_weights(grids.COSINE_WT, ax['tau_points'], ax['freq_points'], *rW,
                                  warn=False)[0]
    C = grids.transform_weights(grids.COSINE_TW, ax['tau_points'], ax['pade_freq'], *rS,
                                warn=False)[0]
    S = grids.transform_weights(grids.SINE_TW, ax['tau_points'], ax['pade_freq'], *rS,
                                warn=False)[0]
    Wt = O.screened_interaction_tau(chi, Ctw)
    ref = O.self_energy_diagonal(
        O.self_energy_branch_sums(inst.X, D, inst.eps, inst.nocc, states, Wt,
                                  ax['tau_points'], ax['mu']), C, S)
    assert nerr(got, ref) < 5e-5, nerr(got, ref)


def test_refusals():
    from mbptcode_amd.space_time import Chi0Builder
    inst = synth.config(0)
    b64 = Chi0Builder(inst.X, inst.eps, inst.nocc, precision='fp64')
    Z = torch.zeros((1, inst.M, inst.M), dtype=torch.float32, device='cuda')
    with pytest.raises(NotImplementedError):
        b64.sigma_sums([0.1], Z, [0])
    b = Chi0Builder(inst.X, inst.eps, inst.nocc, precision='fp32')
    with pytest.raises(ValueError):
        b.sigma_sums([0.1], Z, [inst.nocc + inst.nvir])
    with pytest.raises(ValueError):
        b.sigma_sums([0.1], Z.double(), [0])
    with pytest.raises(ValueError):
        b.sigma_sums([0.1, 0.2], Z, [0])


def test_quasiparticle_energies_on_the_device_chain():
    """G0W0 energies from the device Sigma^c(i w) against the exact GW@RPA
    roots (water size) within 1 meV, with the quasiparticle weights."""
    from mbptcode_amd.gw import quasiparticle_energies_space_time
    from test_gw_oracle import HA_MEV, exact_qp
    inst = synth.config(0)
    D = O.grid_factors(inst.V)
    no = inst.nocc
    states = [no - 2, no - 1, no, no + 1]
    xc = np.array([-0.05, -0.05, 0.02, 0.02])
    print('gw qp energies cfg0', flush=True)
    e, z = quasiparticle_energies_space_time(inst.X, inst.V, inst.eps, no, states, xc,
                                             with_z=True)
    exact = exact_qp(inst, D, states, xc)
    assert np.abs(e - exact[:, 0]).max() * HA_MEV < 1.0
    assert np.abs(z - exact[:, 1]).max() < 1e-3
