"""Parity of the device chi0(i tau) build, the static W and the dRPA energy
with the FP64 oracle.  chi0 tolerances relative to max|chi0|: FP64 1e-11,
FP32 1e-5.  Energies: FP64 1e-9 Ha, FP32 1 microHartree."""
import os

import numpy as np
import pytest
import torch

from mbptcode_amd import grids, synth
from oracle import isdf_oracle as O

pytestmark = pytest.mark.gpu

TOL = {'fp64': 1e-11, 'fp32': 1e-5}
GOLDEN = os.path.join(os.path.dirname(__file__), 'golden', 'cfg0_oracle.npz')


def rel(got, ref):
    return float(np.abs(np.asarray(got) - ref).max() / np.abs(ref).max())


def oracle_grid_chi0(inst, tau, weights):
    X_o, X_v, e_o, e_v, mu = O.split_branches(inst.X, inst.eps, inst.nocc)
    I = np.eye(inst.M)
    return sum(w * O.polarizability_projected_tau(X_o, X_v, e_o, e_v, I, t)
               for t, w in zip(tau, weights))


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_golden_single_time_point(precision):
    from mbptcode_amd.space_time import chi0_grid
    gold = np.load(GOLDEN)
    inst = synth.config(0)
    print(f'chi0 golden {precision}', flush=True)
    got = chi0_grid(inst.X, inst.eps, inst.nocc, [gold['tau'][3]], [1.0],
                    precision=precision).cpu().numpy()
    assert rel(got, gold['proj_t3']) < TOL[precision]
    assert np.array_equal(got, got.T) or rel(got, got.T) < TOL[precision]


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
@pytest.mark.parametrize('shape', [(5, 19, 120), (3, 7, 64), (11, 40, 130),
                                   (21, 93, 1200)])
def test_weighted_tau_sum(shape, precision):
    from mbptcode_amd.space_time import chi0_grid
    nocc, nvir, M = shape
    print(f'chi0 {shape} {precision}', flush=True)
    inst = synth.make_instance(nocc, nvir, M, seed=[5, M])
    g = grids.static_screening_grid(inst.eps, nocc, ntau=8)
    ref = oracle_grid_chi0(inst, g.tau_points, g.cosft_wt[0])
    got = chi0_grid(inst.X, inst.eps, nocc, g.tau_points, g.cosft_wt[0],
                    precision=precision).cpu().numpy()
    assert rel(got, ref) < TOL[precision]


def test_pair_sum_identity_at_benzene_size():
    """Grid chi0(tau) = -2 sum_ia e^{-d_ia tau} C_ia C_ia^T at M = 1200 in FP64."""
    from mbptcode_amd.space_time import chi0_grid
    inst = synth.config(1)
    ref = O.proj_tau_bruteforce(inst.X, np.eye(inst.M), inst.eps, inst.nocc, 0.7)
    got = chi0_grid(inst.X, inst.eps, inst.nocc, [0.7], [1.0],
                    precision='fp64').cpu().numpy()
    assert rel(got, ref) < 1e-11


def test_accumulate_adds_onto_out():
    from mbptcode_amd.space_time import Chi0Builder
    inst = synth.make_instance(4, 9, 70, seed=2)
    b = Chi0Builder(inst.X, inst.eps, 4, precision='fp64')
    one = b.build([0.2], [1.0]).clone()
    two = b.build([0.5], [2.0]).clone()
    out = b.build([0.2], [1.0])
    b.build([0.5], [2.0], out=out, accumulate=True)
    both = b.build([0.2, 0.5], [1.0, 2.0])
    assert float((out - (one + two)).abs().max()) < 1e-13 * float(both.abs().max())
    assert float((both - out).abs().max()) < 1e-12 * float(both.abs().max())


def test_reference_signature_projected_tau():
    """polarizability_projected_tau with a factor D (naux < M), NumPy in/out."""
    from mbptcode_amd.space_time import polarizability_projected_tau, split_branches
    inst = synth.make_instance(4, 12, 50, seed=6)
    D = 0.25 * np.random.default_rng(0).normal(size=(50, 9))
    X_o, X_v, e_o, e_v = split_branches(inst.X, inst.eps, 4)[:4]
    got = polarizability_projected_tau(X_o, X_v, e_o, e_v, D, 0.4)
    ref = O.polarizability_projected_tau(X_o, X_v, e_o, e_v, D, 0.4)
    assert rel(got, ref) < 1e-11
    out = np.full((9, 9), np.nan)
    assert polarizability_projected_tau(X_o, X_v, e_o, e_v, D, 0.4, out=out) is out
    assert rel(out, ref) < 1e-11


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_static_screening_grid_matches_oracle(precision):
    from mbptcode_amd.space_time import screened_interaction_grid
    gold = np.load(GOLDEN)
    inst = synth.config(0)
    g = grids.static_screening_grid(inst.eps, inst.nocc, ntau=8)
    W, _ = screened_interaction_grid(inst.X, inst.V, inst.eps, inst.nocc, g,
                                     precision=precision)
    assert rel(W.cpu().numpy(), gold['W_grid']) < (1e-10 if precision == 'fp64' else 2e-5)


def test_static_screening_reference_signature():
    from mbptcode_amd.space_time import static_screening_matrix
    inst = synth.make_instance(5, 15, 80, seed=9)
    D = O.grid_factors(inst.V)
    g = grids.static_screening_grid(inst.eps, 5, ntau=8)
    got = static_screening_matrix(inst.X, D, inst.eps, 5, grid=g)
    ref = O.static_screening_matrix(inst.X, D, inst.eps, 5, g.tau_points, g.cosft_wt[0])
    assert rel(got, ref) < 1e-10


@pytest.mark.parametrize('precision,tol', [('fp64', 1e-9), ('fp32', 1e-6)])
def test_rpa_energy_golden(precision, tol):
    from mbptcode_amd.space_time import rpa_correlation_energy_grid
    gold = np.load(GOLDEN)
    inst = synth.config(0)
    g = grids.rpa_energy_grid(inst.eps, inst.nocc, ntau=8, nfreq=40)
    e = rpa_correlation_energy_grid(inst.X, inst.V, inst.eps, inst.nocc, g,
                                    precision=precision)
    print(f'E_c {precision} {e!r} vs {float(gold["e_c"])!r}', flush=True)
    assert abs(e - float(gold['e_c'])) < tol
    assert abs(e - float(gold['e_c_casida'])) < 1e-6


def test_rpa_energy_reference_signature():
    from mbptcode_amd.space_time import rpa_correlation_energy_space_time
    inst = synth.make_instance(4, 12, 60, seed=2)
    D = O.grid_factors(inst.V)
    g = grids.rpa_energy_grid(inst.eps, 4, ntau=20, nfreq=40)
    e = rpa_correlation_energy_space_time(inst.X, D, inst.eps, 4, g)
    ref = O.rpa_correlation_energy_space_time(inst.X, D, inst.eps, 4, g.tau_points,
                                              g.cosft_wt, g.omega_weights)
    assert abs(e - ref) < 1e-10


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
@pytest.mark.parametrize('shape,ntau', [((5, 19, 120), 8), ((7, 30, 300), 20),
                                        ((4, 9, 130), 40)])
def test_terms_and_transform_against_the_oracle(shape, ntau, precision):
    """Every proj(tau) built once (`mbpt_chi0_terms`) and transformed to all
    frequencies in one pass (`mbpt_chi0_transform`) against the oracle's
    chi0(i w_k) = sum_tau cosft_wt[k, tau] proj(tau) (space_time.py:607-651),
    on shapes with one, three (mirrored off-diagonal tiles) and two 128-point
    tile rows and 8, 20 and 40 time points; then `accumulate` adds a second
    transform on top."""
    from mbptcode_amd.space_time import Chi0Builder
    no, nv, M = shape
    inst = synth.make_instance(no, nv, M, seed=[11, M])
    print(f'chi0 terms+transform {shape} ntau={ntau} {precision}', flush=True)
    g = grids.rpa_energy_grid(inst.eps, no, ntau=min(ntau, 20), nfreq=12)
    tau = np.sort(np.concatenate([g.tau_points, np.linspace(0.05, 3.0, ntau - g.ntau)])) \
        if ntau > g.ntau else g.tau_points
    W = np.random.default_rng(ntau).normal(size=(12, ntau))
    ref = O.chi0_imaginary_frequency(inst.X, np.eye(M), inst.eps, no, tau, W)
    b = Chi0Builder(inst.X, inst.eps, no, precision=precision).terms(tau)
    got = b.transform(W).cpu().numpy()
    for k in range(len(W)):
        assert rel(got[k], ref[k]) < TOL[precision], k
        assert rel(got[k], got[k].T) < TOL[precision], k
    out = b.transform(W[:3])
    b.transform(W[3:6], out=out, accumulate=True)
    assert rel(out.cpu().numpy(), ref[:3] + ref[3:6]) < TOL[precision]


def test_frequency_sweep_guards_fp32_against_large_weights():
    """The benzene-size instance's unregularized 16-point fit carries cosine
    weights of 3e8 with cancelling signs: FP32 terms would lose chi0
    entirely, so the sweep forms those frequencies from FP64 terms, and
    chi0(i w) matches the oracle as an FP64 build does."""
    from mbptcode_amd.space_time import FrequencySweep, fp32_weights_ok
    inst = synth.config(1)
    g = grids.rpa_energy_grid(inst.eps, inst.nocc, ntau=inst.ntau, nfreq=40)
    ok = fp32_weights_ok(g.cosft_wt)
    assert not ok[:30].any() and ok[-1]
    big = synth.config(2)
    g2 = grids.rpa_energy_grid(big.eps, big.nocc, ntau=big.ntau, nfreq=40)
    assert fp32_weights_ok(g2.cosft_wt).all()
    sweep = FrequencySweep(inst.X, inst.eps, inst.nocc, g.tau_points, g.cosft_wt,
                           precision='fp32')
    assert set(sweep.builders) == {'fp32', 'fp64'}
    print('chi0 sweep guard cfg1', flush=True)
    rows = [0, 3, 38, 39]
    got = sweep.build(rows).cpu().numpy()
    ref = O.chi0_imaginary_frequency(inst.X, np.eye(inst.M), inst.eps, inst.nocc,
                                     g.tau_points, g.cosft_wt[rows])
    # the bar is the rounding of the terms amplified by the weights, in the
    # oracle as here: |error| <= 16 u sum_t |w_t| max|P(t)|, u the unit
    # roundoff of the terms' precision (FP64 rows reach 1e-16 x 1e9)
    proj = O.polarizability_projected_sweep(inst.X, np.eye(inst.M), inst.eps, inst.nocc,
                                            g.tau_points)
    pmax = np.abs(proj).reshape(len(proj), -1).max(axis=1)
    for j, k in enumerate(rows):
        u = 1.1e-16 if sweep.use64[k] else 6e-8
        bound = 16 * u * float(np.abs(g.cosft_wt[k]) @ pmax)
        err = float(np.abs(got[j] - ref[j]).max())
        assert err <= bound, (k, err, bound)
