"""Parity of the device sigma-vector (libmbpt_hip.so through its C ABI) with
the FP64 oracle.  Tolerances are relative to max|A z|: FP64 1e-11, FP32 1e-5
(single-precision accumulation over up to a few thousand grid points)."""
import os

import numpy as np
import pytest
import torch

from mbptcode_amd import grids, synth
from oracle import isdf_oracle as O

pytestmark = pytest.mark.gpu

TOL = {'fp64': 1e-11, 'fp32': 1e-5}
GOLDEN = os.path.join(os.path.dirname(__file__), 'golden', 'cfg0_oracle.npz')


def grid_inputs(inst, seed=11, screened=True):
    """(W_grid, D, W_aux): a symmetric positive screened interaction on the
    grid and its factor-gauge twin for the oracle."""
    D = O.grid_factors(inst.V)
    if screened:
        g = grids.static_screening_grid(inst.eps, inst.nocc, ntau=8)
        W_aux = O.static_screening_matrix(inst.X, D, inst.eps, inst.nocc,
                                          g.tau_points, g.cosft_wt[0])
    else:
        rng = np.random.default_rng(seed)
        B = rng.normal(size=(inst.M, inst.M)) / inst.M
        W_aux = np.eye(inst.M) + 0.5 * (B + B.T)
    return O.grid_screened_interaction(W_aux, D), D, W_aux


def device_action(inst, W_grid, precision, kappa=2.0):
    from mbptcode_amd.block_actions import SigmaOperator
    return SigmaOperator(inst.eps, inst.nocc, inst.X, W_grid, inst.V, kappa=kappa,
                         precision=precision)


def rel(got, ref):
    return float(np.abs(got - ref).max() / np.abs(ref).max())


def apply_np(op, z):
    zt = torch.as_tensor(np.ascontiguousarray(z), device='cuda')
    out = op.apply(zt)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_golden_water_size(precision):
    gold = np.load(GOLDEN)
    inst = synth.config(0)
    print(f'sigma golden cfg0 {precision}', flush=True)
    op = device_action(inst, gold['W_grid'], precision)
    assert rel(apply_np(op, gold['z']), gold['Az']) < TOL[precision]
    sd = op.screened_diagonal_device().cpu().numpy()
    assert rel(sd, gold['sdiag']) < TOL[precision]


# (nocc, nvir, M): tiny, ragged, every exchange width bucket, max nocc
SHAPES = [(1, 3, 10), (2, 5, 64), (17, 29, 130), (33, 20, 65), (60, 70, 257),
          (100, 40, 200), (128, 10, 70)]


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
@pytest.mark.parametrize('shape', SHAPES)
def test_shapes_against_oracle(shape, precision):
    nocc, nvir, M = shape
    print(f'sigma shape {shape} {precision}', flush=True)
    inst = synth.make_instance(nocc, nvir, M, seed=[3, nocc, M])
    W_grid, D, W_aux = grid_inputs(inst, screened=False)
    z = np.random.default_rng(nocc).normal(size=(5, nocc, nvir))
    ref = O.isdf_block_action(inst.eps, nocc, inst.X, D, W_aux)[0](z)
    op = device_action(inst, W_grid, precision)
    assert rel(apply_np(op, z), ref) < TOL[precision]


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_triplet_has_no_hartree(precision):
    inst = synth.make_instance(7, 30, 150, seed=4)
    W_grid, D, W_aux = grid_inputs(inst, screened=False)
    z = np.random.default_rng(1).normal(size=(4, 7, 30))
    ref = O.isdf_block_action(inst.eps, 7, inst.X, D, W_aux, kappa=0.0)[0](z)
    op = device_action(inst, W_grid, precision, kappa=0.0)
    assert rel(apply_np(op, z), ref) < TOL[precision]


def test_many_vectors_then_fewer_reuse_the_workspace():
    """k = 70 (several Hartree batches, several split slots) then k = 3 and
    k = 1 on the same handle: each result is the oracle's."""
    inst = synth.make_instance(9, 25, 300, seed=8)
    W_grid, D, W_aux = grid_inputs(inst, screened=False)
    act = O.isdf_block_action(inst.eps, 9, inst.X, D, W_aux)[0]
    op = device_action(inst, W_grid, 'fp64')
    rng = np.random.default_rng(2)
    for k in (70, 3, 1):
        print(f'sigma k={k}', flush=True)
        z = rng.normal(size=(k, 9, 25))
        assert rel(apply_np(op, z), act(z)) < TOL['fp64']


def test_benzene_size_fp32():
    inst = synth.config(1)
    print('sigma cfg1 fp32', flush=True)
    W_grid, D, W_aux = grid_inputs(inst, screened=False)
    z = np.random.default_rng(5).normal(size=(10, inst.nocc, inst.nvir))
    ref = O.isdf_block_action(inst.eps, inst.nocc, inst.X, D, W_aux)[0](z)
    for precision in ('fp64', 'fp32'):
        op = device_action(inst, W_grid, precision)
        assert rel(apply_np(op, z), ref) < TOL[precision]


def test_linearity_and_symmetry_at_full_size():
    """At the large single-instance size: A(a z1 + b z2) = a A z1 + b A z2 and
    <z1, A z2> = <A z1, z2> (A is symmetric)."""
    inst = synth.make_instance(60, 400, 4000, seed=0)
    rng = np.random.default_rng(0)
    B = rng.normal(size=(inst.M, inst.M)) / inst.M
    W_grid = inst.V + 0.1 * (B + B.T)
    print('sigma linearity cfg2 size', flush=True)
    op = device_action(inst, W_grid, 'fp64')
    z = torch.as_tensor(rng.normal(size=(2, 60, 400)), device='cuda')
    Az = op.apply(z)
    comb = op.apply((0.3 * z[0] - 1.7 * z[1]).unsqueeze(0).contiguous())
    lin = (comb[0] - (0.3 * Az[0] - 1.7 * Az[1])).abs().max() / Az.abs().max()
    sym = abs(float((z[0] * Az[1]).sum() - (Az[0] * z[1]).sum())) / float(Az.abs().max())
    assert float(lin) < 1e-12
    assert sym < 1e-9


def test_empty_batch_and_shape_errors():
    from mbptcode_amd.block_actions import SigmaOperator, isdf_block_action
    inst = synth.make_instance(3, 4, 20, seed=1)
    op = device_action(inst, inst.V, 'fp64')
    out = op.apply(torch.zeros((0, 3, 4), dtype=torch.float64, device='cuda'))
    assert out.shape == (0, 3, 4)
    with pytest.raises(ValueError):
        op.apply(torch.zeros((1, 5), dtype=torch.float64, device='cuda'))
    with pytest.raises(ValueError):
        SigmaOperator(inst.eps, 3, inst.X, inst.V[:10], inst.V)
    big = synth.make_instance(129, 2, 10, seed=1)
    with pytest.raises(ValueError):
        SigmaOperator(big.eps, 129, big.X, big.V, big.V)

    class LR:
        eps = inst.eps
        spin_mode = 'restricted'

    D = O.grid_factors(inst.V)
    Wa = np.eye(20)
    Wa[0, 1] = 1.0
    with pytest.raises(ValueError):
        isdf_block_action(LR, 3, True, Wa, (inst.X, D), tda=True)
    with pytest.raises(ValueError):
        isdf_block_action(LR, 3, True, np.eye(19), (inst.X, D), tda=True)


def test_reference_signature_entry():
    """isdf_block_action with factor-gauge (X, D) and W_aux, NumPy in and out,
    against the oracle on the same arguments."""
    from mbptcode_amd.block_actions import isdf_block_action
    inst = synth.make_instance(6, 14, 90, seed=3)
    rng = np.random.default_rng(3)
    D = 0.3 * rng.normal(size=(90, 11))
    W = rng.normal(size=(11, 11))
    W = 0.5 * (W + W.T) + 11 * np.eye(11)

    class LR:
        eps = inst.eps
        spin_mode = 'restricted'

    z = rng.normal(size=(3, 6, 14))
    for spin, kappa in (('singlet', 2.0), ('triplet', 0.0)):
        apply_AB, d = isdf_block_action(LR, 6, True, W, (inst.X, D), tda=True,
                                        spin=spin, precision='fp64')
        Az, Bz = apply_AB(z)
        assert Bz is None
        ref = O.isdf_block_action(inst.eps, 6, inst.X, D, W, kappa=kappa)[0](z)
        assert rel(Az, ref) < 1e-11
        assert np.allclose(d, O.pair_diagonal(inst.eps, 6))


# Split BF16: the error is measured on the kernel part K z = A z - d o z
# (d o z is exact in every mode).  hi+lo carries ~16 bits per operand, so
# |dKz| / |Kz| ~ 1e-5; plain BF16 (8 bits) would sit near 4e-3.
KTOL = {'fp64': 1e-12, 'fp32': 2e-6, 'bf16x3': 6e-5}


@pytest.mark.parametrize('precision', ['fp64', 'fp32', 'bf16x3'])
@pytest.mark.parametrize('shape', [(1, 3, 10), (5, 19, 120), (21, 93, 1200),
                                   (33, 20, 65), (60, 70, 257), (100, 40, 200),
                                   (128, 10, 70)])
def test_kernel_part_relative_error(shape, precision):
    nocc, nvir, M = shape
    print(f'sigma kernel part {shape} {precision}', flush=True)
    inst = synth.make_instance(nocc, nvir, M, seed=[7, nocc, M])
    W_grid, D, W_aux = grid_inputs(inst, screened=False)
    z = np.random.default_rng(nocc + 1).normal(size=(6, nocc, nvir))
    ref = O.isdf_block_action(inst.eps, nocc, inst.X, D, W_aux)[0](z)
    dz = O.pair_diagonal(inst.eps, nocc)[None] * z
    op = device_action(inst, W_grid, precision)
    got = apply_np(op, z)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref - dz)
    assert err < KTOL[precision], err


def test_bf16x3_golden_and_screened_diagonal():
    gold = np.load(GOLDEN)
    inst = synth.config(0)
    op = device_action(inst, gold['W_grid'], 'bf16x3')
    got = apply_np(op, gold['z'])
    dz = O.pair_diagonal(inst.eps, inst.nocc)[None] * gold['z']
    assert np.linalg.norm(got - gold['Az']) / np.linalg.norm(gold['Az'] - dz) < KTOL['bf16x3']
    sd = op.screened_diagonal_device().cpu().numpy()
    d = O.pair_diagonal(inst.eps, inst.nocc)
    assert np.abs(sd - gold['sdiag']).max() < 1e-6 * np.abs(gold['sdiag'] - d).max() + 1e-12


@pytest.mark.parametrize('precision', ['fp64', 'bf16x3'])
def test_stage_timing_reads_every_timed_apply(precision):
    """Stage timing records events without waiting: applies queued back to
    back are all counted when read, a pause keeps what was timed and leaves
    out what runs meanwhile, and the results match untimed applies."""
    inst = synth.config(1)
    print(f'sigma stage timing cfg1 {precision}', flush=True)
    W, _, _ = grid_inputs(inst)
    op = device_action(inst, W, precision)
    rng = np.random.default_rng(5)
    zs = [torch.as_tensor(rng.normal(size=(k, inst.nocc, inst.nvir)), device='cuda')
          for k in (2, 3, 1, 4)]
    ref = [op.apply(z) for z in zs]
    op.profile(True)
    got = [op.apply(z) for z in zs[:3]]
    op.profile_pause()
    op.apply(zs[3])
    op.profile_resume()
    got.append(op.apply(zs[3]))
    pr = op.profile_read()
    op.profile(False)
    torch.cuda.synchronize()
    for g, r in zip(got, ref):
        assert float((g - r).abs().max() / r.abs().max()) < 1e-12
    assert pr['launches'] == 4
    flops = sum(k * (4.0 * inst.nocc * inst.M ** 2 + inst.M ** 2) for k in (2, 3, 1, 4))
    assert pr['exchange_flops'] == pytest.approx(flops, rel=1e-12)
    assert all(v > 0 for v in pr['ms'].values())
    # many timed applies queued at once: the pending events are collected
    # as they fill up and nothing is lost
    op.profile(True)
    for _ in range(600):
        op.apply(zs[2])
    pr = op.profile_read()
    op.profile(False)
    assert pr['launches'] == 600
