"""The full Casida action's B block on the GPU, (A z, B z) against the FP64
oracle (block_actions.py:563-661 with tda=False; the oracle's B block is
pinned to the dense B by tests/test_oracle.py)."""
import numpy as np
import pytest
import torch

from mbptcode_amd import synth
from oracle import isdf_oracle as O

from test_gpu_sigma import SHAPES, device_action, grid_inputs

pytestmark = pytest.mark.gpu

#: relative error of B z (no diagonal term) against the FP64 oracle
BTOL = {'fp64': 1e-11, 'fp32': 2e-5, 'bf16x3': 2e-5}


def apply_ab_np(op, z):
    zt = torch.as_tensor(np.ascontiguousarray(z), device='cuda')
    a, b = op.apply_ab(zt)
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


def nrel(got, ref):
    return float(np.linalg.norm(got - ref) / np.linalg.norm(ref))


@pytest.mark.parametrize('precision', ['fp64', 'fp32', 'bf16x3'])
@pytest.mark.parametrize('shape', SHAPES)
def test_b_block_against_oracle(shape, precision):
    nocc, nvir, M = shape
    print(f'sigma B block {shape} {precision}', flush=True)
    inst = synth.make_instance(nocc, nvir, M, seed=[5, nocc, M])
    W_grid, D, W_aux = grid_inputs(inst, screened=False)
    z = np.random.default_rng(nocc + 2).normal(size=(5, nocc, nvir))
    Az_ref, Bz_ref = O.isdf_block_action(inst.eps, nocc, inst.X, D, W_aux, tda=False)[0](z)
    op = device_action(inst, W_grid, precision)
    Az, Bz = apply_ab_np(op, z)
    assert nrel(Bz, Bz_ref) < BTOL[precision], nrel(Bz, Bz_ref)
    # the A half is the Tamm-Dancoff action, unchanged by the B request
    Az_alone = op.apply(torch.as_tensor(z, device='cuda')).cpu().numpy()
    assert np.array_equal(Az, Az_alone)


@pytest.mark.parametrize('kappa', [2.0, 0.0])
def test_b_block_symmetric_and_triplet(kappa):
    """y^T B z = z^T B y (B symmetric), and kappa = 0 (triplet) drops the
    Hartree part from B as from A."""
    inst = synth.make_instance(21, 93, 1200, seed=4)
    W_grid, D, W_aux = grid_inputs(inst)
    print(f'sigma B block symmetry kappa={kappa}', flush=True)
    op = device_action(inst, W_grid, 'fp64', kappa=kappa)
    rng = np.random.default_rng(9)
    y, z = rng.normal(size=(2, 1, 21, 93))
    _, By = apply_ab_np(op, y)
    _, Bz = apply_ab_np(op, z)
    lhs, rhs = float(np.vdot(y, Bz)), float(np.vdot(z, By))
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), 1e-300)
    Bz_ref = O.isdf_block_action(inst.eps, 21, inst.X, D, W_aux, kappa=kappa, tda=False)[0](z)[1]
    assert nrel(Bz, Bz_ref) < 1e-11


def test_reference_signature_full_action():
    """isdf_block_action(tda=False): apply_AB(z) -> (A z, B z), NumPy in and
    out, factor-gauge W_aux, against the oracle on the same arguments."""
    from mbptcode_amd.block_actions import isdf_block_action
    inst = synth.make_instance(6, 14, 90, seed=3)
    rng = np.random.default_rng(3)
    D = 0.3 * rng.normal(size=(90, 11))
    W = rng.normal(size=(11, 11))
    W = 0.5 * (W + W.T) + 11 * np.eye(11)

    class LR:
        eps = inst.eps
        spin_mode = 'restricted'

    apply_AB, d = isdf_block_action(LR, 6, True, W, (inst.X, D), tda=False, precision='fp64')
    z = rng.normal(size=(4, 6, 14))
    Az, Bz = apply_AB(z)
    Az_ref, Bz_ref = O.isdf_block_action(inst.eps, 6, inst.X, D, W, tda=False)[0](z)
    assert isinstance(Bz, np.ndarray) and Bz.shape == z.shape
    assert nrel(Az, Az_ref) < 1e-11 and nrel(Bz, Bz_ref) < 1e-11
