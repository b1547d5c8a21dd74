"""The full Casida Davidson on the GPU (run_davidson with tda=False: the
paired loop over the device (A z, B z) action) against the closed-form dense
Casida roots of the oracle (LR/davidson.py:998 `_davidson_root`; the roots,
normalisation and residuals are judged, not the iteration)."""
import numpy as np
import pytest
import torch

from mbptcode_amd import grids, synth
from oracle import isdf_oracle as O

pytestmark = pytest.mark.gpu

HA_MEV = 27211.386245988


def screened(inst):
    D = O.grid_factors(inst.V)
    g = grids.static_screening_grid(inst.eps, inst.nocc, ntau=inst.ntau)
    W_aux = O.static_screening_matrix(inst.X, D, inst.eps, inst.nocc, g.tau_points,
                                      g.cosft_wt[0])
    return D, W_aux, O.grid_screened_interaction(W_aux, D)


def dense(inst, D, W_aux):
    A = O.dense_tda_matrix(inst.eps, inst.nocc, inst.X, D, W_aux)
    B = O.dense_casida_b_matrix(inst.nocc, inst.X, D, W_aux)
    return A, B


def solve(inst, W_grid, precision, tol, preconditioner='screened', refine=True,
          timings=None, **kw):
    from mbptcode_amd.block_actions import grid_block_action
    from mbptcode_amd.davidson import run_davidson
    apply_AB, d = grid_block_action(inst.eps, inst.nocc, inst.X, W_grid, inst.V,
                                    precision=precision, tda=False)
    ref64 = None
    if refine and precision != 'fp64':
        ref64 = grid_block_action(inst.eps, inst.nocc, inst.X, W_grid, inst.V,
                                  precision='fp64', tda=False)[0]
    return run_davidson(apply_AB, d, inst.nroots, tol, 200, tda=False,
                        preconditioner=preconditioner, refine_apply=ref64,
                        timings=timings, **kw)


def check_pairs(A, B, w, X, Y, tol):
    """A X + B Y = w X, B X + A Y = -w Y within tol per root, and
    X^T X - Y^T Y = 1."""
    r = np.sqrt(np.sum((A @ X + B @ Y - X * w) ** 2, 0) + np.sum((B @ X + A @ Y + Y * w) ** 2, 0))
    assert r.max() <= tol, r
    assert np.allclose(np.sum(X * X, 0) - np.sum(Y * Y, 0), 1.0, atol=1e-10)


@pytest.mark.parametrize('preconditioner', ['bare', 'screened'])
def test_fp64_roots_against_dense_casida(preconditioner):
    inst = synth.config(0)
    D, W_aux, W_grid = screened(inst)
    A, B = dense(inst, D, W_aux)
    ref = O.dense_casida_roots(A, B)[0][:inst.nroots]
    print(f'casida davidson cfg0 fp64 {preconditioner}', flush=True)
    w, X, Y = solve(inst, W_grid, 'fp64', 1e-8, preconditioner)
    assert np.abs(w - ref).max() * HA_MEV < 1e-6 * 1e3   # 1e-6 eV
    check_pairs(A, B, w, X, Y, 1e-8)
    # the de-excitation part is there: B couples X to Y
    assert np.abs(Y).max() > 1e-6


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_mixed_with_fp64_recheck_meets_one_mev(precision):
    inst = synth.config(1)
    D, W_aux, W_grid = screened(inst)
    A, B = dense(inst, D, W_aux)
    ref = O.dense_casida_roots(A, B)[0][:inst.nroots]
    timings = {}
    print(f'casida davidson cfg1 {precision}', flush=True)
    w, X, Y = solve(inst, W_grid, precision, inst.tol, timings=timings)
    err = np.abs(w - ref).max() * HA_MEV
    print(f'casida davidson cfg1 {precision} err {err:.2e} meV', flush=True)
    assert err < 1.0
    check_pairs(A, B, w, X, Y, inst.tol * (1 + 1e-6))
    assert max(timings['davidson_residuals']) <= inst.tol
    # below the Tamm-Dancoff roots of the same A (by 0.04 to 0.24 mHa here)
    assert np.all(w <= np.linalg.eigvalsh(A)[:inst.nroots] - 1e-5)


def test_large_instance_residuals_and_warm_start():
    """At 24000 pairs (no dense solve): FP64 residuals of the fp32 solve on
    the FP64 action, the normalisation, and a warm start from the roots."""
    from mbptcode_amd.block_actions import grid_block_action
    inst = synth.config(2)
    _, _, W_grid = screened(inst)
    print('casida davidson cfg2 fp32', flush=True)
    tracked = {}
    w, X, Y = solve(inst, W_grid, 'fp32', inst.tol, tracked=tracked)
    op = grid_block_action(inst.eps, inst.nocc, inst.X, W_grid, inst.V,
                           precision='fp64', tda=False)[0].operator
    shape = (-1, inst.nocc, inst.nvir)
    Xt = torch.as_tensor(X.T.reshape(shape), device='cuda')
    Yt = torch.as_tensor(Y.T.reshape(shape), device='cuda')
    AX, BX = op.apply_ab(Xt)
    AY, BY = op.apply_ab(Yt)
    wt = torch.as_tensor(w, device='cuda')[:, None, None]
    rn = torch.sqrt(((AX + BY - Xt * wt) ** 2).sum((1, 2))
                    + ((BX + AY + Yt * wt) ** 2).sum((1, 2))).cpu().numpy()
    assert rn.max() <= inst.tol, rn
    assert np.allclose(np.sum(X * X, 0) - np.sum(Y * Y, 0), 1.0, atol=1e-8)
    timings = {}
    print('casida davidson cfg2 fp32 warm start', flush=True)
    w2, _, _ = solve(inst, W_grid, 'fp32', inst.tol, guess=(tracked['X'], tracked['Y']),
                     timings=timings)
    assert timings['davidson_iterations'] <= 3
    assert np.abs(w2 - w).max() * HA_MEV < 1e-2


def test_learned_preconditioner_is_refused():
    inst = synth.config(0)
    _, _, W_grid = screened(inst)
    with pytest.raises(NotImplementedError):
        solve(inst, W_grid, 'fp64', 1e-6, preconditioner='learned')


class LR:
    """The two fields of the reference's solver object the entry reads."""

    def __init__(self, eps):
        self.eps = eps
        self.spin_mode = 'restricted'


def test_amb_probe_on_the_device_action():
    """min eig(A - B) of the FP64 device action against eigvalsh(A - B): the
    converged probe (k = 2), its spin invariance (kappa cancels), and the
    sign certificate from the Casida roots (probe.py:23-228)."""
    from mbptcode_amd.block_actions import grid_block_action
    from mbptcode_amd.probe import lowest_amb_from_action
    inst = synth.config(1)
    D, W_aux, W_grid = screened(inst)
    A, B = dense(inst, D, W_aux)
    ref = np.linalg.eigvalsh(A - B)[:2]
    print('amb probe cfg1 k=2', flush=True)
    vals = {}
    for kappa in (2.0, 0.0):
        act, d = grid_block_action(inst.eps, inst.nocc, inst.X, W_grid, inst.V, kappa=kappa,
                                   precision='fp64', tda=False)
        vals[kappa] = lowest_amb_from_action(act, d, k=2)
        assert np.abs(vals[kappa] - ref).max() < 1e-9, (vals[kappa], ref)
    w, X, Y = solve(inst, W_grid, 'fp64', 1e-6)
    act, d = grid_block_action(inst.eps, inst.nocc, inst.X, W_grid, inst.V,
                               precision='fp64', tda=False)
    stats = {}
    theta = lowest_amb_from_action(act, d, sign_only=True, start=X - Y, stats=stats)[0]
    assert stats['probe_sign_proven'] and theta - stats['probe_residual'] > 0
    assert theta >= ref[0] - 1e-12


@pytest.mark.parametrize('tda, probe', [(True, False), (False, 'sign'), (False, True)])
def test_solve_casida_davidson_entry(tda, probe):
    """The reference's entry (LR/davidson.py:252) with factor-gauge W_aux
    and isdf_factors (X, D), fp32 with the FP64 recheck, against the dense
    roots; probe= runs the (A - B) probe after the full solve."""
    from mbptcode_amd.davidson import solve_casida_davidson
    inst = synth.config(1)
    D, W_aux, _ = screened(inst)
    A, B = dense(inst, D, W_aux)
    ref = (np.linalg.eigvalsh(A) if tda else O.dense_casida_roots(A, B)[0])[:5]
    stats = {}
    print(f'solve_casida_davidson cfg1 tda={tda} probe={probe}', flush=True)
    w, X, Y = solve_casida_davidson(LR(inst.eps), inst.nocc, polarizability='BSE',
                                    W_aux=W_aux, isdf_factors=(inst.X, D), tda=tda,
                                    probe=probe, nroots=5, conv_tol=1e-6, stats=stats)
    assert np.abs(w - ref).max() * HA_MEV < 1e-3
    if tda:
        assert not Y.any() and 'min_eig_amb' not in stats
    else:
        amb = np.linalg.eigvalsh(A - B)[0]
        assert stats['min_eig_amb'] > 0
        if probe is True:
            assert abs(stats['min_eig_amb'] - amb) < 1e-9


@pytest.mark.parametrize('tda', [True, False])
def test_dense_completion_against_the_oracle(tda):
    """`dense_casida`, the finish of a pair space the trial space ran out on
    (LR/davidson.py:1147-1193, :1296-1330): A (and B) built whole from the
    device action on unit vectors, then the dense eigensolve; against the
    oracle's dense roots, with the Casida equations' residuals and
    X^T X - Y^T Y = 1 checked on the returned vectors."""
    from mbptcode_amd.block_actions import grid_block_action
    from mbptcode_amd.davidson import dense_casida
    inst = synth.config(0)
    D, W_aux, W_grid = screened(inst)
    A, B = dense(inst, D, W_aux)
    act, d = grid_block_action(inst.eps, inst.nocc, inst.X, W_grid, inst.V, precision='fp64',
                               tda=tda)
    print(f'dense completion tda={tda}', flush=True)
    w, X, Y, conv, rn = dense_casida(act, d, 6, 1e-8, batch=17, tda=tda)
    ref = np.linalg.eigvalsh(A)[:6] if tda else O.dense_casida_roots(A, B)[0][:6]
    assert np.abs(w - ref).max() < 1e-10 and conv.all()
    Xh, Yh = X.cpu().numpy(), Y.cpu().numpy()
    if tda:
        assert not Yh.any()
        assert np.abs(A @ Xh - Xh * w).max() < 1e-10
    else:
        check_pairs(A, B, w, Xh, Yh, 1e-10)


@pytest.mark.parametrize('precision', ['fp64', 'bf16x3'])
def test_pair_symmetry_sectors_on_the_device(precision):
    """An instance with an inversion symmetry (`synth.inversion_instance`):
    the device loop with the pairs' parity labels (x_sym) reaches the dense
    roots of the oracle, each returned root inside its own sector: its weight
    on the other sector is round-off (FP64) or the split-BF16 action's
    rounding of the cancelling mirror terms (below 1e-4 of the root's
    largest component)."""
    from mbptcode_amd.orbital_spaces import get_occ_virt_indices, pair_symmetry
    inst = synth.inversion_instance(12, 40, 240, seed=5, nroots=6)
    print(f'inversion instance {precision}: 12 x 40 pairs, M = 480', flush=True)
    D, W_aux, W_grid = screened(inst)
    A, B = dense(inst, D, W_aux)
    lab = pair_symmetry(inst.extra['orbsym'], *get_occ_virt_indices(inst.eps, inst.nocc))
    ref, _, _ = O.dense_casida_roots(A, B)
    w, X, Y = solve(inst, W_grid, precision, 1e-7, x_sym=lab)[:3]
    assert np.abs(w - ref[:6]).max() * HA_MEV < (1e-5 if precision == 'fp64' else 1e-3)
    leak = 1e-9 if precision == 'fp64' else 1e-4
    for j in range(6):
        top = np.abs(X[:, j]).max()
        sect = [L for L in (0, 1) if np.abs(X[lab != L, j]).max() < leak * top
                and np.abs(Y[lab != L, j]).max() < leak * top]
        assert len(sect) == 1, j
