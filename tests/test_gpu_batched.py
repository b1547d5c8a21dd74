"""Batched path on the GPU: the batched Jacobi eigensolver against LAPACK, the
batched sigma-vectors against the single-instance action (itself pinned to
the oracle) and against the oracle directly, the live-vector and padding
contract, the work-plan invariance, the hipGraph replays, the batched
screened diagonal and learned correction, and whole batched solves against
per-instance FP64 solves.  Tolerances are relative to max|A z|: FP64 1e-11,
FP32 1e-5."""
import numpy as np
import pytest
import torch

from mbptcode_amd import synth
from mbptcode_amd.batched import (BatchedSigma, SigmaGraph, autotune, solve_batched,
                                  syev_batched)
from oracle import isdf_oracle as O

pytestmark = pytest.mark.gpu

TOL = {'fp64': 1e-11, 'fp32': 1e-5}
HA_MEV = 27211.386245988
#: (nocc, nvir, M): every occupied bucket (16, 32, 48, 64), uneven M
SIZES = [(5, 19, 120), (17, 30, 200), (33, 12, 150), (40, 25, 264), (64, 8, 96), (9, 41, 70)]


def random_w(M, seed):
    rng = np.random.default_rng(seed)
    B = rng.normal(size=(M, M)) / M
    return np.eye(M) + 0.5 * (B + B.T)


@pytest.fixture(scope='module')
def batch():
    insts = [synth.make_instance(no, nv, M, [7, i]) for i, (no, nv, M) in enumerate(SIZES)]
    Ws = [random_w(inst.M, 100 + i) for i, inst in enumerate(insts)]
    return insts, Ws


def make_op(insts, Ws, precision, kmax=4):
    return BatchedSigma([i.eps for i in insts], [i.nocc for i in insts], [i.X for i in insts],
                        Ws, [i.V for i in insts], precision=precision, kmax=kmax)


def single_op(inst, W, precision='fp64'):
    from mbptcode_amd.block_actions import SigmaOperator
    return SigmaOperator(inst.eps, inst.nocc, inst.X, W, inst.V, kappa=2.0,
                         precision=precision)


def random_z(op, seed=0):
    rng = np.random.default_rng(seed)
    z = np.zeros((op.n, op.kmax, op.P))
    for b in range(op.n):
        z[b, :, :op.npair[b]] = rng.normal(size=(op.kmax, op.npair[b]))
    return torch.as_tensor(z, device='cuda')


def rel(got, ref):
    return float(np.abs(got - ref).max() / np.abs(ref).max())


@pytest.mark.parametrize('m', [1, 2, 5, 33, 64])
def test_syev_batched_matches_lapack(m):
    print(f'syev_batched m={m}', flush=True)
    rng = np.random.default_rng(m)
    A = rng.normal(size=(37, m, m))
    A[:, np.arange(m), np.arange(m)] += np.linspace(-3, 3, m)
    A[0] = np.diag(np.arange(m, 0, -1.0))          # already diagonal, reversed
    if m > 2:
        A[1, :2, :2] = [[1.0, 0.0], [0.0, 1.0]]     # degenerate pair
    if m > 4:
        # slots shifted far above the rest and decoupled from it (how the
        # batched Davidson parks unused trial-space slots)
        A[2, -2:, :] = 0.0
        A[2, :, -2:] = 0.0
        A[2, -2:, -2:] = np.diag([1e10, 2e10])
    w, V = syev_batched(torch.as_tensor(A, device='cuda'))
    w, V = w.cpu().numpy(), V.cpu().numpy()
    S = 0.5 * (A + A.transpose(0, 2, 1))
    w_ref = np.linalg.eigvalsh(S)
    scale = np.abs(np.where(np.abs(S) < 1e9, S, 0)).max(axis=(1, 2))[:, None]
    assert (np.abs(w - w_ref) < 1e-13 * scale * m).all()
    assert (np.diff(w, axis=1) >= 0).all()
    res = np.einsum('bij,bjk->bik', S, V) - V * w[:, None, :]
    assert (np.abs(res).max(axis=1) < 1e-13 * np.maximum(scale, np.abs(w)) * m).all()
    eye = np.einsum('bji,bjk->bik', V, V)
    assert np.abs(eye - np.eye(m)).max() < 1e-12 * m


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_batched_sigma_matches_single_instance(batch, precision):
    insts, Ws = batch
    print(f'batched sigma vs single {precision}', flush=True)
    op = make_op(insts, Ws, precision)
    z = random_z(op)
    Az = op.apply(z).cpu().numpy()
    zh = z.cpu().numpy()
    for b, inst in enumerate(insts):
        ref_op = single_op(inst, Ws[b])
        np_ = inst.npair
        zb = torch.as_tensor(zh[b, :, :np_].reshape(op.kmax, inst.nocc, inst.nvir), device='cuda')
        ref = ref_op.apply(zb).cpu().numpy().reshape(op.kmax, np_)
        assert rel(Az[b, :, :np_], ref) < TOL[precision], (b, inst.nocc, inst.nvir, inst.M)


def test_batched_sigma_matches_oracle(batch):
    insts, Ws = batch
    print('batched sigma vs oracle', flush=True)
    op = make_op(insts, Ws, 'fp64', kmax=2)
    z = random_z(op, seed=3)
    Az = op.apply(z).cpu().numpy()
    for b in (0, 4):
        inst = insts[b]
        D = O.grid_factors(inst.V)
        W_aux = np.linalg.solve(D, np.linalg.solve(D, Ws[b]).T).T
        zb = z.cpu().numpy()[b, :, :inst.npair].reshape(2, inst.nocc, inst.nvir)
        apply_A, _ = O.isdf_block_action(inst.eps, inst.nocc, inst.X, D, W_aux, kappa=2.0)
        ref = apply_A(zb)
        assert rel(Az[b, :, :inst.npair], ref.reshape(2, -1)) < TOL['fp64']


def test_live_rows_and_padding_untouched(batch):
    insts, Ws = batch
    print('batched sigma kact contract', flush=True)
    # the partial pass runs first on a fresh handle, so no workspace left by a
    # full pass can stand in for rows it skipped
    op = make_op(insts, Ws, 'fp64')
    z = random_z(op, seed=5)
    kact = torch.as_tensor([4, 2, 0, 1, 3, 4], dtype=torch.int32, device='cuda')
    out = torch.full_like(z, float('nan'))
    op.apply_into(z, out, kact)
    got = out.cpu().numpy()
    full = make_op(insts, Ws, 'fp64').apply(z).cpu().numpy()
    for b in range(op.n):
        k, np_ = int(kact[b]), int(op.npair[b])
        assert np.isnan(got[b, k:]).all()
        assert np.isnan(got[b, :, np_:]).all()
        assert rel(got[b, :k, :np_], full[b, :k, :np_]) < 1e-14 if k else True


def test_plans_agree(batch):
    insts, Ws = batch
    print('batched sigma plan invariance', flush=True)
    op = make_op(insts, Ws, 'fp64')
    z = random_z(op, seed=9)
    ref = op.apply(z).cpu().numpy()
    # the last knob is the exchange bucket width: 32 pads n_o = 5, 9, 17 to
    # 32 and 33, 40 to 64 (the K depth of the P product, and which
    # instances share an exchange launch)
    for plan in [(64, 16, 16, 8, 32), (128, 48, 32, 0, 16), (4096, 4096, 4096, 64, 32),
                 (256, 128, 1024, 16, 16), (1024, 512, 512, 0, 32)]:
        op.plan(*plan)
        got = op.apply(z).cpu().numpy()
        for b in range(op.n):
            np_ = op.npair[b]
            assert rel(got[b, :, :np_], ref[b, :, :np_]) < 1e-12, plan


def test_graph_replays_match_eager(batch):
    insts, Ws = batch
    print('batched sigma graph replays', flush=True)
    op = make_op(insts, Ws, 'fp32')
    sg = SigmaGraph(op)
    for rep in range(4):
        z = random_z(op, seed=20 + rep)
        sg.z.copy_(z)
        sg.kact.fill_(op.kmax - (rep % 2))
        got = sg.run().clone()
        torch.cuda.synchronize()
        eager = torch.full_like(z, float('nan'))
        op.apply_into(z, eager, sg.kact)
        torch.cuda.synchronize()
        k = int(sg.kact[0])
        a, e = got[:, :k].cpu().numpy(), eager[:, :k].cpu().numpy()
        mask = ~np.isnan(e)
        assert np.array_equal(a[mask], e[mask]), rep


def test_batched_screened_diagonal(batch):
    insts, Ws = batch
    print('batched screened diagonal', flush=True)
    for precision in ('fp64', 'fp32'):
        op = make_op(insts, Ws, precision)
        sd = op.screened_diagonal().cpu().numpy()
        for b, inst in enumerate(insts):
            ref = single_op(inst, Ws[b]).screened_diagonal_device().cpu().numpy().ravel()
            assert rel(sd[b, :inst.npair], ref) < TOL[precision]
            assert (sd[b, inst.npair:] > 1e5).all()


def test_autotune_picks_a_valid_plan(batch, tmp_path, monkeypatch):
    insts, Ws = batch
    print('autotune', flush=True)
    monkeypatch.setenv('MBPT_TUNE_CACHE', str(tmp_path / 'tune.json'))
    op = make_op(insts, Ws, 'fp64')
    z = random_z(op, seed=4)
    ref = op.apply(z).cpu().numpy()
    space = dict(rows_per_split=(64, 1024), tail_kchunk=(64, 512), hartree_kchunk=(512,),
                 grid_cap=(0, 64))
    plan, info = autotune(op, space=space, reps=2)
    assert info['source'] == 'timed' and (tmp_path / 'tune.json').exists()
    got = op.apply(z).cpu().numpy()
    for b in range(op.n):
        assert rel(got[b, :, :op.npair[b]], ref[b, :, :op.npair[b]]) < 1e-12
    plan2, info2 = autotune(op, space=space)
    assert plan2 == plan and info2['source'].endswith('tune.json')


def test_learned_correction_batched_matches_single(batch):
    from mbptcode_amd.mlp_precond import LearnedPreconditioner
    insts, Ws = batch
    print('batched learned correction', flush=True)
    net = LearnedPreconditioner(seed=3)
    rng = np.random.default_rng(0)
    B, k = len(insts), 3
    P = max(i.npair for i in insts)
    R = np.zeros((B, P, k)); d = np.full((B, P), 1e6); th = rng.uniform(0.2, 1.0, (B, k))
    for b, inst in enumerate(insts):
        R[b, :inst.npair] = rng.normal(size=(inst.npair, k))
        d[b, :inst.npair] = rng.uniform(0.3, 4.0, inst.npair)
    dev = lambda a: torch.as_tensor(a, device='cuda')
    T = net.correction_batched(dev(R), None, dev(th), dev(d),
                               [i.npair for i in insts]).cpu().numpy()
    for b, inst in enumerate(insts):
        n = inst.npair
        ref = net.correction(dev(R[b, :n]), None, dev(th[b]), dev(d[b, :n])).cpu().numpy()
        assert np.allclose(T[b, :n], ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())
        assert (T[b, n:] == 0).all()


def _batch_problem(count, seed=3):
    from mbptcode_amd.grids import static_screening_grid
    from mbptcode_amd.space_time import screened_interaction_grid
    rng = np.random.default_rng(seed)
    insts, Ws = [], []
    for i in range(count):
        no, nv = int(rng.integers(5, 20)), int(rng.integers(15, 60))
        inst = synth.make_instance(no, nv, 4 * (no + nv), [seed, i])
        g = static_screening_grid(inst.eps, inst.nocc, ntau=8)
        W, _ = screened_interaction_grid(inst.X, inst.V, inst.eps, inst.nocc, g,
                                         precision='fp64')
        insts.append(inst)
        Ws.append(W)
    return insts, Ws


@pytest.mark.parametrize('mode', ['fp64', 'fp32+refine'])
def test_solve_batched_matches_per_instance(mode):
    from mbptcode_amd.block_actions import grid_block_action
    from mbptcode_amd.davidson import run_davidson
    print(f'solve_batched {mode}', flush=True)
    insts, Ws = _batch_problem(7)
    nroots, tol = 5, 1e-6
    op64 = make_op(insts, Ws, 'fp64', kmax=10)
    if mode == 'fp64':
        e, X, info = solve_batched(op64, nroots, tol)
    else:
        op32 = make_op(insts, Ws, 'fp32', kmax=10)
        e, X, info = solve_batched(op32, nroots, tol, refine_op=op64)
        assert info['rnorm_fp64_after'] <= tol
    print(info['trace'][-3:], flush=True)
    assert info['converged']
    for b, inst in enumerate(insts):
        act, d = grid_block_action(inst.eps, inst.nocc, inst.X, Ws[b], inst.V, precision='fp64')
        w, _, _ = run_davidson(act, d, nroots, 1e-9, 300, tda=True)
        err = np.abs(e[b] - w).max() * HA_MEV
        assert err < 1e-3, (b, err)
        # eigenvector residual in FP64
        Xb = X[b, :inst.npair].T.contiguous().reshape(nroots, inst.nocc, inst.nvir)
        AX = act.operator.apply(Xb).reshape(nroots, -1).cpu().numpy()
        xb = Xb.reshape(nroots, -1).cpu().numpy()
        rn = np.linalg.norm(AX - xb * e[b][:, None], axis=1)
        assert rn.max() <= 1.01 * tol


def test_batched_subspace_products_match_bmm():
    from mbptcode_amd.batched import bnn, btn
    print('batched subspace products', flush=True)
    rng = np.random.default_rng(8)
    B, P, m, c = 5, 1300, 37, 9
    rows_h = np.array([1300, 700, 0, 1, 513], np.int32)
    rows = torch.as_tensor(rows_h, device='cuda')
    A = torch.zeros((B, P, 64), dtype=torch.float64, device='cuda')
    Bm = torch.zeros((B, P, c), dtype=torch.float64, device='cuda')
    for b in range(B):
        A[b, :rows_h[b], :m] = torch.as_tensor(rng.normal(size=(rows_h[b], m)), device='cuda')
        Bm[b, :rows_h[b]] = torch.as_tensor(rng.normal(size=(rows_h[b], c)), device='cuda')
    Av = A[:, :, :m]                                   # strided view, row stride 64
    ref_tn = torch.bmm(Av.transpose(1, 2).contiguous(), Bm)
    H = torch.full((B, 50, 50), 7.0, dtype=torch.float64, device='cuda')
    btn(Av, Bm, rows, out=H[:, 3:3 + m, 5:5 + c])
    assert torch.allclose(H[:, 3:3 + m, 5:5 + c], ref_tn, rtol=1e-12, atol=1e-10)
    assert (H[:, :3] == 7.0).all() and (H[:, :, :5] == 7.0).all()
    S = torch.as_tensor(rng.normal(size=(B, m, c)), device='cuda')
    ref_nn = torch.bmm(Av.contiguous(), S)
    Y = bnn(Av, S, rows)
    assert torch.allclose(Y, ref_nn, rtol=1e-12, atol=1e-10)
    Z = Bm.clone()
    bnn(Av, S, rows, out=Z, subtract=True)
    assert torch.allclose(Z, Bm - ref_nn, rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize('m,c', [(64, 70), (4, 1), (17, 33)])
def test_batched_nn_product_shapes(m, c):
    """NN over column groups past 64, the narrowest product and ragged tiles;
    padding rows zeroed by assignment and left alone by subtraction."""
    from mbptcode_amd.batched import bnn
    print(f'batched nn m={m} c={c}', flush=True)
    rng = np.random.default_rng(m + c)
    B, P = 3, 300
    rows_h = np.array([300, 65, 1], np.int32)
    rows = torch.as_tensor(rows_h, device='cuda')
    A = torch.zeros((B, P, m), dtype=torch.float64, device='cuda')
    for b in range(B):
        A[b, :rows_h[b]] = torch.as_tensor(rng.normal(size=(rows_h[b], m)), device='cuda')
    S = torch.as_tensor(rng.normal(size=(B, m, c)), device='cuda')
    ref = torch.bmm(A, S)
    Y = torch.full((B, P, c), 5.0, dtype=torch.float64, device='cuda')
    bnn(A, S, rows, out=Y)
    assert torch.allclose(Y, ref, rtol=1e-12, atol=1e-10)
    Z = torch.full((B, P, c), 2.0, dtype=torch.float64, device='cuda')
    bnn(A, S, rows, out=Z, subtract=True)
    for b in range(B):
        n = rows_h[b]
        assert torch.allclose(Z[b, :n], 2.0 - ref[b, :n], rtol=1e-12, atol=1e-10)
        assert (Z[b, n:] == 2.0).all()
