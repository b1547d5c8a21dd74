"""The learned preconditioner's BF16 kernels (libmbpt_hip.so, mlp.hip) against
the FP64 MLP oracle: forward, gradient of every parameter block, the
Davidson-correction mode, training, and the Davidson fallback.

Tolerances: BF16 activations and W2 carry 8 mantissa bits, so outputs and
gradients are compared by relative 2-norm error (3e-2 forward, 6e-2
gradients) and by cosine similarity (> 0.999)."""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def net_and_data(N=777, seed=0):
    from mbptcode_amd.mlp_precond import LearnedPreconditioner, init_params, unflatten
    rng = np.random.default_rng(seed)
    flat = init_params(seed)
    flat[3 * 256:4 * 256] = rng.normal(0, 0.1, 256)          # b1
    net = LearnedPreconditioner(params=flat)
    X = rng.normal(size=(N, 3)).astype(np.float32)
    T = rng.normal(size=N).astype(np.float32)
    p = {k: np.asarray(v, float) for k, v in unflatten(flat).items()}
    return net, p, X, T


def relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def cosine(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))


@pytest.mark.parametrize('N', [1, 15, 777, 20000])
def test_forward_matches_oracle(N):
    net, p, X, T = net_and_data(N)
    print(f'mlp forward N={N}', flush=True)
    y = net.forward(torch.as_tensor(X, device='cuda')).cpu().numpy()
    ref, _ = M.forward(p, X.astype(float))
    assert relerr(y, ref) < 3e-2
    if N > 10:
        assert cosine(y, ref) > 0.999


@pytest.mark.parametrize('N', [64, 5000])
def test_gradient_matches_oracle(N):
    from mbptcode_amd.mlp_precond import unflatten
    net, p, X, T = net_and_data(N, seed=1)
    print(f'mlp gradient N={N}', flush=True)
    loss, g = net.gradient(torch.as_tensor(X, device='cuda'), torch.as_tensor(T, device='cuda'))
    ref_loss, ref = M.backward(p, X.astype(float), T.astype(float))
    assert abs(loss - ref_loss) < 3e-2 * ref_loss
    got = unflatten(g.cpu().numpy())
    for name in ('W1', 'b1', 'W2', 'b2', 'w3', 'b3'):
        a, b = np.asarray(got[name], float), np.asarray(ref[name], float)
        if name == 'b3':
            assert abs(float(a) - float(b)) < 6e-2 * abs(float(b)) + 1e-6, name
            continue
        assert relerr(a, b) < 6e-2, (name, relerr(a, b))
        assert cosine(a, b) > 0.998, (name, cosine(a, b))


def test_correction_mode_matches_oracle():
    from mbptcode_amd.mlp_precond import LearnedPreconditioner, unflatten
    net, p, _, _ = net_and_data(10)
    rng = np.random.default_rng(4)
    n, k = 900, 5
    R = rng.normal(size=(n, k)) * 1e-3
    d = rng.uniform(0.3, 3.0, n)
    th = np.sort(rng.uniform(0.3, 0.6, k))
    lp = LearnedPreconditioner(params=net.params(), d=d)
    T = lp.correction(torch.as_tensor(R, device='cuda'), None,
                      torch.as_tensor(th, device='cuda')).cpu().numpy()
    ref = M.correction(p, R, d, th)
    assert relerr(T, ref) < 3e-2 and cosine(T, ref) > 0.999


def test_training_fits_a_smooth_target():
    """Adam on the device kernels learns t = r / (d - theta + 0.5) (a damped
    diagonal correction) to a small fraction of its variance."""
    from mbptcode_amd.mlp_precond import LearnedPreconditioner
    rng = np.random.default_rng(5)
    N = 200000
    X = np.stack([rng.normal(size=N), rng.uniform(0.0, 3.0, N), rng.uniform(0.2, 1.0, N)], 1)
    T = X[:, 0] / (X[:, 1] + 0.5)
    net = LearnedPreconditioner(seed=3)
    Xd = torch.as_tensor(X, dtype=torch.float32, device='cuda')
    Td = torch.as_tensor(T, dtype=torch.float32, device='cuda')
    hist = net.fit(Xd, Td, steps=400, batch=16384, lr=2e-3, lr_final=2e-4)
    print(f'mlp training loss {hist[0]:.3e} -> {hist[-1]:.3e}, var {T.var():.3e}', flush=True)
    assert hist[-1] < 0.05 * T.var()
    assert hist[-1] < 0.1 * hist[0]


def test_davidson_with_learned_correction_and_fallback():
    """An untrained network gives useless corrections: the solve must detect
    the stall, fall back to the screened Jacobi-Davidson correction and still
    reach the dense roots."""
    import os
    from mbptcode_amd import synth
    from mbptcode_amd.block_actions import grid_block_action
    from mbptcode_amd.davidson import run_davidson
    from mbptcode_amd.mlp_precond import LearnedPreconditioner
    gold = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'cfg0_oracle.npz'))
    inst = synth.config(0)
    apply_AB, d = grid_block_action(inst.eps, inst.nocc, inst.X, gold['W_grid'], inst.V,
                                    precision='fp64')
    net = LearnedPreconditioner(seed=11, d=d)
    net.residual_of_jd = False          # the bare (random) network alone
    tims = {}
    w, X, _ = run_davidson(apply_AB, d, inst.nroots, 1e-7, 200, tda=True,
                           preconditioner='learned', learned=net, stall_cycles=2,
                           learned_mode='replace', learned_until=None, timings=tims)
    assert np.abs(w - gold['roots']).max() * 27211.386 < 1e-3
    assert net.calls >= 1
    assert 'learned_fallback_cycle' in tims['davidson_flags']


def test_trained_gain():
    """What the shipped network gains over the method without it -- the
    Jacobi-Davidson correction on the screened diagonal, which it augments --
    on 32 held-out instances of the batched workload (seed family 1; the
    network was trained on family 2) and on the 100/700/6000 instance, FP64
    sigma (and split-BF16 sigma with the FP64 recheck at 100/700/6000).

    Asserted: every learned solve reaches the screened solve's roots
    (<= 1e-3 meV); on the held-out set it takes fewer Davidson cycles than
    the screened correction and than the bare diagonal.  Measured and
    printed, not a gain: it applies sigma to more vectors than the screened
    correction (its augment cycle adds a learned vector beside every
    Jacobi-Davidson one), and at 100/700/6000 it takes as many cycles; the
    counts are checked against the record shipped with the weights, so a
    change of either shows."""
    import json
    import os
    from mbptcode_amd import synth
    from mbptcode_amd.mlp_precond import DEFAULT_WEIGHTS, LearnedPreconditioner
    from mbptcode_amd.train_precond import evaluate, evaluate_single
    net = LearnedPreconditioner.load()
    sizes = synth.batched_sizes()
    held = [synth.batched_instance(i, sizes) for i in range(32)]
    res = evaluate(net.params(), held, residual_of_jd=net.residual_of_jd)
    single = evaluate_single(net.params(), synth.config(4), repeats=1,
                             residual_of_jd=net.residual_of_jd)
    print('[trained gain] held-out ' + json.dumps(res), flush=True)
    print('[trained gain] 100/700/6000 ' + json.dumps(single), flush=True)
    for name, r in list(res.items()) + list(single.items()):
        assert r['max_err_meV'] < 1e-3, name
    L, S, B = res['learned'], res['screened'], res['bare']
    assert L['cycles'] < S['cycles'] and L['cycles'] < B['cycles']
    with open(os.path.splitext(DEFAULT_WEIGHTS)[0] + '_train.json') as f:
        rec = json.load(f)
    for mode in ('bare', 'screened', 'learned'):
        for key in ('cycles', 'vectors'):
            assert abs(res[mode][key] - rec['eval'][mode][key]) <= 0.03 * rec['eval'][mode][key], \
                (mode, key, res[mode][key], rec['eval'][mode][key])
    for mode, r in single.items():
        ref = rec['eval_single']['results'][mode]
        assert (r['cycles'], r['vectors']) == (ref['cycles'], ref['vectors']), (mode, r, ref)
