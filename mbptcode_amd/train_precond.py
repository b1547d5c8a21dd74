"""Train the learned Davidson preconditioner and compare it with the diagonal ones.

    python -m mbptcode_amd.train_precond [--instances 512] [--steps 3000]
                                         [--eval 32] [--target jd|span]
                                         [--out PATH] [--eval-only]

Training set: plain FP64 Davidson runs (screened Jacobi-Davidson correction,
tol 1e-5) on `--instances` random instances sized like the batched workload
(n_occ ~ U[5, 40], n_vir ~ U[20, 200], M = 8 (n_occ + n_vir), 5 roots),
drawn from seed family 2 so that the batched workload's own instances
(seed family 1) stay unseen.  Every cycle and unconverged root contributes
the features of 256 random pairs and their ideal correction (see
`mlp_precond.harvest`).  The network is fitted with the device kernels (Adam,
65536-row minibatches) and saved with its training record.

Evaluation: the first `--eval` instances of the batched workload, each solved
with the 'bare', 'screened' and 'learned' corrections at FP64, and the
100/700/6000 instance with the screened and learned corrections in FP64 and
split-BF16 sigma; the cycles, sigma-vectors and wall time of each are
written with the training figures to the record beside the weights
(`<out>_train.json`), roots checked against the screened FP64 solve.
`--eval-only` re-evaluates the weights at `--out` and rewrites the record's
evaluation.
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

from . import synth
from .block_actions import grid_block_action
from .davidson import run_davidson
from .grids import static_screening_grid
from .mlp_precond import DEFAULT_WEIGHTS, LearnedPreconditioner, train_on_instances
from .space_time import screened_interaction_grid

HA_MEV = 27211.386245988


def training_instances(count, seed=2):
    sizes = synth.batched_sizes(count, seed=seed)
    return [synth.batched_instance(i, sizes, seed=seed) for i in range(count)]


def evaluate(net_params, instances, precision='fp64', residual_of_jd=True, learned_until=1000.0):
    """Per preconditioner: total cycles, sigma-vectors and seconds over the
    instances, and the worst root error (meV) against the screened solve."""
    modes = ('bare', 'screened', 'learned', 'learned_replace')
    out = {p: {'cycles': 0, 'vectors': 0, 'seconds': 0.0, 'max_err_meV': 0.0,
               'fallbacks': 0} for p in modes}
    for inst in instances:
        g = static_screening_grid(inst.eps, inst.nocc, ntau=inst.ntau)
        W, _ = screened_interaction_grid(inst.X, inst.V, inst.eps, inst.nocc, g,
                                         precision='fp32')
        act, d = grid_block_action(inst.eps, inst.nocc, inst.X, W, inst.V,
                                   precision=precision)
        act.operator.screened_diagonal_device()
        net = LearnedPreconditioner(params=net_params, d=d)
        net.residual_of_jd = residual_of_jd
        ref = None
        for pre in ('screened', 'bare', 'learned', 'learned_replace'):
            tims = {}
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            w, _, _ = run_davidson(act, d, inst.nroots, inst.tol, 200, tda=True,
                                   preconditioner=pre.split('_')[0], learned=net,
                                   timings=tims, learned_until=learned_until,
                                   learned_mode='replace' if pre.endswith('replace')
                                   else 'augment')
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            if ref is None:
                ref = w
            o = out[pre]
            o['cycles'] += tims['davidson_iterations']
            o['vectors'] += tims['davidson_vectors']
            o['seconds'] += dt
            o['max_err_meV'] = max(o['max_err_meV'], float(np.abs(w - ref).max() * HA_MEV))
            o['fallbacks'] += int('learned_fallback_cycle' in tims['davidson_flags'])
    return out


def evaluate_single(net_params, inst, precisions=('fp64', 'bf16x3'), residual_of_jd=True,
                    learned_until=1000.0, repeats=2):
    """The learned and screened corrections on one single instance (W from
    FP32 chi0 terms), per sigma precision: cycles, sigma-vectors, the best of
    `repeats` wall times (after one untimed solve) and the worst root error
    (meV) against the screened FP64 solve.  Mixed precisions recheck their
    roots in FP64, as the headline solver does."""
    g = static_screening_grid(inst.eps, inst.nocc, ntau=inst.ntau)
    W, _ = screened_interaction_grid(inst.X, inst.V, inst.eps, inst.nocc, g,
                                     precision='fp32')
    acts = {}
    for prec in set(precisions) | {'fp64'}:
        act, d = grid_block_action(inst.eps, inst.nocc, inst.X, W, inst.V, precision=prec)
        act.operator.screened_diagonal_device()
        acts[prec] = act
    net = LearnedPreconditioner(params=net_params, d=d)
    net.residual_of_jd = residual_of_jd
    out, ref = {}, None
    for prec in ('fp64',) + tuple(p for p in precisions if p != 'fp64'):
        for pre in ('screened', 'learned'):
            refine = None if prec == 'fp64' else acts['fp64']
            best = None
            for r in range(repeats + 1):
                tims = {}
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                w, _, _ = run_davidson(acts[prec], d, inst.nroots, inst.tol, 200, tda=True,
                                       preconditioner=pre, learned=net, timings=tims,
                                       learned_until=learned_until, refine_apply=refine)
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                if r > 0:
                    best = dt if best is None else min(best, dt)
            if ref is None:
                ref = w
            out[f'{pre}_{prec}'] = {
                'cycles': int(tims['davidson_iterations']),
                'vectors': int(tims['davidson_vectors']), 'seconds': best,
                'max_err_meV': float(np.abs(w - ref).max() * HA_MEV),
                'fallbacks': int('learned_fallback_cycle' in tims['davidson_flags'])}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--instances', type=int, default=512)
    ap.add_argument('--steps', type=int, default=3000)
    ap.add_argument('--eval', type=int, default=32)
    ap.add_argument('--out', default=DEFAULT_WEIGHTS)
    ap.add_argument('--target', default='jd', choices=['jd', 'span'])
    ap.add_argument('--eval-only', action='store_true',
                    help='evaluate the weights at --out (held-out batched instances '
                         'and the 100/700/6000 instance) and rewrite the record\'s '
                         'evaluation')
    args = ap.parse_args(argv)
    torch.cuda.set_device(0)
    rec_path = os.path.splitext(args.out)[0] + '_train.json'
    if args.eval_only:
        net = LearnedPreconditioner.load(args.out)
        with open(rec_path) as f:
            rec = json.load(f)
    else:
        t0 = time.time()
        insts = training_instances(args.instances)
        print(f'[train] {len(insts)} instances generated in {time.time() - t0:.1f} s',
              flush=True)
        net, info = train_on_instances(insts, steps=args.steps, log_every=500,
                                       target=args.target)
        print(f'[train] {json.dumps(info)}', flush=True)
        net.save(args.out, rows=info['rows'], steps=args.steps, instances=args.instances,
                 loss_first=info['loss_first'], loss_last=info['loss_last'])
        rec = {'train': info}
    sizes = synth.batched_sizes()
    evals = [synth.batched_instance(i, sizes) for i in range(args.eval)]
    res = evaluate(net.params(), evals, residual_of_jd=net.residual_of_jd)
    print(f'[eval] held-out: {json.dumps(res)}', flush=True)
    single = evaluate_single(net.params(), synth.config(4), residual_of_jd=net.residual_of_jd)
    print(f'[eval] 100/700/6000: {json.dumps(single)}', flush=True)
    rec.update(eval_instances=args.eval, eval=res,
               eval_single={'instance': '100/700/6000', 'results': single})
    with open(rec_path, 'w') as f:
        json.dump(rec, f, indent=1)
    return 0


if __name__ == '__main__':
    sys.exit(main())
