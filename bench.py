"""Benchmark: Tamm-Dancoff BSE Davidson roots/s within 1 meV of FP64, and the
chi0 build's MFMA fraction, on one MI355X.

A step is one complete Davidson solve of the benzene/def2-SVP-size instance
(21 occupied, 93 virtual, 1200 grid points, 10 lowest singlet roots,
residual tol 1e-5) on device-resident inputs: sigma-vectors on the device
kernels, the FP64 subspace work on the device, the FP64 residual recheck of
the converged roots included.  The static W it needs (16 minimax time
points, FP64 Cholesky Dyson step) is built once before the timed region and
timed apart.

Printed: one JSON line (rank 0).  `value` is roots/s of the headline solver
(mixed precision: FP32 MFMA sigma + FP64 subspace + FP64 residual recheck);
`plain_fp64` is the same solve with FP64 sigma-vectors.  `roofline` is the
exchange kernel (the sigma-vector's dominant launch) timed with HIP events
inside the timed steps; `chi0_roofline` the fused chi0 kernel on the large
single instance (60 occupied, 400 virtual, 4000 points, 20 time points).
`cpu_baseline` is the FP64 NumPy oracle (the reference algorithm restated,
one trial vector at a time) solving the same instance on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HA_MEV = 27211.386245988
PEAK_TFLOPS = {'fp64': 78.6, 'fp32': 157.3, 'bf16x3': 2500.0}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=5)
    p.add_argument('--warmup', type=int, default=2)
    p.add_argument('--precision', default='fp32', choices=['fp32', 'fp64'])
    p.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline')
    p.add_argument('--no-chi0', action='store_true', help='skip the large chi0 leg')
    return p.parse_args()


def log(msg):
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def solve(apply_AB, d, inst, refine_apply=None, timings=None):
    from mbptcode_amd.davidson import run_davidson
    return run_davidson(apply_AB, d, inst.nroots, inst.tol, 200, tda=True,
                        preconditioner='screened', refine_apply=refine_apply,
                        timings=timings)


def cpu_baseline(inst, W_grid, e_ref):
    """The oracle's FP64 solve of the same instance on the host cores."""
    from oracle import isdf_oracle as O
    try:
        from threadpoolctl import threadpool_info
        blas = [(i.get('internal_api'), i.get('num_threads')) for i in threadpool_info()
                if i.get('user_api') == 'blas']
    except Exception:
        blas = []
    cores = len(os.sched_getaffinity(0))
    D = O.grid_factors(inst.V)
    W_aux = np.linalg.solve(D, np.linalg.solve(D, W_grid).T).T
    W_aux = 0.5 * (W_aux + W_aux.T)
    apply_A, d = O.isdf_block_action(inst.eps, inst.nocc, inst.X, D, W_aux)
    no, nv = inst.nocc, inst.nvir

    def mm(V):
        return apply_A(V.T.reshape(-1, no, nv)).reshape(V.shape[1], -1).T

    t0 = time.perf_counter()
    sd = apply_A.screened_diagonal()
    e, X, conv, napplied = O.davidson_tda(mm, d, inst.nroots, tol=inst.tol,
                                          precond_diag=sd)
    dt = time.perf_counter() - t0
    good = int(np.sum(np.abs(e - e_ref) * HA_MEV <= 1.0))
    return {'value': good / dt, 'unit': 'roots/s', 'cores': cores, 'kind': 'port',
            'sample': (f'one FP64 Davidson solve of the same instance '
                       f'({napplied} sigma-vectors, one at a time as the '
                       f'reference does), {dt:.2f} s, {good}/{inst.nroots} roots '
                       f'within 1 meV'),
            'blas': blas, 'seconds': dt}


def chi0_leg(reps=3):
    """MFMA fraction of the fused chi0 kernel on the large single instance."""
    import torch
    from mbptcode_amd import grids, synth
    from mbptcode_amd.space_time import Chi0Builder
    inst = synth.config(2)
    g = grids.static_screening_grid(inst.eps, inst.nocc, ntau=inst.ntau)
    out = {}
    for precision in ('fp32',):
        b = Chi0Builder(inst.X, inst.eps, inst.nocc, precision=precision)
        chi = b.build(g.tau_points, g.cosft_wt[0])      # warm
        torch.cuda.synchronize()
        b.profile(True)
        for _ in range(reps):
            b.build(g.tau_points, g.cosft_wt[0], out=chi)
        torch.cuda.synchronize()
        pr = b.profile_read()
        b.profile(False)
        avg_ms = pr['ms'] / pr['launches']
        achieved = pr['flops'] / pr['launches'] / (avg_ms * 1e-3) / 1e12
        M, no, nv = inst.M, inst.nocc, inst.nvir
        out = {'kernel': 'k_chi0<float>', 'bound': 'mfma', 'achieved': achieved,
               'peak': PEAK_TFLOPS[precision], 'unit': 'TFLOP/s',
               'frac': achieved / PEAK_TFLOPS[precision], 'traffic': None,
               'avg_ms': avg_ms, 'launches': pr['launches'],
               'workload': f'static chi0(0), M={M}, nocc={no}, nvir={nv}, '
                           f'ntau={inst.ntau}',
               'flops_per_launch': pr['flops'] / pr['launches'],
               'flop_definition': 'ntau * M(M+1)(nocc+nvir+1): upper triangle'}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU: this build has no multi-GPU path '
              f'(asked for --gpus {args.gpus}).', file=sys.stderr)
        sys.exit(2)
    import torch
    from mbptcode_amd import grids, synth
    from mbptcode_amd.block_actions import grid_block_action
    from mbptcode_amd.space_time import screened_interaction_grid

    torch.cuda.set_device(0)
    inst = synth.config(1)
    log(f'instance nocc={inst.nocc} nvir={inst.nvir} M={inst.M} roots={inst.nroots}')
    g = grids.static_screening_grid(inst.eps, inst.nocc, ntau=inst.ntau)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    W_grid, _ = screened_interaction_grid(inst.X, inst.V, inst.eps, inst.nocc, g,
                                          precision='fp64')
    torch.cuda.synchronize()
    w_build_s = time.perf_counter() - t0
    log(f'W built in {w_build_s:.3f} s')

    act64, d = grid_block_action(inst.eps, inst.nocc, inst.X, W_grid, inst.V,
                                 precision='fp64')
    # FP64 reference eigenvalues: the FP64 device solve at a tight residual
    from mbptcode_amd.davidson import run_davidson
    e_ref, _, _ = run_davidson(act64, d, inst.nroots, 1e-8, 300, tda=True)
    log(f'reference roots (eV): {np.round(e_ref * 27.211386, 4).tolist()}')

    def timed(apply_AB, refine, steps, warmup, profile_op=None):
        for _ in range(warmup):
            solve(apply_AB, d, inst, refine)
        torch.cuda.synchronize()
        if profile_op is not None:
            profile_op.profile(True)
        good, errs, tims = 0, [], {}
        t0 = time.perf_counter()
        for s in range(steps):
            w, X, _ = solve(apply_AB, d, inst, refine, timings=tims)
            err = np.abs(w - e_ref) * HA_MEV
            good += int(np.sum(err <= 1.0))
            errs.append(err)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        prof = profile_op.profile_read() if profile_op is not None else None
        if profile_op is not None:
            profile_op.profile(False)
        return good, dt, np.max(errs, axis=0), tims, prof

    act, _ = grid_block_action(inst.eps, inst.nocc, inst.X, W_grid, inst.V,
                               precision=args.precision)
    log(f'headline solver: {args.precision} sigma + FP64 subspace/recheck')
    good, dt, err, tims, prof = timed(act, act64 if args.precision != 'fp64' else None,
                                      args.steps, args.warmup, act.operator)
    roots_per_s = good / dt
    log(f'{args.precision}: {roots_per_s:.1f} roots/s, max err {err.max():.2e} meV, '
        f'{tims.get("davidson_iterations")} cycles')
    good64, dt64, err64, tims64, _ = timed(act64, None, args.steps, args.warmup)
    plain = good64 / dt64
    log(f'plain fp64: {plain:.1f} roots/s, {tims64.get("davidson_iterations")} cycles')

    ex_ms = prof['ms']['exchange'] / prof['launches']
    ex_flops = prof['exchange_flops'] / prof['launches']
    achieved = ex_flops / (ex_ms * 1e-3) / 1e12
    roofline = {'bound': 'mfma', 'achieved': achieved, 'peak': PEAK_TFLOPS[args.precision],
                'unit': 'TFLOP/s', 'frac': achieved / PEAK_TFLOPS[args.precision],
                'traffic': None, 'kernel': f'k_exchange<{"float" if args.precision == "fp32" else "double"}>',
                'avg_ms': ex_ms, 'launches': prof['launches'],
                'flops_per_launch': ex_flops,
                'flop_definition': 'k (4 nocc M^2 + M^2) per launch, k = vectors in the batch',
                'stage_ms_per_apply': {k: v / prof['launches'] for k, v in prof['ms'].items()}}

    result = {
        'metric': 'BSE Davidson roots/s at <=1 meV vs FP64 (learned+mixed vs plain); chi0 MFMA %',
        'value': roots_per_s, 'unit': 'roots/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': dt / args.steps * 1e3,
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': args.precision, 'data': 'synthetic',
        'config': {'workload': 'benzene/def2-SVP-size ISDF BSE-TDA singlet Davidson',
                   'nocc': inst.nocc, 'nvir': inst.nvir, 'M': inst.M,
                   'nroots': inst.nroots, 'tol': inst.tol, 'ntau_W': inst.ntau,
                   'preconditioner': 'screened', 'seed': 0},
        'max_err_meV': float(err.max()),
        'plain_fp64': {'roots_per_s': plain, 'ms_per_step': dt64 / args.steps * 1e3,
                       'max_err_meV': float(err64.max()),
                       'cycles': tims64.get('davidson_iterations')},
        'speedup_vs_plain_fp64': roots_per_s / plain if plain else None,
        'cycles': tims.get('davidson_iterations'),
        'W_build_s': w_build_s,
        'roofline': roofline,
    }
    if not args.no_chi0:
        log('chi0 leg (M=4000, 20 time points)')
        result['chi0_roofline'] = chi0_leg()
        log(f'chi0: {result["chi0_roofline"]["achieved"]:.1f} TFLOP/s')
    if not args.no_cpu:
        log('CPU baseline (FP64 oracle)')
        result['cpu_baseline'] = cpu_baseline(inst, W_grid.cpu().numpy(), e_ref)
        log(f'cpu: {result["cpu_baseline"]["value"]:.2f} roots/s')
    print(json.dumps(result), flush=True)


if __name__ == '__main__':
    main()
