"""Benchmark: Tamm-Dancoff BSE Davidson roots/s within 1 meV of FP64 (learned
preconditioner + mixed precision against plain FP64), and the chi0 build's
MFMA fraction, on one MI355X.

Headline step: one complete Davidson solve of the learned-preconditioner
instance (100 occupied, 700 virtual, 6000 grid points, 20 lowest singlet
roots, residual tol 1e-5) on device-resident inputs:

* learned+mixed (the headline `value`): split-BF16 MFMA sigma-vectors, FP64
  subspace work on the device, the learned MLP correction (trained weights
  shipped in mbptcode_amd/data) with the screened Jacobi-Davidson fallback,
  and the FP64 recheck of every requested root (A x - theta x with FP64
  sigma-vectors, iterated in FP64 until the FP64 residual meets the tolerance);
* plain: the same Davidson with FP64 sigma-vectors and the screened
  correction, timed the same way (`plain_fp64`), whose eigenvalues judge the
  errors (`max_err_meV`, `err_meV_per_root`).

The static W (FP64 chi0 at 20 minimax points, FP64 Cholesky Dyson step) is
built once before the timed region and timed apart.  `roofline` is the
split-BF16 exchange kernel, the dominant launch of the headline solve, timed
with HIP events on its stream inside the timed steps.  Further legs:
the benzene-size instance (plain / FP32 / split-BF16), the large instance
(per-root errors of FP32 and split-BF16 against FP64, the RPA correlation
energy error in micro-Hartree, and the fused chi0 kernel's MFMA fraction in
`chi0_roofline`), and the batched workload (256 uneven instances in one
captured batched pass per cycle).  `cpu_baseline` is the FP64 NumPy oracle
(the reference algorithm restated: one trial vector at a time) solving the
headline instance on the host cores, one whole solve timed;
`cpu_baseline_chi0` its proj(tau) at one time point of the headline
instance, scaled to the build's time points.
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
KERNEL_NAMES = {'fp64': 'k_exchange<double>', 'fp32': 'k_exchange<float>',
                'bf16x3': 'k_exchange_bf16'}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=5)
    p.add_argument('--warmup', type=int, default=2)
    p.add_argument('--config', type=int, default=4, choices=[1, 2, 4],
                   help='single-instance workload of the headline')
    p.add_argument('--precision', default='bf16x3', choices=['bf16x3', 'fp32'],
                   help='sigma-vector precision of the mixed solver')
    p.add_argument('--learned-mode', default='augment', choices=['augment', 'replace'])
    p.add_argument('--learned-until', type=float, default=1000.0,
                   help='learned correction while the largest residual exceeds this x tol')
    p.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline')
    p.add_argument('--no-legs', action='store_true', help='headline only')
    p.add_argument('--leg', choices=['benzene', 'large', 'casida', 'gw', 'fit', 'batched',
                                     'ritz'],
                   help='run this side leg alone and print its JSON (no headline line)')
    p.add_argument('--kernel-probe', nargs='?', const='k20', choices=['k20', 'mix'],
                   help='counter-run workload, then one 256 MiB device copy: k20, three '
                        'sigma applies of the headline instance at k = 20 in the mixed '
                        'precision; mix, one learned+mixed and one screened+mixed solve '
                        '(the launch sequence the timed steps repeat), printing the k of '
                        'every mixed-precision apply in order')
    return p.parse_args()


#: HBM traffic of the headline exchange kernel per launch, from the counter
#: runs over the bench's own launch sequence recorded in
#: profiles/exchange_bf16_bench_mix_pmc.md (FETCH_SIZE + WRITE_SIZE of every
#: dispatch, scaled by the same counters on a copy of known size, averaged
#: over the sequence)
TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'exchange_bf16_bench_mix_pmc.md')


def recorded_traffic():
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith('traffic_bytes_per_launch:'):
                    return float(line.split(':', 1)[1])
    except OSError:
        pass
    return None


class _ApplyLog:
    """Stands in for the library object of one operator and records the k of
    every mbpt_sigma_apply call (one exchange launch each) in order."""

    def __init__(self, lib):
        self._lib, self.ks = lib, []

    def __getattr__(self, name):
        fn = getattr(self._lib, name)
        if name != 'mbpt_sigma_apply':
            return fn

        def apply(h, k, *rest):
            self.ks.append(int(k))
            return fn(h, k, *rest)
        return apply


def kernel_probe(args):
    """Counter-run workload on the headline instance, then a 256 MiB
    device-to-device copy (the calibration kernel for FETCH_SIZE /
    WRITE_SIZE).  'k20': three mixed-precision sigma applies at k = 20.
    'mix': one learned+mixed and one screened+mixed solve, the sequence of
    exchange launches the bench's timed steps repeat; the k of each
    mixed-precision apply is printed in launch order."""
    import torch
    from mbptcode_amd.block_actions import SigmaOperator, grid_block_action
    inst, W, _ = build_problem(args.config)
    if args.kernel_probe == 'k20':
        op = SigmaOperator(inst.eps, inst.nocc, inst.X, W, inst.V, precision=args.precision)
        z = torch.randn((20, inst.nocc, inst.nvir), dtype=torch.float64, device='cuda')
        for _ in range(3):
            op.apply(z)
    else:
        from mbptcode_amd.mlp_precond import LearnedPreconditioner
        act64, d = grid_block_action(inst.eps, inst.nocc, inst.X, W, inst.V, precision='fp64')
        actm, _ = grid_block_action(inst.eps, inst.nocc, inst.X, W, inst.V,
                                    precision=args.precision)
        net = LearnedPreconditioner.load()
        rec = _ApplyLog(actm.operator._lib)
        actm.operator._lib = rec
        for s in (Solver('learned', actm, d, inst, refine=act64, preconditioner='learned',
                         learned=net, learned_mode=args.learned_mode,
                         learned_until=args.learned_until),
                  Solver('screened', actm, d, inst, refine=act64)):
            s()
        sync()
        log(f'mixed-precision applies in launch order, k = {json.dumps(rec.ks)}')
    sync()
    a = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device='cuda')
    b = torch.empty_like(a)
    b.copy_(a)
    sync()
    log(f'kernel probe done ({args.kernel_probe}, {args.precision}), one 256 MiB copy')


def log(msg):
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def sync():
    import torch
    torch.cuda.synchronize()


def build_problem(cfg, w_precision='fp64'):
    from mbptcode_amd import grids, synth
    from mbptcode_amd.space_time import screened_interaction_grid
    inst = synth.config(cfg)
    g = grids.static_screening_grid(inst.eps, inst.nocc, ntau=inst.ntau)
    sync()
    t0 = time.perf_counter()
    W, _ = screened_interaction_grid(inst.X, inst.V, inst.eps, inst.nocc, g,
                                     precision=w_precision)
    sync()
    return inst, W, time.perf_counter() - t0


class Solver:
    """One way of solving an instance: sigma precision, preconditioner,
    FP64 recheck."""

    def __init__(self, name, act, d, inst, refine=None, preconditioner='screened',
                 learned=None, learned_mode='augment', learned_until=1000.0):
        self.name, self.act, self.d, self.inst = name, act, d, inst
        self.refine, self.pre = refine, preconditioner
        self.learned, self.mode, self.until = learned, learned_mode, learned_until

    def __call__(self, tol=None, timings=None):
        from mbptcode_amd.davidson import run_davidson
        return run_davidson(self.act, self.d, self.inst.nroots,
                            self.inst.tol if tol is None else tol, 200, tda=True,
                            preconditioner=self.pre, refine_apply=self.refine,
                            learned=self.learned, learned_mode=self.mode,
                            learned_until=self.until, timings=timings)


def time_solver(solver, steps, warmup, e_ref, profile_op=None):
    """(roots/s counting roots within 1 meV of e_ref, ms/step, max error per
    root (meV), timings of the last step, exchange profile)."""
    for _ in range(warmup):
        solver()
    sync()
    if profile_op is not None:
        profile_op.profile(True)
    good, errs, tims = 0, [], {}
    t0 = time.perf_counter()
    for _ in range(steps):
        w, _, _ = solver(timings=tims)
        err = np.abs(w - e_ref) * HA_MEV
        good += int(np.sum(err <= 1.0))
        errs.append(err)
    sync()
    dt = time.perf_counter() - t0
    prof = None
    if profile_op is not None:
        prof = profile_op.profile_read()
        profile_op.profile(False)
    return good / dt, dt / steps * 1e3, np.max(errs, axis=0), tims, prof


def exchange_roofline(prof, precision, nocc, M):
    ex_ms = prof['ms']['exchange'] / prof['launches']
    ex_flops = prof['exchange_flops'] / prof['launches']
    achieved = ex_flops / (ex_ms * 1e-3) / 1e12
    out = {'bound': 'mfma', 'achieved': achieved, 'peak': PEAK_TFLOPS[precision],
           'unit': 'TFLOP/s', 'frac': achieved / PEAK_TFLOPS[precision], 'traffic': None,
           'kernel': KERNEL_NAMES[precision], 'shape': f'nocc={nocc}, M={M}', 'avg_ms': ex_ms,
           'launches': prof['launches'], 'flops_per_launch': ex_flops,
           'flop_definition': 'k (4 nocc M^2 + M^2) per launch, k = trial vectors in '
                              'the launch (algorithmic FP32-equivalent flops)',
           'launch_set': 'the timed steps\' launches that have the device to themselves: '
                         'the last cycle\'s launch of each solve, which runs beside the '
                         'FP64 recheck on a second stream, is left out',
           'stage_ms_per_apply': {k: v / prof['launches'] for k, v in prof['ms'].items()}}
    if precision == 'bf16x3':
        # MFMAs the kernel issues per (vector, 32-row chunk, 16-column tile):
        # P products 2 x (3 per full 32-wide K block + the last block's 1-2
        # side-by-side MFMAs, or 3), T products 3 per 16-row tile of T that
        # holds real rows; each a 16x16x32 BF16 MFMA of 16384 flops
        nkb = -(-nocc // 32)
        r = nocc - 32 * (nkb - 1)
        tail = 1 if 3 * r <= 32 else (2 if 3 * r <= 64 else 3)
        nfb = nkb if tail == 3 else nkb - 1
        per_tile = 2 * (3 * nfb + (tail if tail < 3 else 0)) + 3 * -(-nocc // 16)
        k_avg = ex_flops / (4.0 * nocc * M * M + M * M)
        hw = k_avg * -(-M // 32) * 4 * -(-M // 64) * per_tile * 16384
        out['mfma_hw_tflops'] = hw / (ex_ms * 1e-3) / 1e12
        out['mfma_util'] = out['mfma_hw_tflops'] / PEAK_TFLOPS['bf16x3']
        out['mfma_util_definition'] = ('issued BF16 MFMA flops/s (3 per FP32 product, the '
                                       'padded K block and T rows trimmed) / 2.5 PFLOP/s')
    return out


def chi0_roofline(inst, reps=5):
    """The chi0(0) build (prep, fused main kernel, reduce) of an instance's
    static W, FP32 MFMA, timed with HIP events on its stream."""
    from mbptcode_amd import grids
    from mbptcode_amd.space_time import Chi0Builder
    g = grids.static_screening_grid(inst.eps, inst.nocc, ntau=inst.ntau)
    b = Chi0Builder(inst.X, inst.eps, inst.nocc, precision='fp32')
    chi = b.build(g.tau_points, g.cosft_wt[0])
    sync()
    b.profile(True)
    for _ in range(reps):
        b.build(g.tau_points, g.cosft_wt[0], out=chi)
    sync()
    pr = b.profile_read()
    b.profile(False)
    avg_ms = pr['ms'] / pr['launches']
    achieved = pr['flops'] / pr['launches'] / (avg_ms * 1e-3) / 1e12
    return {'kernel': 'k_chi0_prep + k_chi0_f32 + k_chi0_reduce (one build)', 'bound': 'mfma',
            'achieved': achieved, 'peak': PEAK_TFLOPS['fp32'], 'unit': 'TFLOP/s',
            'frac': achieved / PEAK_TFLOPS['fp32'], 'traffic': None, 'avg_ms': avg_ms,
            'launches': pr['launches'], 'flops_per_launch': pr['flops'] / pr['launches'],
            'workload': f'static chi0(0), M={inst.M}, nocc={inst.nocc}, nvir={inst.nvir}, '
                        f'ntau={inst.ntau}',
            'flop_definition': 'ntau M(M+1)(nocc+nvir+1): the upper triangle with diagonal',
            'mfma_instruction': 'v_mfma_f32_16x16x4f32 (peak 157.3 TFLOP/s)'}


def headline(args):
    import torch
    from mbptcode_amd.block_actions import grid_block_action
    from mbptcode_amd.mlp_precond import LearnedPreconditioner
    inst, W, w_s = build_problem(args.config)
    log(f'headline instance cfg{args.config}: nocc={inst.nocc} nvir={inst.nvir} '
        f'M={inst.M} roots={inst.nroots}; W built in {w_s:.2f} s')
    act64, d = grid_block_action(inst.eps, inst.nocc, inst.X, W, inst.V, precision='fp64')
    actm, _ = grid_block_action(inst.eps, inst.nocc, inst.X, W, inst.V,
                                precision=args.precision)
    net = LearnedPreconditioner.load()
    plain = Solver('plain_fp64', act64, d, inst)
    mixed = Solver(f'{args.precision}+screened', actm, d, inst, refine=act64)
    learned = Solver(f'{args.precision}+learned', actm, d, inst, refine=act64,
                     preconditioner='learned', learned=net, learned_mode=args.learned_mode,
                     learned_until=args.learned_until)
    e_plain, _, _ = plain()
    e_tight, _, _ = Solver('tight', act64, d, inst)(tol=1e-8)
    log(f'roots (eV): {np.round(e_plain * 27.211386, 4).tolist()}')
    out = {}
    prof_all = None
    for s in (learned, mixed, plain):
        # the exchange launches of both mixed solvers' timed steps make up the
        # roofline average (the kernel trace of the whole run sees the same
        # launch sequence: warm-up and timed solves repeat it)
        prof_op = actm.operator if s is not plain else None
        rps, ms, err, tims, prof = time_solver(s, args.steps, args.warmup, e_plain, prof_op)
        w, _, _ = s()
        out[s.name] = {'roots_per_s': rps, 'ms_per_step': ms,
                       'max_err_meV': float(err.max()),
                       'err_meV_per_root': [float(x) for x in err],
                       'max_err_vs_tight_fp64_meV': float(np.abs(w - e_tight).max() * HA_MEV),
                       'cycles': tims.get('davidson_iterations'),
                       'sigma_vectors': tims.get('davidson_vectors'),
                       'fp64_recheck': tims.get('davidson_refine'),
                       'learned_fallback': 'learned_fallback_cycle' in tims.get('davidson_flags', {})}
        if prof is not None:
            if prof_all is None:
                prof_all = prof
            else:
                prof_all = {'ms': {k: prof_all['ms'][k] + prof['ms'][k] for k in prof['ms']},
                            'launches': prof_all['launches'] + prof['launches'],
                            'exchange_flops': prof_all['exchange_flops'] + prof['exchange_flops']}
        log(f'{s.name}: {rps:.1f} roots/s, {ms:.1f} ms/solve, max err {err.max():.2e} meV, '
            f'{tims.get("davidson_iterations")} cycles, {tims.get("davidson_vectors")} vectors')
    out['plain_roots_Ha'] = [float(x) for x in e_plain]
    out['roofline'] = exchange_roofline(prof_all, args.precision, inst.nocc, inst.M)
    tr = recorded_traffic()
    default_mix = (args.precision == 'bf16x3' and args.config == 4
                   and args.learned_mode == 'augment' and args.learned_until == 1000.0)
    if tr is not None and default_mix:
        # the counter runs replayed this launch sequence (k = 40, 60, 30, ...)
        out['roofline']['traffic'] = tr
        out['roofline']['traffic_note'] = ('HBM bytes per launch: FETCH_SIZE + WRITE_SIZE of '
                                           'each dispatch of one learned+mixed and one '
                                           'screened+mixed solve (the timed steps\' launch '
                                           'sequence), averaged, from ' +
                                           os.path.relpath(TRAFFIC_FILE, ROOT))
    out['chi0_roofline'] = chi0_roofline(inst)
    log(f'chi0 (headline instance): {out["chi0_roofline"]["achieved"]:.1f} TF '
        f'({100 * out["chi0_roofline"]["frac"]:.0f}% of FP32 MFMA peak)')
    out['end_to_end'] = end_to_end(inst, e_plain, w_s, out, args, net)
    return inst, W, out, w_s


def end_to_end(inst, e_plain, w64_s, out, args, net):
    """W build plus solve, per solver.  The mixed solvers build W on the
    precision plan -- chi0(0) from FP32 terms, FP64 Cholesky Dyson step --
    and solve on it; the plain solver's W is FP64 throughout.  Roots are
    judged against the plain FP64 solve on the FP64 W."""
    from mbptcode_amd.block_actions import grid_block_action
    build_problem(args.config, w_precision='fp32')          # warm-up
    _, W32, w32_s = build_problem(args.config, w_precision='fp32')
    act64, d = grid_block_action(inst.eps, inst.nocc, inst.X, W32, inst.V, precision='fp64')
    actm, _ = grid_block_action(inst.eps, inst.nocc, inst.X, W32, inst.V,
                                precision=args.precision)
    learned = Solver('learned', actm, d, inst, refine=act64, preconditioner='learned',
                     learned=net, learned_mode=args.learned_mode,
                     learned_until=args.learned_until)
    mixed = Solver('mixed', actm, d, inst, refine=act64)
    res = {'W_fp64_chi0_s': w64_s, 'W_fp32_chi0_s': w32_s,
           'W_definition': 'chi0(0) on the 20-point static grid, FP64 Cholesky of V and '
                           'of I - L^T chi0 L; the FP32 build forms chi0 from FP32 terms'}
    for key, s, name in (('learned+mixed', learned, f'{args.precision}+learned'),
                         ('screened+mixed', mixed, f'{args.precision}+screened')):
        s()
        sync()
        t0 = time.perf_counter()
        w, _, _ = s()
        sync()
        dt = time.perf_counter() - t0
        err = float(np.abs(w - e_plain).max() * HA_MEV)
        res[key] = {'solve_ms_fp32_W': dt * 1e3, 'max_err_meV_vs_plain_fp64_W': err,
                    'roots_per_s_end_to_end': inst.nroots / (w32_s + dt)}
    pl = out['plain_fp64']
    res['plain_fp64'] = {'roots_per_s_end_to_end': inst.nroots / (w64_s + pl['ms_per_step'] * 1e-3)}
    log(f'end to end: W fp32 {w32_s * 1e3:.0f} ms (fp64 {w64_s * 1e3:.0f} ms); learned+mixed '
        f'{res["learned+mixed"]["roots_per_s_end_to_end"]:.1f} roots/s '
        f'({res["learned+mixed"]["max_err_meV_vs_plain_fp64_W"]:.1e} meV), plain '
        f'{res["plain_fp64"]["roots_per_s_end_to_end"]:.1f} roots/s')
    return res


def leg_benzene():
    from mbptcode_amd.block_actions import grid_block_action
    inst, W, _ = build_problem(1)
    act64, d = grid_block_action(inst.eps, inst.nocc, inst.X, W, inst.V, precision='fp64')
    plain = Solver('plain_fp64', act64, d, inst)
    e_plain, _, _ = plain()
    res = {}
    for p in ('fp64', 'fp32', 'bf16x3'):
        if p == 'fp64':
            s = plain
        else:
            act, _ = grid_block_action(inst.eps, inst.nocc, inst.X, W, inst.V, precision=p)
            s = Solver(p, act, d, inst, refine=act64)
        rps, ms, err, tims, _ = time_solver(s, 5, 2, e_plain)
        res[p] = {'roots_per_s': rps, 'ms_per_step': ms, 'max_err_meV': float(err.max()),
                  'cycles': tims.get('davidson_iterations')}
    log(f'benzene-size: ' + ', '.join(f'{k} {v["roots_per_s"]:.0f} roots/s' for k, v in res.items()))
    return {'workload': 'nocc=21, nvir=93, M=1200, 10 roots, tol 1e-5', 'solvers': res}


def leg_large():
    """Per-root errors of FP32 and split-BF16 sigma against FP64, RPA E_c
    error, and the chi0 kernel's MFMA fraction on the large instance."""
    import torch
    from mbptcode_amd import grids
    from mbptcode_amd.block_actions import grid_block_action
    from mbptcode_amd.space_time import rpa_correlation_energy_grid
    inst, W, _ = build_problem(2)
    act64, d = grid_block_action(inst.eps, inst.nocc, inst.X, W, inst.V, precision='fp64')
    e64, _, _ = Solver('fp64', act64, d, inst)()
    res = {'workload': 'nocc=60, nvir=400, M=4000, 16 roots, tol 1e-5, 20 time points'}
    for p in ('fp32', 'bf16x3'):
        act, _ = grid_block_action(inst.eps, inst.nocc, inst.X, W, inst.V, precision=p)
        s = Solver(p, act, d, inst, refine=act64)
        rps, ms, err, tims, _ = time_solver(s, 3, 1, e64)
        res[p] = {'roots_per_s': rps, 'ms_per_step': ms,
                  'err_meV_per_root': [float(x) for x in err], 'max_err_meV': float(err.max())}
    s = Solver('fp64', act64, d, inst)
    rps, ms, _, _, _ = time_solver(s, 3, 1, e64)
    res['fp64'] = {'roots_per_s': rps, 'ms_per_step': ms}
    grid = grids.rpa_energy_grid(inst.eps, inst.nocc, ntau=inst.ntau)
    ec = {}
    for p in ('fp64', 'fp32'):
        rpa_correlation_energy_grid(inst.X, inst.V, inst.eps, inst.nocc, grid, precision=p)
        sync()
        t0 = time.perf_counter()
        ec[p] = rpa_correlation_energy_grid(inst.X, inst.V, inst.eps, inst.nocc, grid,
                                            precision=p)
        sync()
        ec[p + '_s'] = time.perf_counter() - t0
    split = {}
    rpa_correlation_energy_grid(inst.X, inst.V, inst.eps, inst.nocc, grid, precision='fp32',
                                timings=split)
    terms_tf = split['terms_flops'] / (split['terms_ms'] * 1e-3) / 1e12
    res['rpa_ec'] = {'fp64_Ha': ec['fp64'], 'fp32_Ha': ec['fp32'],
                     'err_uHa': abs(ec['fp32'] - ec['fp64']) * 1e6,
                     'seconds': {'fp64': ec['fp64_s'], 'fp32': ec['fp32_s']},
                     'nfreq': grid.nfreq, 'ntau': grid.ntau,
                     'fp32_split': {'chi0_terms_ms': split['terms_ms'],
                                    'chi0_transform_ms': split['transform_ms'],
                                    'chi0_share_of_Ec': (split['terms_ms'] + split['transform_ms'])
                                    * 1e-3 / split['seconds'],
                                    'seconds_profiled': split['seconds'],
                                    'fp64_frequencies': split['fp64_frequencies'],
                                    'chi0_terms_tflops': terms_tf,
                                    'chi0_terms_frac_fp32_peak': terms_tf / PEAK_TFLOPS['fp32'],
                                    'note': 'each of the ntau terms -2 Go o Gv is built once '
                                            '(one launch of the fused kernel over all time '
                                            'points) and transformed to the nfreq '
                                            'frequencies in FP64'}}
    # chi0 kernel MFMA fraction: the static chi0(0) build, 20 time points
    res['chi0_roofline'] = chi0_roofline(inst)
    log(f'large: fp32 {res["fp32"]["max_err_meV"]:.1e} meV, bf16x3 '
        f'{res["bf16x3"]["max_err_meV"]:.1e} meV, E_c err {res["rpa_ec"]["err_uHa"]:.2e} uHa, '
        f'E_c fp32 {ec["fp32_s"] * 1e3:.0f} ms (chi0 terms {split["terms_ms"]:.1f} ms at '
        f'{terms_tf:.1f} TF, transforms {split["transform_ms"]:.1f} ms), static chi0 '
        f'{res["chi0_roofline"]["achieved"]:.1f} TF')
    return res


def leg_full_action(inst, W, k=20, reps=5):
    """The full Casida action (A z, B z) against the Tamm-Dancoff A z alone on
    the headline instance: per-apply times of k vectors and the B block's
    extra cost (its exchange is 4 n_o M^2 flops per vector, in split BF16 on
    the split-BF16 handle), with B z's error against FP64."""
    import torch
    from mbptcode_amd.block_actions import SigmaOperator
    z = torch.randn((k, inst.nocc, inst.nvir), dtype=torch.float64, device='cuda')
    res, Bs = {}, {}
    for prec in ('fp64', 'bf16x3'):
        op = SigmaOperator(inst.eps, inst.nocc, inst.X, W, inst.V, precision=prec)
        times = {}
        for name, fn in (('A', lambda: op.apply(z)), ('AB', lambda: op.apply_ab(z))):
            fn()
            sync()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            sync()
            times[name] = e0.elapsed_time(e1) / reps
        Bs[prec] = op.apply_ab(z)[1]
        extra = times['AB'] - times['A']
        res[prec] = {'A_ms': times['A'], 'AB_ms': times['AB'], 'B_extra_ms': extra,
                     'B_exchange_tflops': k * 4.0 * inst.nocc * inst.M ** 2 / (extra * 1e-3) / 1e12}
        del op
    err = float((Bs['bf16x3'] - Bs['fp64']).norm() / Bs['fp64'].norm())
    log(f'full action (k={k}): fp64 A {res["fp64"]["A_ms"]:.2f} / A+B {res["fp64"]["AB_ms"]:.2f} ms, '
        f'bf16x3 A {res["bf16x3"]["A_ms"]:.2f} / A+B {res["bf16x3"]["AB_ms"]:.2f} ms; '
        f'B rel err bf16x3 vs fp64 {err:.1e}')
    # the full Casida Davidson on the same instance: split-BF16 action with the
    # FP64 recheck against the plain FP64 solve
    from mbptcode_amd.block_actions import grid_block_action
    from mbptcode_amd.davidson import run_davidson
    solves, roots = {}, {}
    for name, prec in (('bf16x3+fp64_recheck', 'bf16x3'), ('plain_fp64', 'fp64')):
        act, d = grid_block_action(inst.eps, inst.nocc, inst.X, W, inst.V, precision=prec,
                                   tda=False)
        ref64 = None
        if prec != 'fp64':
            ref64 = grid_block_action(inst.eps, inst.nocc, inst.X, W, inst.V,
                                      precision='fp64', tda=False)[0]
        run_davidson(act, d, inst.nroots, inst.tol, 200, tda=False, refine_apply=ref64)
        tm = {}
        sync()
        t0 = time.perf_counter()
        w, _, _ = run_davidson(act, d, inst.nroots, inst.tol, 200, tda=False,
                               refine_apply=ref64, timings=tm)
        sync()
        dt = time.perf_counter() - t0
        roots[name] = w
        solves[name] = {'s': dt, 'roots_per_s': inst.nroots / dt,
                        'cycles': tm['davidson_iterations'], 'vectors': tm['davidson_vectors'],
                        'max_residual_fp64': max(tm['davidson_residuals'])}
        log(f'full Casida Davidson {name}: {dt * 1e3:.1f} ms, {inst.nroots / dt:.1f} roots/s, '
            f'{tm["davidson_iterations"]} cycles, {tm["davidson_vectors"]} vectors')
        del act, ref64
    derr = float(np.abs(roots['bf16x3+fp64_recheck'] - roots['plain_fp64']).max() * HA_MEV)
    return {'workload': f'headline instance, k = {k} trial vectors per apply',
            'precisions': res, 'B_rel_err_bf16x3_vs_fp64': err,
            'davidson': {'nroots': inst.nroots, 'tol': inst.tol, 'solvers': solves,
                         'max_root_diff_meV': derr,
                         'speedup_vs_plain_fp64': solves['bf16x3+fp64_recheck']['roots_per_s']
                         / solves['plain_fp64']['roots_per_s']}}


def leg_gw():
    """The space-time GW correlation self-energy Sigma^c_pp(i w) on the 16
    Pade frequencies: stage times and the fused branch-sum kernel's FP32 MFMA
    rate, for a 10-state gap window and for every state, at the large and the
    learned-preconditioner instances."""
    from mbptcode_amd import synth
    from mbptcode_amd.gw import self_energy_diagonal_space_time
    res = {}
    for cfg in (2, 4):
        inst = synth.config(cfg)
        window = list(range(inst.nocc - 5, inst.nocc + 5))
        for name, states in (('window10', window),
                             ('all', list(range(inst.nocc + inst.nvir)))):
            self_energy_diagonal_space_time(inst.X, inst.V, inst.eps, inst.nocc, states[:2])
            sync()
            tm = {}
            t0 = time.perf_counter()
            sig, ax = self_energy_diagonal_space_time(inst.X, inst.V, inst.eps, inst.nocc,
                                                      states, timings=tm)
            sync()
            wall = time.perf_counter() - t0
            k = tm['sigma_kernel']
            tf = k['flops'] / (k['ms'] * 1e-3) / 1e12
            key = f'cfg{cfg}_{name}'
            res[key] = {'workload': f'nocc={inst.nocc}, nvir={inst.nvir}, M={inst.M}, '
                                    f'{len(states)} states, ntau={ax["ntau"]}',
                        'seconds': wall,
                        'stages_s': {n: tm[n] for n in ('transforms', 'screening', 'to_tau',
                                                        'sigma')},
                        'sigma_kernel_ms': k['ms'], 'sigma_kernel_tflops': tf,
                        'sigma_kernel_frac_fp32_peak': tf / PEAK_TFLOPS['fp32'],
                        'max_abs_sigma_Ha': float(np.abs(sig).max())}
            log(f'gw {key}: {wall * 1e3:.0f} ms (screening {tm["screening"] * 1e3:.0f} ms, '
                f'sigma kernel {k["ms"]:.2f} ms, {tf:.1f} TF = '
                f'{100 * tf / PEAK_TFLOPS["fp32"]:.0f}% FP32 peak), ntau {ax["ntau"]}')
    return res


def leg_fit():
    """The separable fit at the learned-preconditioner instance's size: 6000
    points, a 60000-column test set, 2000 auxiliary functions (synthetic D and
    F: the integrals behind them need pyscf).  FP64 throughout."""
    import torch
    from mbptcode_amd import isdf_fit as F
    nk, nrho, naux = 6000, 60000, 2000
    g = torch.Generator(device='cuda').manual_seed(0)
    D = torch.randn((nk, nrho), generator=g, dtype=torch.float64, device='cuda')
    Ff = torch.randn((naux, nrho), generator=g, dtype=torch.float64, device='cuda')
    B = torch.randn((naux, naux), generator=g, dtype=torch.float64, device='cuda')
    V = B @ B.T / naux + 0.05 * torch.eye(naux, dtype=torch.float64, device='cuda')
    F.fit_M_stable(D[:500], Ff)
    sync()
    t0 = time.perf_counter()
    M = F.fit_M_stable(D, Ff)
    sync()
    t_fit = time.perf_counter() - t0
    t0 = time.perf_counter()
    Dsep = F.separable_factor(M, V)
    sync()
    t_gauge = time.perf_counter() - t0
    # the streaming form over 6000-column blocks of the same test set
    blk = 6000
    t0 = time.perf_counter()
    Ms = F.fit_M_streaming((D[:, j:j + blk], Ff[:, j:j + blk]) for j in range(0, nrho, blk))
    sync()
    t_stream = time.perf_counter() - t0
    stream_diff = float((Ms - M).abs().max() / M.abs().max())
    flops = 2.0 * nk * nk * nrho + 2.0 * naux * nk * nrho
    res = {'workload': f'nk={nk}, nrho={nrho}, naux={naux}, FP64',
           'fit_s': t_fit, 'gauge_s': t_gauge,
           'fit_gemm_tflops': flops / t_fit / 1e12,
           'streaming_fit_s': t_stream, 'streaming_block': blk,
           'streaming_vs_whole_rel': stream_diff,
           'factor_shape': list(Dsep.shape)}
    log(f'fit: {t_fit * 1e3:.0f} ms ({res["fit_gemm_tflops"]:.1f} TF FP64 over its two '
        f'products), streaming {t_stream * 1e3:.0f} ms ({stream_diff:.1e} from the whole '
        f'fit), gauge {t_gauge * 1e3:.0f} ms')
    return res


def leg_batched():
    """The batched workload: 256 instances, 5 roots each, tol 1e-5."""
    import torch
    from mbptcode_amd import synth, grids
    from mbptcode_amd.batched import BatchedSigma, autotune, solve_batched, time_pass
    from mbptcode_amd.space_time import screened_interaction_grid
    sizes = synth.batched_sizes(256)
    insts = [synth.batched_instance(i, sizes) for i in range(256)]
    Ws = []
    for inst in insts:
        g = grids.static_screening_grid(inst.eps, inst.nocc, ntau=inst.ntau)
        Ws.append(screened_interaction_grid(inst.X, inst.V, inst.eps, inst.nocc, g,
                                            precision='fp32')[0])
    args = ([i.eps for i in insts], [i.nocc for i in insts], [i.X for i in insts], Ws,
            [i.V for i in insts])
    op64 = BatchedSigma(*args, precision='fp64', kmax=10)
    op32 = BatchedSigma(*args, precision='fp32', kmax=10)
    tune = {}
    for op in (op64, op32):
        plan, info = autotune(op)
        tune[op.precision] = {'plan': plan, 'pass_ms': info.get('ms'), 'source': info['source']}
    # one captured sigma pass at kmax vectors per instance against the MFMA
    # peak of its precision (algorithmic flops of every stage)
    no = np.asarray(op32.nocc, float)
    nv = np.asarray(op32.nvir, float)
    Mg = np.asarray(op32.M, float)
    per_vec = 4 * no * Mg ** 2 + 3 * Mg ** 2 + 4 * no * nv * Mg + 2 * no * Mg
    pass_flops = float(op32.kmax * per_vec.sum())
    passes = {}
    for op in (op32, op64):
        ms = time_pass(op, reps=7)
        tf = pass_flops / (ms * 1e-3) / 1e12
        peak = PEAK_TFLOPS['fp32' if op.precision == 'fp32' else 'fp64']
        passes[op.precision] = {'pass_ms': ms, 'achieved': tf, 'peak': peak,
                                'unit': 'TFLOP/s', 'frac': tf / peak, 'bound': 'mfma'}
    runs = {}
    e_plain = None
    for name, op, ref in (('plain_fp64', op64, None), ('fp32+fp64_recheck', op32, op64)):
        solve_batched(op, 5, 1e-5, refine_op=ref)          # warm-up
        sync()
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            e, _, info = solve_batched(op, 5, 1e-5, refine_op=ref)
        sync()
        dt = (time.perf_counter() - t0) / reps
        if e_plain is None:
            e_plain = e
        err = float(np.abs(e - e_plain).max() * HA_MEV)
        runs[name] = {'roots_per_s': 256 * 5 / dt, 'ms_per_solve': dt * 1e3,
                      'cycles': info['cycles'], 'refine_cycles': info.get('refine_cycles'),
                      'converged': info['converged'], 'max_err_meV': err}
    log('batched: ' + ', '.join(f'{k} {v["roots_per_s"]:.0f} roots/s' for k, v in runs.items())
        + '; sigma pass ' + ', '.join(f'{k} {v["pass_ms"]:.2f} ms ({100 * v["frac"]:.0f}% of '
                                      f'peak)' for k, v in passes.items()))
    return {'workload': '256 instances, nocc~U[5,40], nvir~U[20,200], M=8(nocc+nvir), '
                        '5 roots, tol 1e-5', 'solvers': runs, 'plan': tune,
            'sigma_pass': dict(passes, kmax=op32.kmax, flops_per_pass=pass_flops,
                               flop_definition='kmax vectors per instance, each '
                                               '4 nocc M^2 + 3 M^2 + 4 nocc nvir M + '
                                               '2 nocc M (exchange, Hartree, head and '
                                               'tail products)')}


def leg_ritz(reps=20):
    """The Davidson's Rayleigh-Ritz step at the subspace sizes of the
    headline solves and above: the device eigensolver (ritz.hip, the lowest
    30 pairs and their quadratic forms, timed with HIP events on its stream)
    against host LAPACK's eigh on one thread (timed on the host clock),
    random symmetric matrices with the Ritz matrices' scale."""
    import torch
    from threadpoolctl import threadpool_limits
    from mbptcode_amd.subspace import ritz_eigh
    rng = np.random.default_rng(0)
    out = []
    for m in (40, 70, 100, 130, 161, 192):
        A = rng.normal(size=(m, m))
        H = A + A.T
        d = torch.as_tensor(np.concatenate([H, H @ H], 1), device='cuda')
        k = 30
        ritz_eigh(d[:, :m], d[:, m:], k)
        sync()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            _, wq = ritz_eigh(d[:, :m], d[:, m:], k)
        e.record()
        sync()
        dev_us = s.elapsed_time(e) / reps * 1e3
        err = float(np.abs(wq.cpu().numpy()[:k] - np.linalg.eigvalsh(H)[:k]).max())
        with threadpool_limits(1):
            np.linalg.eigh(H)
            t0 = time.perf_counter()
            for _ in range(reps):
                np.linalg.eigh(H)
            host_us = (time.perf_counter() - t0) / reps * 1e6
        out.append({'m': m, 'k': k, 'device_us': dev_us, 'host_lapack_us': host_us,
                    'max_eig_err': err})
        log(f'ritz m={m}: device {dev_us:.0f} us, host LAPACK {host_us:.0f} us, '
            f'eigenvalue err {err:.1e}')
    return out


def _blas_threads():
    try:
        from threadpoolctl import threadpool_info
        return [(i.get('internal_api'), i.get('num_threads')) for i in threadpool_info()
                if i.get('user_api') == 'blas']
    except Exception:
        return []


def cpu_baseline(inst, W_grid, e_ref):
    """The oracle's FP64 solve of the headline instance on the host cores:
    its block action (the reference's algorithm, one trial vector at a time)
    driven by its block Jacobi-Davidson with the screened diagonal, timed
    whole."""
    from oracle import isdf_oracle as O
    cores = len(os.sched_getaffinity(0))
    D = O.grid_factors(inst.V)
    W_aux = np.linalg.solve(D, np.linalg.solve(D, W_grid).T).T
    W_aux = 0.5 * (W_aux + W_aux.T)
    apply_A, d = O.isdf_block_action(inst.eps, inst.nocc, inst.X, D, W_aux)
    no, nv = inst.nocc, inst.nvir
    n = no * nv

    def mm(V):
        k = V.shape[1]
        return apply_A(np.ascontiguousarray(V.T).reshape(k, no, nv)).reshape(k, n).T
    t0 = time.perf_counter()
    pdiag = apply_A.screened_diagonal()
    e, _, conv, applied = O.davidson_tda(mm, d, inst.nroots, tol=inst.tol,
                                         precond_diag=pdiag)
    dt = time.perf_counter() - t0
    err = float(np.abs(np.asarray(e) - e_ref).max() * HA_MEV) if e_ref is not None else None
    return {'value': inst.nroots / dt, 'unit': 'roots/s', 'cores': cores, 'kind': 'port',
            'sample': (f'one whole FP64 solve of the headline instance by the oracle '
                       f'(block action one trial vector at a time, block Jacobi-Davidson '
                       f'on the screened diagonal, tol {inst.tol:g}): {dt:.1f} s, '
                       f'{applied} sigma-vectors; W built on the GPU beforehand'),
            'seconds': dt, 'sigma_vectors': int(applied),
            'converged': bool(np.all(conv)), 'max_err_meV_vs_gpu_plain': err,
            'blas': _blas_threads()}


def cpu_baseline_chi0(inst):
    """The oracle's proj(tau) (the reference's polarizability_projected_tau,
    grid rows tiled) on the host cores for one time point of the headline
    instance's static grid in the grid gauge (D = I), scaled to the build's
    n_tau points."""
    from mbptcode_amd import grids
    from oracle import isdf_oracle as O
    cores = len(os.sched_getaffinity(0))
    g = grids.static_screening_grid(inst.eps, inst.nocc, ntau=inst.ntau)
    X_o, X_v, e_o, e_v, _ = O.split_branches(inst.X, inst.eps, inst.nocc)
    D = np.eye(inst.M)
    t0 = time.perf_counter()
    O.polarizability_projected_tau(X_o, X_v, e_o, e_v, D, float(g.tau_points[0]))
    dt = time.perf_counter() - t0
    build_s = dt * inst.ntau
    flops = inst.ntau * inst.M * (inst.M + 1.0) * (inst.nocc + inst.nvir + 1.0)
    return {'value': flops / build_s / 1e12, 'unit': 'TFLOP/s', 'cores': cores,
            'kind': 'port', 'seconds_per_build': build_s,
            'flop_definition': 'ntau M(M+1)(nocc+nvir+1), as the GPU chi0 roofline',
            'sample': (f'proj(tau) of the oracle at one of the {inst.ntau} time points, '
                       f'M={inst.M}, D = I (its two projections by D are part of the '
                       f'reference algorithm and are timed): {dt:.2f} s, scaled to '
                       f'{inst.ntau} points'),
            'blas': _blas_threads()}


def main():
    args = parse()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU: this build has no multi-GPU path '
              f'(asked for --gpus {args.gpus}).', file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    if args.kernel_probe:
        kernel_probe(args)
        return
    if args.leg:
        legs = {'benzene': leg_benzene, 'large': leg_large, 'gw': leg_gw,
                'fit': leg_fit, 'batched': leg_batched, 'ritz': leg_ritz}
        if args.leg == 'casida':
            inst, W, _ = build_problem(4)
            print(json.dumps({'casida': leg_full_action(inst, W)}), flush=True)
            return
        print(json.dumps({args.leg: legs[args.leg]()}), flush=True)
        return
    inst, W, out, w_s = headline(args)
    lm = out[f'{args.precision}+learned']
    pl = out['plain_fp64']
    result = {
        'metric': 'BSE Davidson roots/s at <=1 meV vs FP64 (learned+mixed vs plain); chi0 MFMA %',
        'value': lm['roots_per_s'], 'unit': 'roots/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': lm['ms_per_step'],
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': args.precision, 'data': 'synthetic',
        'config': {'workload': 'learned-preconditioner instance: ISDF BSE-TDA singlet Davidson',
                   'nocc': inst.nocc, 'nvir': inst.nvir, 'M': inst.M, 'nroots': inst.nroots,
                   'tol': inst.tol, 'ntau_W': inst.ntau, 'seed': 0,
                   'solver': f'{args.precision} sigma + learned correction '
                             f'({args.learned_mode}, above {args.learned_until:g} tol) + '
                             f'FP64 subspace + FP64 recheck'},
        'max_err_meV': lm['max_err_meV'],
        'speedup_vs_plain_fp64': lm['roots_per_s'] / pl['roots_per_s'],
        'solvers': {k: v for k, v in out.items()
                    if k not in ('roofline', 'chi0_roofline', 'end_to_end', 'plain_roots_Ha')},
        'W_build_s': w_s,
        'end_to_end': out['end_to_end'],
        'roofline': out['roofline'],
        'chi0_roofline': out['chi0_roofline'],
    }
    if not args.no_legs:
        log('leg: benzene-size instance')
        result['benzene_size'] = leg_benzene()
        log('leg: large instance (errors, E_c, chi0)')
        result['large_instance'] = leg_large()
        log('leg: full Casida action (B block)')
        result['full_casida_action'] = leg_full_action(inst, W)
        log('leg: space-time GW self-energy')
        result['gw_self_energy'] = leg_gw()
        log('leg: separable fit')
        result['isdf_fit'] = leg_fit()
        log('leg: batched workload')
        result['batched'] = leg_batched()
    if not args.no_cpu:
        log('CPU baseline (FP64 oracle)')
        result['cpu_baseline'] = cpu_baseline(inst, W.cpu().numpy(),
                                              np.asarray(out['plain_roots_Ha']))
        log(f'cpu: {result["cpu_baseline"]["value"]:.3f} roots/s '
            f'({result["cpu_baseline"]["seconds"]:.1f} s solve, '
            f'{result["cpu_baseline"]["sigma_vectors"]} sigma-vectors)')
        result['cpu_baseline_chi0'] = cpu_baseline_chi0(inst)
        log(f'cpu chi0: {result["cpu_baseline_chi0"]["value"]:.3f} TFLOP/s algorithmic')
    print(json.dumps(result), flush=True)


if __name__ == '__main__':
    main()
