#!/usr/bin/env python
"""Benchmark: propagated density-matrix steps per second on one MI355X.

Workload: the 2-D (third-order, photon-echo '-++', GSB + ESE + ESA) spectrum
of a 7-site aggregate (site energies U(12200, 12700) cm^-1 and couplings
N(0, 60) cm^-1 from RandomState(1), FMO unit dipoles), non-secular Redfield
in the g/e/f manifolds (21 double excitations), Drude bath lambda = 35 cm^-1,
cutoff 106 cm^-1, 77 K; t1 x t3 = 128 x 128 at 2 fs, t2 in {0, 100, 200,
500} fs; every t1 branch is a right-hand side of the t2 stage.

One benchmark step is one whole 2-D spectrum (third_order_response plus the
two FFTs), with the Liouvillian blocks already built and resident in HBM.
One propagated density-matrix step is one RK4 step (4 Liouvillian
applications) of one state vector; a 2-D spectrum is 3 * (127 + 128*250 +
127) = 96,762 of them.

Legs: 'mixed' (complex64 Liouvillian products, FP64 diagonal / state / RK
update; the reported value), 'plain' (all complex128).  The CPU baseline is
the oracle's restatement of the reference path (scipy zvode at default
tolerance, serial loop over right-hand sides) on one host thread.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

METRIC = 'propagated density-matrix steps/sec, batched-FP32/64 vs plain FP64, @1e-6 rel err'
PEAK_TFLOPS = {'fp64': 78.6, 'mixed': 157.3}   # MI355X dense FP64 / FP32 (vector and MFMA)


def parse_args(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--cpu-baseline', choices=('auto', 'off'), default='auto')
    ap.add_argument('--cpu-budget', type=float, default=30.0,
                    help='seconds of CPU work allowed for the baseline sample')
    return ap.parse_args(argv)


def run_plain(qa, model, work):
    """The drop-in path, pathway by pathway, classical RK4 (four Liouvillian
    products per step), all complex128."""
    return qa.two_dimensional_spectra(
        model, work['coherence_time_max'], population_times=work['population_times'],
        geometry=work['geometry'], include_signal=work['include_signal'], dt=work['dt'],
        precision='fp64', schedule='rk4', plan='pathways')[1]


class FusedSpectrum(object):
    """The batched path: fused plan (merged pathways, t3 on a second stream,
    RK4 propagator schedules: cached propagator powers for the t2 stage and
    time-chunked stepping for t1/t3, complex64 products with FP64 state), the
    whole 2-D spectrum including both FFTs replayed from one HIP graph."""

    def __init__(self, model, work, precision='mixed'):
        from qspectra_amd.simulate.fused import ThirdOrderPlan
        self.model = model
        self.plan = ThirdOrderPlan(model, work['coherence_time_max'],
                                   population_times=work['population_times'],
                                   geometry=work['geometry'],
                                   include_signal=work['include_signal'], dt=work['dt'],
                                   precision=precision).capture(two_dimensional=True)

    def __call__(self):
        return self.plan.run()[0]


def time_leg(torch, fn, steps):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0, out


def dominant_kernel(summary):
    """Largest total time among the propagation launches; algorithmic work
    8 M^2 real flops per complex matrix-vector product: 4 per RK4 step, or 1
    per step with the RK4 propagator."""
    prop = {k: v for k, v in summary.items() if k.startswith('propagate')}
    label = max(prop, key=lambda k: prop[k]['total_ms'])
    info = prop[label]
    f = dict(re.findall(r'(\w+)=(\w+)', label))
    M, R, B, steps = int(f['M']), int(f['R']), int(f['B']), int(f['steps'])
    products = 4 if ' rk4 ' in label else 1
    flops = 8.0 * products * M * M * R * B * steps
    return label, info, flops


def cpu_baseline(budget, steps_per_spectrum, min_seconds=10.0):
    """The oracle's zvode restatement of the reference path, one thread.

    The sample is the whole 2-D spectrum, repeated until ``min_seconds`` of
    CPU work; if one spectrum would exceed ``budget``, whole pathways are run
    until the budget is spent and the rate is taken over their steps."""
    from threadpoolctl import threadpool_limits
    from oracle import system as S, liouville as Lv, propagate as P
    from qspectra_amd import workloads as W
    H1, d, extra = W.aggregate7_hamiltonian()
    ham = S.ElectronicHamiltonian(H1, bath=S.DebyeBath(S.CM_K * 77, 35, 106), dipoles=d,
                                  energy_spread_extra=extra)
    model = Lv.LiouvilleModel(ham, unit_convert=S.CM_FS, secular=False)
    work = W.AGG7_2D
    per_path = steps_per_spectrum // 3
    done_steps, elapsed, runs = 0, 0.0, 0
    with threadpool_limits(limits=1):
        model.evolution_super_operator        # Liouvillian built outside the timed part
        while elapsed < min_seconds and elapsed < budget:
            for name in ('GSB', 'ESE', 'ESA'):
                t0 = time.perf_counter()
                P.third_order_response(model, work['coherence_time_max'],
                                       population_times=work['population_times'],
                                       geometry=work['geometry'], include_signal=name,
                                       integrator='zvode')
                elapsed += time.perf_counter() - t0
                done_steps += per_path
                if elapsed > budget:
                    break
            runs += 1
    return dict(value=done_steps / elapsed, unit='steps/s', cores=1, kind='port',
                sample='%d repetitions of the benchmark 2-D spectrum (%d RK4-equivalent steps) '
                       'through the oracle restatement of the reference path: scipy zvode at '
                       'rtol 1e-6, serial loop over t1 branches, Liouvillian prebuilt; %.1f s on '
                       'one host thread' % (runs, done_steps, elapsed))


def main(argv=None):
    args = parse_args(argv)
    if args.gpus != 1:
        print('bench.py measures one GPU: this build runs on a single MI355X (the path '
              'shards by ensemble member; see DESIGN.md). Use --gpus 1.', file=sys.stderr)
        return 2
    import torch
    import qspectra_amd as qa
    from qspectra_amd import workloads as W, timing
    qa.require_native()
    model = W.aggregate7_model(secular=False)
    work = W.AGG7_2D
    steps_per_spectrum = W.rk4_steps_2d()

    fused = FusedSpectrum(model, work, 'mixed')
    for _ in range(max(args.warmup, 1)):
        run_plain(qa, model, work)
        fused()
    fused.plan.check_status()

    t_plain, X64 = time_leg(torch, lambda: run_plain(qa, model, work), args.steps)
    t_mixed, X32 = time_leg(torch, fused, args.steps)
    X32 = X32.cpu().numpy()
    err = float(np.max(np.abs(X32 - X64)) / np.max(np.abs(X64)))
    # per-kernel durations: the same launch sequence without the graph, with
    # HIP events around every launch on the stream it runs on
    with timing.kernel_timer() as kt:
        for _ in range(args.steps):
            fused.plan._launch()
    summary = kt.summary()
    label, info, flops = dominant_kernel(summary)
    achieved = flops / (info['mean_ms'] * 1e-3) / 1e12
    value = steps_per_spectrum * args.steps / t_mixed
    plain = steps_per_spectrum * args.steps / t_plain

    line = {
        'metric': METRIC, 'value': value, 'unit': 'steps/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': 1e3 * t_mixed / args.steps, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32 products / f64 state',
        'data': 'synthetic',
        'config': {'workload': '2-D spectrum (GSB+ESE+ESA, -++) of a 7-site aggregate, '
                               'non-secular Redfield g/e/f, t1 x t3 = 128 x 128 at 2 fs, '
                               't2 in {0,100,200,500} fs, all t1 branches as RHS',
                   'rk4_steps_per_spectrum': steps_per_spectrum, 'parallelism': 'single GPU'},
        'plain_fp64': {'value': plain, 'ms_per_step': 1e3 * t_plain / args.steps},
        'speedup_vs_plain': value / plain,
        'rel_err_vs_fp64': err,
        'roofline': {'bound': 'mfma', 'kernel': label, 'achieved': achieved,
                     'peak': PEAK_TFLOPS['mixed'], 'unit': 'TFLOP/s',
                     'frac': achieved / PEAK_TFLOPS['mixed'], 'traffic': None,
                     'launch_ms': info['mean_ms'], 'algorithmic_flop': flops},
        'kernels': {k: {'count': v['count'], 'mean_ms': round(v['mean_ms'], 5)}
                    for k, v in summary.items()},
    }
    if args.cpu_baseline == 'auto':
        line['cpu_baseline'] = cpu_baseline(args.cpu_budget, steps_per_spectrum)
        line['vs_cpu'] = value / line['cpu_baseline']['value']
    else:
        line['cpu_baseline'] = None
    print(json.dumps(line))
    return 0


if __name__ == '__main__':
    sys.exit(main())
