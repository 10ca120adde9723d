#!/usr/bin/env python
"""Benchmark: propagated density-matrix steps per second on one MI355X.

Workload: the 2-D (third-order, photon-echo '-++', GSB + ESE + ESA) spectrum
of a 7-site aggregate (site energies U(12200, 12700) cm^-1 and couplings
N(0, 60) cm^-1 from RandomState(1), FMO unit dipoles), non-secular Redfield
in the g/e/f manifolds (21 double excitations), Drude bath lambda = 35 cm^-1,
cutoff 106 cm^-1, 77 K; t1 x t3 = 128 x 128 at 2 fs, t2 in {0, 100, 200,
500} fs; every t1 branch is a right-hand side of the t2 stage.

One benchmark step is one whole 2-D spectrum (third_order_response plus the
two FFTs) computed from the Liouvillian blocks resident in HBM: every replay
of the captured plan rebuilds the RK4 propagators, their powers and the
mixed-precision split from the blocks, then propagates and contracts.  One
propagated density-matrix step is one 2 fs step of one state vector on the
spectrum's time grid; a 2-D spectrum is 3 * (127 + 128*250 + 127) = 96,762
of them.  Each 2 fs step is integrated as two classical RK4 steps of 1 fs,
so that the spectrum is within 1e-6 of the reference's tight-tolerance
result (tests/golden/spectra_2d_full.npz; RK4 at 2 fs is 4.4e-6 from it).

Legs: 'mixed' (complex64 Liouvillian products, FP64 diagonal / state / RK
update; the reported value), 'plain' (pathway by pathway, classical RK4 at
the same 1 fs step, all complex128).  The CPU baseline is the oracle's
restatement of the reference path (scipy zvode at default tolerance, serial
loop over right-hand sides), timed on all host cores (default BLAS
threading) and on one thread.
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
REFERENCE_2D = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'tests', 'golden',
                            'spectra_2d_full.npz')


def parse_args(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=100)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--cpu-baseline', choices=('auto', 'off'), default='auto')
    ap.add_argument('--cpu-budget', type=float, default=40.0,
                    help='seconds of CPU work allowed for the baseline samples (all cores, '
                         'then one thread)')
    ap.add_argument('--workloads', default='all',
                    help="'all', 'none' or a comma list of batched_2d, ensemble_absorption, "
                         "adjoint_gradient, heom, pump, example_dimer_2d, zofe_fmo, "
                         "jonas_vibronic: the further workloads to measure")
    ap.add_argument('--counters', action='store_true',
                    help='no timing: run a fixed kernel sequence for hardware-counter '
                         'profiles (a 256 MiB elementwise negation for calibration, the 2-D spectrum '
                         'launch sequence, the 256-member ensemble plan, the 5120-system '
                         'absorption ensemble plan, two HEOM RK4 steps of 256 members)')
    return ap.parse_args(argv)


def counter_sequence(torch, qa):
    """Kernel sequence profiled with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE):
    the calibration kernel (an elementwise negation of 256 MiB) reads and
    writes exactly 256 MiB per dispatch; then
    three launch sequences of the headline 2-D plan (MIXED, the plan's
    chunk length 16), two of the 256-member ensemble plan (MIXED, rebuilt
    from the Hamiltonians), two of the 5120-system absorption ensemble plan
    and two of its propagation as one launch over every wavefront, and two
    HEOM RK4 steps of 256 members (FP64 and MIXED)."""
    from qspectra_amd import workloads as W
    from qspectra_amd.simulate.utils import propagate_heom
    from qspectra_amd.simulate.fused import ThirdOrderPlan
    # calibration: a kernel no other step launches at this size (device
    # copies share one copy kernel whose grid does not depend on the size)
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device='cuda')     # 256 MiB
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    work = W.AGG7_2D
    kw = dict(population_times=work['population_times'], geometry=work['geometry'],
              include_signal=work['include_signal'], dt=work['dt'], precision='mixed',
              fourier=True)
    plan = ThirdOrderPlan(W.aggregate7_model(secular=False), work['coherence_time_max'],
                          chunk=16, **kw)
    for _ in range(3):
        plan._launch()
    torch.cuda.synchronize()
    m = W.aggregate7_model(secular=False, disorder_sigma=100.0)
    mb = m.with_hamiltonians(list(m.hamiltonian.sample_ensemble(256)))
    ens = ThirdOrderPlan(mb, work['coherence_time_max'], ensemble_mean=True, rebuild=True, **kw)
    for _ in range(2):
        ens._launch()
    torch.cuda.synchronize()
    from qspectra_amd.batched import AbsorptionEnsemblePlan
    absorb = AbsorptionEnsemblePlan(W.ensemble_absorption_hamiltonians(), 4096.0,
                                    rw_freq=12450.0, dt=1.0, graph=False)
    for _ in range(2):
        absorb._launch()
    torch.cuda.synchronize()
    from qspectra_amd import _native
    for _ in range(2):       # the propagation over every wavefront in one launch
        _native.response_buckets_into(absorb.waves, absorb.offsets, absorb.P, absorb.y0,
                                      absorb.W, absorb.signal, absorb.status, J=absorb.J)
    torch.cuda.synchronize()
    m = W.heom_aggregate7_model()
    eom = m.with_hamiltonians([m.hamiltonian.sample(n) for n in range(256)]) \
        .equation_of_motion('ee')
    y0 = torch.zeros((256, eom.size), dtype=torch.complex128, device='cuda')
    y0[:, 0] = 1.0
    for prec in ('fp64', 'mixed'):
        propagate_heom(eom, y0, np.array([0.0, 1.0]), dt=0.5, precision=prec)
    torch.cuda.synchronize()
    print('counter sequence done', file=sys.stderr)


def run_plain(qa, model, work):
    """The drop-in path, pathway by pathway, classical RK4 (four Liouvillian
    products per step), all complex128."""
    return qa.two_dimensional_spectra(
        model, work['coherence_time_max'], population_times=work['population_times'],
        geometry=work['geometry'], include_signal=work['include_signal'], dt=work['dt'],
        precision='fp64', schedule='rk4', plan='pathways')[1]


class FusedSpectrum(object):
    """The batched path: fused plan (one shared t1 trajectory; the t2 stage
    on the rows of the V1 map with cached RK4 propagator powers; t3
    trajectories on streams of their own; time-chunked stepping with complex64
    products and FP64 state; both Fourier transforms as GEMMs with the
    transform matrices), the whole 2-D spectrum replayed from one HIP graph."""

    def __init__(self, model, work, precision='mixed'):
        from qspectra_amd.simulate.fused import ThirdOrderPlan, plan_chunk
        self.model = model
        kw = dict(population_times=work['population_times'], geometry=work['geometry'],
                  include_signal=work['include_signal'], dt=work['dt'], precision=precision,
                  fourier=True)
        from qspectra_amd import schedules
        schedules._load_cache()
        before = set(schedules._TUNED)
        chunk = plan_chunk(model, work['coherence_time_max'], **kw)
        self.chunk = chunk
        # candidate times of the plan tuner (empty when the decision came
        # from the on-disk cache)
        self.chunk_ms = {}
        for k in set(schedules._TUNED) - before:
            if k.split(' | ', 1)[-1].startswith('plan '):
                self.chunk_ms.update(schedules._TUNED[k].get('ms', {}))
        self.plan = ThirdOrderPlan(model, work['coherence_time_max'], chunk=chunk,
                                   **kw).capture()

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


# launch labels that are one kernel each (the others -- 'rk4 propagator',
# 'matrix powers' of blocks beyond one work-group -- enqueue several GEMMs)
SINGLE_KERNEL = ('propagate ', 'chunks ', 'small propagator', 'zgemm ')
# launches whose work is one kind of kernel repeated (the split propagator:
# three Horner GEMMs and the squarings on the FP32 matrix cores, beside
# element-wise kernels a hundred times shorter)
MFMA_LAUNCHES = ('split propagator',)
# kernel behind each single-kernel label, as rocprofv3 names it
KERNEL_NAMES = {'chunks': 'chunk_kernel', 'propagate powers': 'small_powers_apply_kernel',
                'small propagator': 'small_program_kernel', 'zgemm': 'zgemm_kernel',
                'propagate': 'step_kernel', 'response buckets': 'response_buckets_kernel',
                'split propagator': 'split_gemm_kernel', 'heom': 'heom_stage_kernel'}


def dominant_kernel(summary, relaunch, timing, per):
    """The single-kernel launch with the largest total device time per
    spectrum (``per`` spectra in the summary); its duration is measured back
    to back from a captured graph (its own time, without host gaps)."""
    best = None
    for label, info in summary.items():
        if not label.startswith(SINGLE_KERNEL + MFMA_LAUNCHES) or label not in relaunch:
            continue
        ms = timing.back_to_back_ms(relaunch[label])
        total = ms * info['count'] / per
        if best is None or total > best[3]:
            best = (label, info, ms, total)
    return best


def _gemms_power(e):
    """Products of binary powering to exponent e >= 1."""
    e = int(e)
    return (e.bit_length() - 1) + (bin(e).count('1') - 1)


def launch_flops(label):
    """Algorithmic flops of a launch from its label (8 real flops per complex
    multiply-add):
      propagate (stepper): 8 M^2 per matrix-vector product, 4 per RK4 step
        or 1 per propagator step, R B right-hand sides x `steps`;
      chunks: per problem nc chunks of `steps` products plus the nc - 1
        products of the chain of chunk starts (each work-group recomputing
        its own start is not counted as work);
      small propagator: 8 M^3 per product: Horner's three, the powering to
        q, to c;
      propagate powers: Horner's three and the squarings (8 M^3 each), and
        8 M^2 R per state product of every set bit of every step count;
      zgemm: 8 m n k per batch entry."""
    f = dict(re.findall(r'(\w+)=([\w,]+)', label))
    if label.startswith('zgemm'):
        return 8.0 * int(f['m']) * int(f['n']) * int(f['k']) * int(f['batch'])
    M, B = int(f['M']), int(f['B'])
    if label.startswith('split propagator'):
        # three Horner GEMMs and one squaring per doubling of q
        return 8.0 * M ** 3 * B * (3 + int(f['q']).bit_length() - 1)
    if label.startswith('small propagator'):
        q, c = int(f['q']), int(f['c'])
        n = 3 + _gemms_power(q) + (_gemms_power(c) if c > 0 else 0)
        return 8.0 * M ** 3 * B * n
    if label.startswith('propagate powers'):
        ex = [int(x) for x in f['e'].split(',')]
        sq = max(ex).bit_length() - 1 if max(ex) > 0 else 0
        apply = sum(bin(x).count('1') for x in ex)
        return 8.0 * B * (M ** 3 * (3 + sq) + M * M * int(f['R']) * apply)
    steps = int(f['steps'])
    if label.startswith('chunks'):
        nc = int(f['nc'])
        products = nc * steps + (nc - 1)
    else:
        products = int(f['R']) * steps * (4 if ' rk4 ' in label else 1)
    return 8.0 * M * M * B * products


def kernel_bound(label):
    """(bound, peak TFLOP/s) of a kernel: the FP64 GEMM of enough tiles runs
    on the FP64 matrix cores; every other kernel here computes on the vector
    ALUs (FP32 for MIXED stepping, FP64 otherwise) -- the MI355X vector peaks
    equal the dense MFMA ones (157.3 FP32, 78.6 FP64)."""
    f = dict(re.findall(r'(\w+)=([\w,]+)', label))
    if label.startswith('zgemm'):
        tiles32 = -(-int(f['m']) // 32) * -(-int(f['n']) // 32) * int(f['batch'])
        fp64 = ' fp64 ' in label
        if fp64 and tiles32 >= 192 and int(f['m']) >= 16 and int(f['n']) >= 16:
            return 'mfma', PEAK_TFLOPS['fp64']
        return 'valu', PEAK_TFLOPS['fp64' if fp64 else 'mixed']
    if label.startswith(('small propagator', 'propagate powers')):
        return 'valu', PEAK_TFLOPS['fp64']
    if label.startswith('split propagator'):
        return 'mfma', PEAK_TFLOPS['mixed']
    return 'valu', PEAK_TFLOPS['fp64' if ' fp64 ' in label else 'mixed']


def host_info():
    """The host the CPU baseline runs on."""
    model = None
    try:
        with open('/proc/cpuinfo') as fh:
            for line in fh:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    from threadpoolctl import threadpool_info
    blas = [dict(api=p.get('internal_api'), threads=p.get('num_threads'))
            for p in threadpool_info()]
    return dict(nproc=os.cpu_count(), sched_getaffinity=affinity, cpu_model=model, blas=blas)


def _cpu_sample(model, work, per_path, budget, min_seconds):
    """Whole pathways of the 2-D spectrum through the oracle's zvode path,
    repeated until min_seconds of work (or the budget is spent)."""
    from oracle import propagate as P
    done_steps, elapsed, runs = 0, 0.0, 0
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
    return done_steps, elapsed, runs


def cpu_baseline(budget, steps_per_spectrum, min_seconds=10.0):
    """The oracle's zvode restatement of the reference path (scipy zvode at
    its default rtol 1e-6, serial loop over t1 branches, Liouvillian built
    outside the timed part), timed (i) on every host core with numpy's
    default BLAS threading -- the reported value -- and (ii) on one thread.
    Each sample is the whole 2-D spectrum, repeated until ``min_seconds``
    of work; if a spectrum would exceed ``budget``, whole pathways are run
    until the budget is spent and the rate is taken over their steps."""
    from threadpoolctl import threadpool_limits
    from oracle import system as S, liouville as Lv
    from qspectra_amd import workloads as W
    H1, d, extra = W.aggregate7_hamiltonian()
    ham = S.ElectronicHamiltonian(H1, bath=S.DebyeBath(S.CM_K * 77, 35, 106), dipoles=d,
                                  energy_spread_extra=extra)
    model = Lv.LiouvilleModel(ham, unit_convert=S.CM_FS, secular=False)
    model.evolution_super_operator        # Liouvillian built outside the timed part
    work = W.AGG7_2D
    per_path = steps_per_spectrum // 3
    info = host_info()
    steps_all, sec_all, runs_all = _cpu_sample(model, work, per_path, budget / 2, min_seconds)
    with threadpool_limits(limits=1):
        steps_one, sec_one, runs_one = _cpu_sample(model, work, per_path, budget / 2,
                                                   min_seconds)
    cores = info['sched_getaffinity'] or info['nproc']
    return dict(value=steps_all / sec_all, unit='steps/s', cores=cores, kind='port',
                sample='%d repetitions of the benchmark 2-D spectrum (%d steps of the 2 fs '
                       'grid) through the oracle restatement of the reference path: scipy '
                       'zvode at rtol 1e-6, serial loop over t1 branches, Liouvillian prebuilt; '
                       '%.1f s with numpy default BLAS threading on %s cores'
                       % (runs_all, steps_all, sec_all, cores),
                ms_per_spectrum=1e3 * sec_all * per_path * 3 / steps_all,
                one_thread=dict(value=steps_one / sec_one, cores=1,
                                ms_per_spectrum=1e3 * sec_one * per_path * 3 / steps_one,
                                sample='%d repetitions, %.1f s on one thread'
                                       % (runs_one, sec_one)),
                host=info,
                reference_default_tolerance_rel_err=_reference_default_err())


def _fixed_step_model(ham, rw):
    """An oracle Liouville-space model on the 'ge' space whose time step is
    the absorption workload's 1 fs grid."""
    from oracle import system as S, liouville as Lv

    class FixedStep(Lv.LiouvilleModel):
        time_step = 1.0
    return FixedStep(ham, rw_freq=rw, hilbert_subspace='ge', unit_convert=S.CM_FS)


def cpu_extrapolated(budget=10.0):
    """Extrapolated CPU figures of the two long workloads (too long to run in
    full on the host), each from a bounded sample timed through the oracle
    restatement of the reference path on all host cores and scaled linearly:
      absorption ensemble: members of the 4096 + 1024 set, absorption over
        4096 points at 1 fs through scipy zvode (rtol 1e-6), one member at a
        time as the reference's ensemble loop runs them, Liouvillian built in
        the timed part (the reference rebuilds it per member);
      HEOM: one member's 3060-ADO hierarchy propagated through scipy zvode on
        the sparse generator (built outside the timed part) over as many
        whole 0.5 fs output steps of the 1 ps trajectory as the budget
        allows, scaled to 2000 steps x 256 members."""
    from oracle import system as S, propagate as P, liouville as Lv
    from qspectra_amd import workloads as W
    out = {}
    hams = W.ensemble_absorption_hamiltonians()
    parts = {}
    for name, idx in (('disorder', range(0, 4096, 16)), ('aggregates', range(4096, 5120, 4))):
        done, sec = 0, 0.0
        for i in idx:
            h = hams[i]
            oh = S.ElectronicHamiltonian(h.H1_nonrotating,
                                         bath=S.DebyeBath(S.CM_K * 77, 35, 106), dipoles=h.dipoles)
            a = time.perf_counter()
            P.absorption_spectra(_fixed_step_model(oh, 12450.0), 4096.0, integrator='zvode')
            sec += time.perf_counter() - a
            done += 1
            if sec > budget / 2:
                break
        parts[name] = (done, sec)
    (nd, sd), (na, sa) = parts['disorder'], parts['aggregates']
    total = sd / nd * 4096 + sa / na * 1024
    out['ensemble_absorption'] = {
        'seconds_extrapolated': total, 'steps_per_s': len(hams) * 4095 / total,
        'sample': '%d of the 4096 disorder members (%.1f s) and %d of the 1024 aggregates of '
                  '2..24 sites (%.1f s), absorption over 4096 points through zvode (rtol 1e-6), '
                  'one member at a time; each part scaled by its count' % (nd, sd, na, sa)}
    H1, d, _ = W.aggregate7_hamiltonian()
    oh = S.ElectronicHamiltonian(H1, bath=S.DebyeBath(S.CM_K * 300, 35, 106),
                                 dipoles=d).in_rotating_frame()
    L = Lv.heom_evolve_matrix(oh, 'ge', 'ee', S.CM_FS, K=1, level_cutoff=5)
    y0 = np.zeros(L.shape[0], complex)
    y0[0] = 1.0
    f = lambda t, v: L @ v
    # a short run sizes the timed one (saved every 0.5 fs, as the trajectory)
    a = time.perf_counter()
    P.integrate_zvode(f, y0, 0.5 * np.arange(11))
    rate = 10 / (time.perf_counter() - a)
    steps = int(min(2000, max(10, rate * budget)))
    a = time.perf_counter()
    P.integrate_zvode(f, y0, 0.5 * np.arange(steps + 1))
    sec = time.perf_counter() - a
    out['heom'] = {'seconds_extrapolated': sec / steps * 2000 * 256,
                   'hierarchy_steps_timed': steps, 'seconds_timed': sec,
                   'hierarchy_steps_per_s': steps / sec,
                   'sample': '%d output steps of 0.5 fs of one 3060-ADO hierarchy through zvode '
                             '(rtol 1e-6) on the sparse generator, %.1f s; x 2000 steps x 256 '
                             'members' % (steps, sec)}
    return out


def _reference_default_err():
    try:
        return float(np.load(REFERENCE_2D, allow_pickle=False)['default_tol_rel_err'])
    except (OSError, KeyError):
        return None


# HEOM bytes per state element and RK4 step.  Algorithmic: four stages each
# reading and writing the stage operand (16 B FP64, 8 B MIXED) plus one read
# and one write of the FP64 state: 160 / 96 B.  The stage kernel as written
# moves more (no accumulator buffer): FP64 reads y in stages 2..4 and the
# stage vectors z2, z3 again in stage 4 (48 + 48 + 48 + 64 = 208 B); MIXED
# writes the stage derivatives k1..k3 in complex64 and reads them, the FP64
# state and the complex64 copy of y back (24 + 32 + 32 + 72 = 160 B).
HEOM_ALGORITHMIC_BYTES = {'fp64': 160, 'mixed': 96}
HEOM_BYTES_PER_ELEMENT_STEP = {'fp64': 208, 'mixed': 160}

# HBM bytes per launch (read + written) measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE over the `bench.py --counters` sequence, FETCH
# scaled by 2.0 and WRITE by 1.0: the factors that make a 256 MiB device copy
# read and write 256 MiB.  WRITE x 1.0 is also what the kernels with exactly
# known outputs show (the response signals, 5120 x 4096 x 16 B, and every HEOM
# stage's 16 B per element read back to the byte).  Keyed by launch label; a
# launch of several kernels counts all of them.
PMC_TRAFFIC = {
    # one RK4 stage over 256 members, summed over its sample passes (10 of 24
    # members and one of 16; FP64 70.8 B per element-stage, MIXED 50.4 B)
    'heom stage fp64': 2.718e9,
    'heom stage mixed': 1.936e9,
    # the headline plan's t2 operator program (one work-group, M = 49)
    'propagate powers M=49 B=1 R=7 e=0,100,200,500': 8.64e4,
    'chunks prop mixed M=147 nc=8 B=1 steps=15 nsave=16': 4.30e6,
    # the 256-member ensemble plan
    'propagate powers M=49 B=256 R=7 e=0,100,200,500': 1.708e7,
    'propagate prop mixed M=147 R=1 B=256 steps=127 nsave=128': 1.235e8,
    # prep 1.366e8 + three Horner GEMMs 9.716e8 each + squaring 5.307e8
    # + diagonal 1.2e6 + final split 9.25e7
    'split propagator M=147 B=256 q=2': 3.676e9,
    # the 5120-system absorption ensemble, one launch over every wavefront
    # (eight steps per pass): P^8 and projection rows 15.0 MB in, signals
    # 335.6 MB out
    'response buckets B=5120 waves=1800 nt=4096': 3.506e8,
}


def launch_traffic(label):
    """Measured HBM bytes of a launch (None if not profiled)."""
    return PMC_TRAFFIC.get(label)


def _host(x):
    return x.cpu().numpy() if hasattr(x, 'cpu') else np.asarray(x)


def _sync_time(torch, fn):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0, out


def batched_2d_leg(torch, members=256, replays=10):
    """The batched 2-D path: the ensemble-mean 2-D spectrum of `members`
    static-disorder members (sigma = 100 cm^-1) of the 7-site aggregate on
    the headline's grid, one captured graph per ensemble spectrum, end to end
    from the members' Hamiltonians resident in HBM: the Liouvillian blocks
    (eigensystems, Redfield tensors, basis changes), the RK4 propagators
    (MIXED: the 147-element block's split on the FP32 matrix cores), the
    trajectories, the contraction over members and both Fourier transforms.
    FP64 and MIXED; error of the MIXED spectrum against the FP64 one."""
    from qspectra_amd import workloads as W, timing
    from qspectra_amd.simulate.fused import ThirdOrderPlan
    work = W.AGG7_2D
    m = W.aggregate7_model(secular=False, disorder_sigma=100.0)
    mb = m.with_hamiltonians(list(m.hamiltonian.sample_ensemble(members)))
    steps = members * W.rk4_steps_2d()
    out = {'workload': '2-D spectrum (GSB+ESE+ESA) ensemble mean over %d static-disorder '
                       'members (sigma 100 cm^-1) of the 7-site aggregate, 128 x 4 x 128 grid, '
                       'RK4 at 1 fs; end to end from the Hamiltonians in one graph' % members,
           'members': members, 'steps_per_ensemble_spectrum': steps}
    res, plans = {}, {}
    for prec in ('mixed', 'fp64'):
        plan = ThirdOrderPlan(mb, work['coherence_time_max'],
                              population_times=work['population_times'],
                              geometry=work['geometry'], include_signal=work['include_signal'],
                              dt=work['dt'], precision=prec, fourier=True, ensemble_mean=True,
                              rebuild=True).capture()
        for _ in range(2):
            plan.run()
        t, X = time_leg(torch, plan.run, replays)
        plan.check_status()
        res[prec] = (t / replays, X.cpu().numpy())
        plans[prec] = plan
    X32, X64 = res['mixed'][1], res['fp64'][1]
    out['rel_err_mixed_vs_fp64'] = float(np.abs(X32 - X64).max() / np.abs(X64).max())
    for prec in ('mixed', 'fp64'):
        out[prec] = {'ms_per_ensemble_spectrum': 1e3 * res[prec][0],
                     'value': steps / res[prec][0], 'unit': 'steps/s'}
    with timing.kernel_timer() as kt:
        for _ in range(3):
            plans['mixed']._launch()
    summary = kt.summary()
    label, info, launch_ms, per = dominant_kernel(summary, kt.relaunch, timing, 3)
    flops = launch_flops(label)
    bound, peak = kernel_bound(label)
    achieved = flops / (launch_ms * 1e-3) / 1e12
    out['roofline'] = {'bound': bound, 'kernel': label, 'achieved': achieved, 'peak': peak,
                       'unit': 'TFLOP/s', 'frac': achieved / peak, 'launch_ms': launch_ms,
                       'kernel_ms_per_spectrum': per, 'algorithmic_flop': flops,
                       'traffic': launch_traffic(label)}
    out['kernels'] = {k: {'launches': v['count'] / 3, 'mean_ms': round(v['mean_ms'], 5)}
                      for k, v in summary.items()}
    out['device_ms_sum_per_spectrum'] = sum(v['total_ms'] for v in summary.values()) / 3
    return out


def ensemble_absorption_leg(torch, replays=5):
    """4096 static-disorder members (sigma = 100 cm^-1) of the 7-site aggregate
    plus 1024 aggregates of 2..24 sites, Redfield 'ge', absorption over 4096
    points at 1 fs, end to end from the Hamiltonians resident in HBM: one
    captured graph holds, per size class (concurrent streams), the Redfield
    builds and propagators of every member and one size-bucketed propagation
    launch, then the Fourier transform of all 5120 signals; the spectra stay
    on the device.  Beside it, the per-bucket
    path (absorption_batch: one build and one propagation launch per site
    count, host work between them) in FP64 and MIXED."""
    from qspectra_amd import workloads as W, timing, _native
    from qspectra_amd.batched import absorption_batch, AbsorptionEnsemblePlan
    hams = W.ensemble_absorption_hamiltonians()
    rw, tmax = 12450.0, 4096.0
    t0 = time.perf_counter()
    plan = AbsorptionEnsemblePlan(hams, tmax, rw_freq=rw, dt=1.0)
    build_s = time.perf_counter() - t0
    plan.run()
    tp, Xp = time_leg(torch, plan.run, replays)
    plan.check_status()
    Xp = Xp.cpu().numpy()
    for prec in ('mixed', 'fp64'):          # warm-up: schedule tuning, allocations
        absorption_batch(hams, tmax, rw_freq=rw, dt=1.0, precision=prec)
    res = {}
    for prec in ('mixed', 'fp64'):
        res[prec] = _sync_time(torch, lambda: absorption_batch(hams, tmax, rw_freq=rw, dt=1.0,
                                                               precision=prec))
    X32, X64 = res['mixed'][1][1], res['fp64'][1][1]
    err = np.max(np.abs(X32 - X64), axis=1) / np.max(np.abs(X64), axis=1)
    errp = np.max(np.abs(Xp - X64), axis=1) / np.max(np.abs(X64), axis=1)
    steps = len(hams) * (int(tmax) - 1)
    nt = len(plan.t)
    # per-launch durations with the launches one after another on one stream
    # (in the graph the size classes overlap)
    plan._launch(sequential=True)
    torch.cuda.synchronize()
    with timing.kernel_timer() as kts:
        for _ in range(2):
            plan._launch(sequential=True)
    summary = kts.summary()
    # each launch's duration from back-to-back replays of it (events around
    # single eager launches have read tens of milliseconds for sub-millisecond
    # kernels here); the event mean where a launch cannot be repeated
    per_launch = {k: (timing.back_to_back_ms(kts.relaunch[k]) if k in kts.relaunch
                      else v['mean_ms']) for k, v in summary.items()}
    # the graph propagates each size class in its own launch, concurrent with
    # the other classes' builds; the kernel's roofline is taken on one launch
    # over every wavefront, alone on the stream
    with timing.kernel_timer() as kt:
        for _ in range(2):
            _native.response_buckets_into(plan.waves, plan.offsets, plan.P, plan.y0, plan.W,
                                          plan.signal, plan.status, J=plan.J)
    alone = kt.summary()
    label = [k for k in alone if k.startswith('response buckets')][0]
    ms = alone[label]['mean_ms']
    # executed: J = plan.J steps per pass, each pass M rows of P^J and J
    # projection rows of M entries per problem
    J = plan.J
    flops = sum(8.0 * bk['B'] * (bk['M'] ** 2 + J * bk['M']) * -(-nt // J) for bk in plan.buckets)
    achieved = flops / (ms * 1e-3) / 1e12
    return {'workload': 'absorption of 4096 disorder members (sigma 100 cm^-1) of the 7-site '
                        'aggregate + 1024 aggregates of N ~ U{2..24} sites, Redfield ge, 4096 '
                        'points at 1 fs',
            'rk4_steps': steps, 'value': steps / (tp / replays), 'unit': 'steps/s',
            'scope': 'end to end from the Hamiltonians in HBM: one captured graph',
            'ms_per_ensemble': 1e3 * tp / replays, 'plan_build_seconds': build_s,
            'rel_err_vs_per_bucket_fp64_max': float(errp.max()),
            'buckets': len(plan.buckets), 'propagation_waves': int(plan.waves.shape[0]),
            'propagation_kernel': {'label': label, 'ms': ms, 'algorithmic_flop': flops,
                                   'achieved_TFLOPs': achieved, 'bound': 'valu',
                                   'peak_TFLOPs': PEAK_TFLOPS['fp64'],
                                   'frac': achieved / PEAK_TFLOPS['fp64'],
                                   'traffic': launch_traffic(label)},
            'kernels': {k: {'launches': v['count'] / 2, 'mean_ms': round(per_launch[k], 5)}
                        for k, v in summary.items()},
            'per_bucket_path': {'seconds_mixed': res['mixed'][0], 'seconds_fp64': res['fp64'][0],
                                'value_mixed': steps / res['mixed'][0],
                                'value_fp64': steps / res['fp64'][0],
                                'rel_err_mixed_vs_fp64_max': float(err.max()),
                                'rel_err_mixed_vs_fp64_median': float(np.median(err))}}


def gradient_leg(torch):
    """Gradient of sum_w (S - S_target)^2 for the ensemble-averaged absorption
    spectrum (512 disorder members, 4096 RK4 steps at 1 fs, checkpoints every
    64 steps) with respect to the 7 site energies, 21 couplings and lambda;
    checked against central differences (h = 1e-3 cm^-1) of the FP64
    forward pass."""
    from qspectra_amd import gradients as G, workloads as W
    from qspectra_amd.constants import CM_K, GAUSSIAN_SD_FWHM
    (E0, J0, l0), (E1, J1, l1) = W.gradient_target_parameters()
    _, d, _ = W.aggregate7_hamiltonian()
    spec = G.RedfieldAbsorption(d, 106.0, CM_K * 77, 12450.0, 4096.0, 1.0,
                                disorder=100.0 / GAUSSIAN_SD_FWHM, ensemble_size=512, ckpt=64)
    dev = torch.device('cuda')
    mk = lambda x, g=False: torch.tensor(np.asarray(x, float), device=dev, requires_grad=g)
    with torch.no_grad():
        target = spec(mk(E1), mk(J1), mk(l1))[1]

    def loss(E, J, lam):
        return ((spec(E, J, lam)[1] - target) ** 2).sum()

    def fwd_bwd():
        th = [mk(E0, True), mk(J0, True), mk(l0, True)]
        loss(*th).backward()
        return torch.cat([th[0].grad, th[1].grad, th[2].grad[None]])

    fwd_bwd()
    t_fb, grad = _sync_time(torch, fwd_bwd)
    t_f, _ = _sync_time(torch, lambda: loss(mk(E0), mk(J0), mk(l0)).item())
    theta = [mk(E0), mk(J0), mk(l0)]
    fd, h = [], 1e-3
    with torch.no_grad():
        for w, tt in enumerate(theta):
            for i in range(tt.numel()):
                vals = []
                for sgn in (1.0, -1.0):
                    th = [x.clone() for x in theta]
                    th[w].reshape(-1)[i] += sgn * h
                    vals.append(float(loss(*th)))
                fd.append((vals[0] - vals[1]) / (2 * h))
    fd = np.array(fd)
    g = grad.cpu().numpy()
    steps = 512 * 4095
    return {'workload': 'gradient of the squared spectral error of a 512-member ensemble '
                        'absorption spectrum (7-site Redfield ge, 4096 RK4 steps at 1 fs, '
                        'checkpoints every 64) w.r.t. 7 E, 21 J, lambda',
            'seconds_forward_backward': t_fb, 'seconds_forward': t_f,
            'value': 3 * steps / t_fb, 'unit': 'steps/s (forward + reverse, 3 per step)',
            'grad_rel_err_vs_central_fd': float(np.abs(g - fd).max() / np.abs(fd).max()),
            'fd_step_cm': h, 'parameters': int(g.size)}


def heom_leg(torch):
    """HEOM dynamics of the 7-site aggregate (300 K, K = 1, depth 4: 3060
    ADOs, 'ee' block, M = 49), 256 disorder members, 1 ps at 0.5 fs from a
    site-1 excitation, FP64 and mixed."""
    from qspectra_amd import workloads as W
    from qspectra_amd.simulate.utils import propagate_heom
    m = W.heom_aggregate7_model()
    members = [m.hamiltonian.sample(n) for n in range(256)]
    eom = m.with_hamiltonians(members).equation_of_motion('ee')
    y0 = torch.zeros((256, eom.size), dtype=torch.complex128, device='cuda')
    y0[:, 0] = 1.0                                # |1><1|: element (0, 0) of ADO 0
    t = np.arange(0, 1000.0 + 1e-9, 10.0)
    propagate_heom(eom, y0, t[:2], dt=0.5)
    res = {}
    for prec in ('fp64', 'mixed'):
        res[prec] = _sync_time(torch, lambda: propagate_heom(eom, y0, t, dt=0.5,
                                                             precision=prec))
    r64, r32 = res['fp64'][1], res['mixed'][1]
    err = float((r32 - r64).abs().max() / r64.abs().max())
    S = eom.size
    steps = 2000
    out = {'workload': 'HEOM 7-site aggregate, 300 K, K=1, depth 4 (3060 ADOs x 49), 256 '
                       'disorder members, 2000 RK4 steps at 0.5 fs',
           'ados': int(eom.n_ado), 'state_elements': int(S),
           'hierarchy_steps': 256 * steps, 'rel_err_mixed_vs_fp64': err}
    for prec in ('fp64', 'mixed'):
        sec = res[prec][0]
        algo = HEOM_ALGORITHMIC_BYTES[prec]
        gbs = 256 * S * algo * steps / sec / 1e9
        per_launch = 256 * S * algo / 4
        traffic = PMC_TRAFFIC['heom stage ' + prec]
        out[prec] = {'seconds': sec, 'hierarchy_steps_per_s': 256 * steps / sec,
                     'achieved_GBps': gbs, 'hbm_frac': gbs / 8000.0,
                     'algorithmic_bytes_per_element_step': algo,
                     'kernel_bytes_per_element_step': HEOM_BYTES_PER_ELEMENT_STEP[prec],
                     'kernel_GBps': gbs * HEOM_BYTES_PER_ELEMENT_STEP[prec] / algo,
                     'traffic_bytes_per_stage_launch': traffic,
                     'algorithmic_bytes_per_stage_launch': per_launch,
                     'traffic_over_algorithmic': traffic / per_launch}
    return out


def pump_leg(torch):
    """simulate_pump of 1024 static-disorder members (sigma = 100 cm^-1) of
    the 7-site aggregate in 'gg,ge,eg,ee' (M = 64): Gaussian pump (FWHM 20 fs
    at 12450 cm^-1, peak field 0.01 fs^-1), output every 2 fs up to 1 ps
    after the pulse, at the default RK4 step; FP64 and mixed; end to end
    from the Hamiltonians (Liouvillians, dipole maps, field samples, the
    driven kernel)."""
    import qspectra_amd as qa
    from qspectra_amd import workloads as W, timing
    m = W.aggregate7_model(disorder_sigma=100.0)
    pump = qa.GaussianPulse(12450., 20., scale=0.01, freq_convert=qa.CM_FS)
    kw = dict(time_extra=1000.0, ensemble_size=1024)
    qa.simulate_pump(m, pump, 'x', **kw)               # warm-up
    res, kern = {}, {}
    for prec in ('fp64', 'mixed'):
        with timing.kernel_timer() as kt:
            res[prec] = _sync_time(torch, lambda: qa.simulate_pump(m, pump, 'x', precision=prec,
                                                                   **kw))
        # the main launch: the driven launch with the most steps (the others
        # are the step calibration's probe runs)
        driven = [(k, v) for k, v in kt.summary().items() if k.startswith('driven')]
        label, info = max(driven, key=lambda kv: int(re.search(r'steps=(\d+)', kv[0]).group(1)))
        kern[prec] = (label, info['mean_ms'])
    y64, y32 = res['fp64'][1][1], res['mixed'][1][1]
    err = float(np.abs(y32 - y64).max() / np.abs(y64).max())
    steps = int(re.search(r'steps=(\d+)', kern['fp64'][0]).group(1))
    # accuracy of the default step: 16 members against a 0.05 fs step
    small = dict(kw, ensemble_size=16)
    _, yd = qa.simulate_pump(m, pump, 'x', **small)
    _, yf = qa.simulate_pump(m, pump, 'x', dt=0.05, **small)
    yd, yf = _host(yd), _host(yf)
    out = {'workload': 'simulate_pump, 1024 disorder members of the 7-site aggregate, '
                       "'gg,ge,eg,ee' (M = 64), Gaussian pump FWHM 20 fs, 1 ps after the pulse "
                       'at 2 fs output, default RK4 step',
           'rk4_steps_per_member': steps, 'rel_err_mixed_vs_fp64': err,
           'rel_err_default_step_vs_0.05fs_16_members':
               float(np.abs(yd - yf).max() / np.abs(yf).max())}
    for prec in ('fp64', 'mixed'):
        sec = res[prec][0]
        out[prec] = {'seconds_end_to_end': sec, 'kernel': kern[prec][0],
                     'kernel_ms': kern[prec][1],
                     'steps_per_s_end_to_end': 1024 * steps / sec,
                     'steps_per_s_kernel': 1024 * steps / (kern[prec][1] * 1e-3)}
    return out


def example_dimer_2d_leg(torch):
    """two_dimensional_spectra of the reference's example dimer (Redfield and
    HEOM, g/e/f, 1 ps coherence times, 50 population times to 1 ps, GSB +
    ESE + ESA) at the default step; FP64 and mixed; end to end from the
    models."""
    import qspectra_amd as qa
    from qspectra_amd import workloads as W
    red, heom = W.example_dimer_models()
    out = {'workload': "2-D spectra of the reference's HEOM / Redfield example dimer: "
                       "g/e/f, t1 = t3 = 1 ps at the model step, 50 population times to 1 ps, "
                       "GSB+ESE+ESA; HEOM level_cutoff 3, K 1"}
    for name, model in (('redfield', red), ('heom', heom)):
        res = {}
        for prec in ('fp64', 'mixed'):
            qa.two_dimensional_spectra(model, precision=prec, **W.EXAMPLE_DIMER_2D)   # warm-up
            res[prec] = _sync_time(torch, lambda: qa.two_dimensional_spectra(
                model, precision=prec, **W.EXAMPLE_DIMER_2D))
        X64, X32 = res['fp64'][1][1], res['mixed'][1][1]
        out[name] = {'seconds_fp64': res['fp64'][0], 'seconds_mixed': res['mixed'][0],
                     'shape': list(X64.shape),
                     'rel_err_mixed_vs_fp64': float(np.abs(X32 - X64).max() / np.abs(X64).max())}
    return out


def zofe_fmo_leg(torch):
    """The reference's FMO ZOFE example: dynamics in the one-exciton space
    from site 1 for 1 ps, a Gaussian pump (FWHM 50 fs) in g + e, and the
    absorption spectrum over 2 ps, at the default step (FP64)."""
    import qspectra_amd as qa
    from qspectra_amd import workloads as W
    me, mge = W.fmo_zofe_model('e'), W.fmo_zofe_model('ge')
    pump = qa.GaussianPulse(12500, 50, scale=1e-3, freq_convert=qa.CM_FS)
    calls = {
        'dynamics_e_1ps': lambda: qa.simulate_dynamics(me, [1, 0, 0, 0, 0, 0, 0], 1000),
        'pump_ge': lambda: qa.simulate_pump(mge, pump, 'x', 1000),
        'absorption_ge_2ps': lambda: qa.absorption_spectra(mge, 2000),
    }
    out = {'workload': "the reference's FMO ZOFE example (7 sites, 16 pseudomodes): "
                       "dynamics 1 ps ('e'), pump ('ge'), absorption 2 ps ('ge'); default step"}
    for name, fn in calls.items():
        fn()                                                  # warm-up
        out[name + '_seconds'] = _sync_time(torch, fn)[0]
    return out


def jonas_leg(torch):
    """The reference's vibronic examples end to end at the default step:
    Jonas dimer dynamics (1 ps from site 1 with no vibrational quanta),
    absorption (10 ps), the third-order GSB response with the exact
    isotropic average (t1 = t3 = 1 ps, 100 population times to 1 ps; FP64
    and mixed), the monomer's absorption, and the unitary vibronic monomer
    (dynamics 1 ps, absorption 10 ps with 1 ps decay)."""
    import qspectra_amd as qa
    from qspectra_amd import workloads as W
    dimer, monomer, unitary = W.jonas_models()
    e0 = np.zeros(8)
    e0[0] = 1
    u0 = np.zeros(5)
    u0[0] = 1
    calls = {
        'dimer_dynamics_1ps': lambda: qa.simulate_dynamics(dimer, e0, 1000),
        'dimer_absorption_10ps': lambda: qa.absorption_spectra(dimer, 10000),
        'dimer_gsb_isotropic_fp64': lambda: qa.third_order_response(dimer, **W.JONAS_GSB),
        'dimer_gsb_isotropic_mixed': lambda: qa.third_order_response(
            dimer, precision='mixed', **W.JONAS_GSB),
        'monomer_absorption_10ps': lambda: qa.absorption_spectra(monomer, 10000),
        'unitary_dynamics_1ps': lambda: qa.simulate_dynamics(unitary, u0, 1000),
        'unitary_absorption_10ps': lambda: qa.absorption_spectra(
            unitary, 10000, correlation_decay_time=1000),
    }
    out = {'workload': "the reference's vibronic examples (Jonas dimer: 2 sites x 2 two-level "
                       "modes, Redfield g/e/f; vibronic monomers), default step"}
    res = {}
    for name, fn in calls.items():
        # first call: eigensystems, Liouvillian builds and plan set-up included
        out[name + '_first_call_seconds'] = _sync_time(torch, fn)[0]
        out[name + '_seconds'], res[name] = _sync_time(torch, fn)
    X64 = res['dimer_gsb_isotropic_fp64'][1]
    X32 = res['dimer_gsb_isotropic_mixed'][1]
    out['gsb_shape'] = list(X64.shape)
    out['gsb_rel_err_mixed_vs_fp64'] = float(np.abs(X32 - X64).max() / np.abs(X64).max())
    return out


LEGS = (('batched_2d', batched_2d_leg),
        ('ensemble_absorption', ensemble_absorption_leg), ('adjoint_gradient', gradient_leg),
        ('heom', heom_leg), ('pump', pump_leg), ('example_dimer_2d', example_dimer_2d_leg),
        ('zofe_fmo', zofe_fmo_leg), ('jonas_vibronic', jonas_leg))


def main(argv=None):
    args = parse_args(argv)
    if args.gpus != 1:
        print('bench.py measures one GPU: the propagation runs on a single MI355X (across '
              'GPUs it would shard by ensemble member). Use --gpus 1.', file=sys.stderr)
        return 2
    import torch
    import qspectra_amd as qa
    from qspectra_amd import workloads as W, timing
    qa.require_native()
    if args.counters:
        counter_sequence(torch, qa)
        return 0
    model = W.aggregate7_model(secular=False)
    work = W.AGG7_2D
    steps_per_spectrum = W.rk4_steps_2d()                  # steps of the 2 fs grid
    rk4_per_spectrum = W.rk4_steps_2d(dt=work['dt'])       # RK4 steps of 1 fs
    ref = np.load(REFERENCE_2D, allow_pickle=False)

    fused = FusedSpectrum(model, work, 'mixed')
    for _ in range(max(args.warmup, 1)):
        run_plain(qa, model, work)
        fused()
    fused.plan.check_status()
    f1, _, f3 = fused.plan.frequencies()
    assert np.allclose(f1, ref['f1']) and np.allclose(f3, ref['f3'])

    t_plain, X64 = time_leg(torch, lambda: run_plain(qa, model, work), args.steps)
    t_mixed, X32 = time_leg(torch, fused, args.steps)
    fused.plan.check_status()
    X32 = X32.cpu().numpy()

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    # device time of the plain leg: its launches, summed
    with timing.kernel_timer() as kt:
        run_plain(qa, model, work)
    plain_device_ms = sum(v['total_ms'] for v in kt.summary().values())
    # per-kernel durations of the fused plan: the same launch sequence without
    # the graph, with HIP events around every launch on the stream it runs on
    with timing.kernel_timer() as kt:
        for _ in range(args.steps):
            fused.plan._launch()
    summary = kt.summary()
    label, info, launch_ms, kernel_ms_per_spectrum = dominant_kernel(summary, kt.relaunch,
                                                                     timing, args.steps)
    flops = launch_flops(label)
    bound, peak = kernel_bound(label)
    achieved = flops / (launch_ms * 1e-3) / 1e12
    value = steps_per_spectrum * args.steps / t_mixed
    plain = steps_per_spectrum * args.steps / t_plain
    per_spec = {k: {'launches': v['count'] / args.steps, 'mean_ms': round(v['mean_ms'], 5)}
                for k, v in summary.items()}

    line = {
        'metric': METRIC, 'value': value, 'unit': 'steps/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': 1e3 * t_mixed / args.steps, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32 products / f64 state',
        'data': 'synthetic',
        'config': {'workload': '2-D spectrum (GSB+ESE+ESA, -++) of a 7-site aggregate, '
                               'non-secular Redfield g/e/f, t1 x t3 = 128 x 128 at 2 fs, '
                               't2 in {0,100,200,500} fs, all t1 branches as RHS; RK4 at 1 fs',
                   'steps_per_spectrum': steps_per_spectrum, 'grid_step_fs': 2.0,
                   'rk4_step_fs': work['dt'], 'parallelism': 'single GPU'},
        'timed_region': 'one captured graph per spectrum: RK4 propagators, their powers and '
                        'the mixed split rebuilt from the Liouvillian blocks, trajectories, '
                        'contractions, both Fourier transforms',
        'ms_per_spectrum': 1e3 * t_mixed / args.steps,
        'rk4_steps_per_spectrum_reference_algorithm': rk4_per_spectrum,
        'executed_per_spectrum': dict(fused.plan.executed_products(),
                                      launches=sum(v['launches'] for v in per_spec.values()),
                                      device_ms_sum_over_streams=sum(
                                          v['total_ms'] for v in summary.values()) / args.steps),
        'plain_fp64': {'value': plain, 'ms_per_step': 1e3 * t_plain / args.steps,
                       'device_ms_per_spectrum': plain_device_ms,
                       'rel_err_vs_reference': rel(X64, ref['S2d'])},
        'speedup_vs_plain': value / plain,
        'rel_err_vs_fp64': rel(X32, X64),
        'rel_err_vs_reference': rel(X32, ref['S2d']),
        'reference': 'tests/golden/spectra_2d_full.npz: the reference two_dimensional_spectra at '
                     'zvode rtol 1e-12, atol 1e-14',
        'roofline': {'bound': bound, 'kernel': label,
                     'rocprof_kernel': next((v for k, v in sorted(KERNEL_NAMES.items(),
                                                                  key=lambda kv: -len(kv[0]))
                                             if label.startswith(k)), None),
                     'achieved': achieved, 'peak': peak, 'unit': 'TFLOP/s',
                     'frac': achieved / peak, 'traffic': launch_traffic(label),
                     'launch_ms': launch_ms, 'launch_ms_in_sequence': info['mean_ms'],
                     'kernel_ms_per_spectrum': kernel_ms_per_spectrum,
                     'algorithmic_flop': flops},
        'kernels': per_spec,
        'plan_chunk': {'choice': fused.chunk, 'ms_per_spectrum': fused.chunk_ms},
    }
    wanted = ([n for n, _ in LEGS] if args.workloads == 'all' else
              [] if args.workloads == 'none' else args.workloads.split(','))
    unknown = set(wanted) - set(n for n, _ in LEGS)
    if unknown:
        raise SystemExit('unknown workloads: %s' % sorted(unknown))
    if wanted:
        line['workloads'] = {}
        for name, leg in LEGS:
            if name in wanted:
                print('bench: %s' % name, file=sys.stderr, flush=True)
                line['workloads'][name] = leg(torch)
    if args.cpu_baseline == 'auto':
        line['cpu_baseline'] = cpu_baseline(args.cpu_budget, steps_per_spectrum)
        line['vs_cpu'] = value / line['cpu_baseline']['value']
        line['vs_cpu_one_thread'] = value / line['cpu_baseline']['one_thread']['value']
        if wanted:
            print('bench: cpu extrapolations', file=sys.stderr, flush=True)
            line['cpu_baseline']['extrapolated'] = cpu_extrapolated()
    else:
        line['cpu_baseline'] = None
    print(json.dumps(line))
    return 0


if __name__ == '__main__':
    sys.exit(main())
