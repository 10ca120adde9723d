"""Step schedules of the propagation engine (exact classical RK4 up to
rounding; P = P4(h L) is the one-step RK4 propagator of a time-invariant L).
"""
import copy
import json
import os

import numpy as np
import torch

from . import _native
from .engine import rk4_schedule
from .lti import propagator_power, transposed
from .timing import capture


def _mode(precision):
    if precision in ('fp64', 'double', None):
        return _native.QSX_FP64
    if precision != 'mixed':
        raise ValueError("precision must be 'fp64' or 'mixed'")
    return _native.QSX_MIXED


def stepper_operator(eom, schedule, h0, precision):
    """Operator (and FP64 diagonal for MIXED) the register stepper applies,
    cached on the equation of motion: L for 'rk4', P4(h0 L) (or P - I for
    MIXED) for 'propagator'."""
    from .lti import rk4_propagator, rk4_increment, split_diagonal
    key = (schedule, float(h0) if schedule == 'propagator' else None, precision)
    cache = eom.__dict__.setdefault('_operators', {})
    if key not in cache:
        if schedule == 'propagator':
            A = (rk4_propagator(eom.L, h0) if _mode(precision) == _native.QSX_FP64
                 else rk4_increment(eom.L, h0))
        else:
            A = eom.L
        cache[key] = (A, None) if _mode(precision) == _native.QSX_FP64 else split_diagonal(A)
    return cache[key]


# 'chunks' in one launch (qsx_propagate_chunks: every work-group computes its
# own chunk start, then steps its chunk) instead of a launch for the chain of
# chunk starts followed by one for the chunks; both give the same values.
FUSED_CHUNKS = True


class Stage(object):
    """One propagation with its launch schedule resident on the device.

    Schedules (all exact classical RK4 up to rounding, with P = P4(hL)):
      'powers'   few save points: Y_s = P^(n_s - n_{s-1}) Y_{s-1}, one batched
                 GEMM per save interval with a cached power (FP64);
      'chunks'   one right-hand side saved at every one of many equally spaced
                 points: chunk starts y_{kc} = (P^c)^k y_0 with P^c, then all
                 chunks stepped in parallel with P, so the sequential chain
                 is n/c + c products instead of n (one launch: work-group k
                 computes its own start, then steps its chunk);
      'doubling' the same trajectory by doubling: rows [y_0 .. y_{m-1}] give
                 rows [y_m .. y_{2m-1}] in one GEMM with the cached power
                 (P^q)^m, so ceil(log2 n) GEMMs (FP64);
      'dchunks'  'chunks' with the chunk starts computed by doubling (powers
                 of P^c, log2(n/c) GEMMs) instead of a stepper chain;
      'stepper'  the register stepper with P (or L for schedule 'rk4').
    With schedule 'auto' a trajectory of one right-hand side picks among
    'doubling', 'chunks' (several chunk lengths) and 'stepper' by timing
    them on the device (schedules.autotune), once per problem shape."""

    def __init__(self, eom, t, t0, dt, precision, schedule, R, chunk=None, kind=None,
                 graph=False):
        n, h = rk4_schedule(t, t0, dt)
        hs = h[n > 0]
        uniform = hs.size == 0 or np.all(np.abs(hs - hs[0]) <= 1e-12 * abs(hs[0]))
        prop = schedule in ('propagator', 'blocked') or (schedule == 'auto' and uniform
                                                        and hs.size)
        if prop and not uniform:
            raise ValueError('the propagator schedules need a uniform step')
        if eom.L.shape[-1] > _native.MAX_REGISTER_BLOCK:
            raise ValueError('block of %d states exceeds the register stepper'
                             % eom.L.shape[-1])
        self.eom = eom
        self.h0 = float(hs[0]) if hs.size else 0.0
        self.n = n
        self.h = h
        self.nsave = len(n)
        self.steps = int(n.sum())
        self.precision = _mode(precision)
        self.prec_name = precision
        self.mode = _native.QSX_MODE_PROPAGATOR if prop else _native.QSX_MODE_RK4
        self.graph = graph
        trajectory = (prop and hs.size and R == 1 and n[0] == 0 and np.all(n[1:] == n[1])
                      and n[1] > 0 and self.nsave > 8)
        if kind is None:
            kind, chunk = 'stepper', chunk
            if prop and schedule in ('auto', 'blocked') and hs.size:
                if self.nsave <= 8:
                    kind = 'powers'
                elif trajectory:
                    kind, chunk = self._choose(schedule, chunk)
        elif kind in ('chunks', 'dchunks', 'doubling') and not trajectory:
            raise ValueError("schedule %r needs one right-hand side saved at equally "
                             "spaced points" % kind)
        self._setup(kind, chunk)

    # ---------------------------------------------------------- selection
    def _candidates(self):
        cands = [('doubling', None)]
        for c in (4, 8, 16, 32, 64):
            if 2 * c < self.nsave:
                cands.append(('chunks', c))
                cands.append(('dchunks', c))
        if self.steps <= 512:
            cands.append(('stepper', None))
        return cands

    def _choose(self, schedule, chunk):
        if chunk is not None:
            return 'chunks', chunk
        default = ('chunks', max(4, int(np.ceil(np.sqrt(self.nsave)))))
        if schedule != 'auto' or not autotune_enabled():
            return default
        M = self.eom.L.shape[-1]
        key = 'traj M=%d B=%d nsave=%d q=%d %s %s' % (
            M, self.eom.L.shape[0], self.nsave, int(self.n[1]), self.prec_name,
            'graph' if self.graph else 'eager')

        def trial(cand):
            st = copy.copy(self)
            st._setup(cand[0], cand[1])
            y0 = torch.ones((self.eom.L.shape[0], 1, M), dtype=torch.complex128,
                            device=self.eom.L.device) / np.sqrt(M)
            return lambda: st.run(y0, [])
        best = autotune(key, self._candidates(), trial, graph=self.graph)
        return best[0], best[1]

    def _setup(self, kind, chunk):
        eom, n, h = self.eom, self.n, self.h
        dev = eom.L.device
        self.kind = kind
        if kind == 'powers':
            self.mats = [None if k == 0 else
                         transposed(eom, ('Pn', self.h0, int(k)),
                                    propagator_power(eom, self.h0, int(k))) for k in n]
        elif kind == 'chunks':
            q = int(n[1])                       # steps per output interval
            self.c = chunk
            self.nc = -(-self.nsave // chunk)
            self.Pc = propagator_power(eom, self.h0, q * chunk)
            self.starts_n = torch.tensor([0] + [1] * (self.nc - 1), dtype=torch.int32,
                                         device=dev)
            self.starts_h = torch.full((self.nc,), self.h0 * q * chunk, dtype=torch.float64,
                                       device=dev)
            self.chunk_n = torch.tensor([0] + [q] * (chunk - 1), dtype=torch.int32, device=dev)
            self.chunk_h = torch.full((chunk,), self.h0, dtype=torch.float64, device=dev)
            self.chunk_steps = q * (chunk - 1)          # products of one chunk
            self.A, self.diag = stepper_operator(eom, 'propagator', self.h0, self.prec_name)
        elif kind == 'dchunks':
            q = int(n[1])
            self.c = chunk
            self.nc = -(-self.nsave // chunk)
            self.pw = []
            m = 1
            while m < self.nc:
                self.pw.append(propagator_power(eom, self.h0, q * chunk * m))
                m *= 2
            self.chunk_n = torch.tensor([0] + [q] * (chunk - 1), dtype=torch.int32, device=dev)
            self.chunk_h = torch.full((chunk,), self.h0, dtype=torch.float64, device=dev)
            self.chunk_steps = q * (chunk - 1)
            self.A, self.diag = stepper_operator(eom, 'propagator', self.h0, self.prec_name)
        elif kind == 'doubling':
            q = int(n[1])
            self.pw = []
            m = 1
            while m < self.nsave:
                self.pw.append(propagator_power(eom, self.h0, q * m))
                m *= 2
        elif kind == 'stepper':
            prop = self.mode == _native.QSX_MODE_PROPAGATOR
            self.A, self.diag = stepper_operator(eom, 'propagator' if prop else 'rk4',
                                                 self.h0, self.prec_name)
            self.nsub = torch.as_tensor(n.astype(np.int32)).to(dev)
            self.hs = torch.as_tensor(h.astype(np.float64)).to(dev)
        else:
            raise ValueError('unknown schedule kind %r' % kind)

    # ---------------------------------------------------------------- run
    def run(self, y0, statuses, status=None):
        """y0 [B, R, M] -> states [B, R, nsave, M] (for 'chunks' a view of the
        padded [B, 1, nc * c, M] output).  status: optional device slot shared
        by the launches (cleared by the caller)."""
        if self.kind == 'powers':
            B, R, M = y0.shape
            outs, y = [], y0
            for A in self.mats:
                if A is not None:
                    y = _native.zgemm(y, A)           # rows: y <- y P^T
                outs.append(y)
            return torch.stack(outs, dim=2)
        if self.kind == 'doubling':
            B, R, M = y0.shape
            out = torch.empty((B, 1, self.nsave, M), dtype=torch.complex128, device=y0.device)
            _doubling(out[:, 0], y0[:, 0], self.pw)
            return out
        if self.kind == 'dchunks':
            B, R, M = y0.shape
            starts = torch.empty((B, self.nc, M), dtype=torch.complex128, device=y0.device)
            _doubling(starts, y0[:, 0], self.pw)
            st1, _, s1 = _native.propagate_states(self.mode, self.A, starts, self.chunk_n,
                                                   self.chunk_h, self.precision, self.diag,
                                                   steps=self.chunk_steps, status=status)
            statuses.append(s1)
            return st1.reshape(B, 1, self.nc * self.c, M)[:, :, :self.nsave]
        if self.kind == 'chunks' and FUSED_CHUNKS:
            B, R, M = y0.shape
            st1, s1 = _native.propagate_chunks(self.Pc, self.A, y0, self.nc, self.chunk_n,
                                               self.chunk_h, self.precision, self.diag,
                                               steps=self.chunk_steps, status=status)
            statuses.append(s1)
            return st1.reshape(B, 1, self.nc * self.c, M)[:, :, :self.nsave]
        if self.kind == 'chunks':
            B, R, M = y0.shape
            st0, _, s0 = _native.propagate_states(_native.QSX_MODE_PROPAGATOR, self.Pc, y0,
                                                   self.starts_n, self.starts_h,
                                                   steps=self.nc - 1, status=status)
            starts = st0.reshape(B, self.nc, M)
            st1, _, s1 = _native.propagate_states(self.mode, self.A, starts, self.chunk_n,
                                                   self.chunk_h, self.precision, self.diag,
                                                   steps=self.chunk_steps, status=status)
            statuses += [s0] if status is not None else [s0, s1]
            return st1.reshape(B, 1, self.nc * self.c, M)[:, :, :self.nsave]
        out, _, st = _native.propagate_states(self.mode, self.A, y0, self.nsub, self.hs,
                                              self.precision, self.diag, steps=self.steps,
                                              status=status)
        statuses.append(st)
        return out


def _doubling(o, y0, powers):
    """Rows of o [B, n, M]: o[:, k] = A^k y0 where powers[j] = A^(2^j):
    rows m..2m-1 are rows 0..m-1 times (A^m)^T, one GEMM per doubling."""
    n = o.shape[1]
    o[:, 0].copy_(y0)
    m = 1
    for Pm in powers:
        r = min(m, n - m)
        _native.zgemm(o[:, :r], Pm, o[:, m:m + r], transb=True)
        m += r


# ---------------------------------------------------------------- autotuner
_TUNED = {}


def autotune_enabled():
    return os.environ.get('QSX_AUTOTUNE', '1') != '0'


def _cache_path():
    return os.environ.get('QSX_TUNE_CACHE',
                          os.path.join(os.path.expanduser('~'), '.cache', 'qspectra_amd',
                                       'schedules.json'))


def _load_cache():
    if _TUNED:
        return
    _TUNED['_loaded'] = True
    try:
        with open(_cache_path()) as f:
            _TUNED.update(json.load(f))
    except (OSError, ValueError):
        pass


def _save_cache():
    try:
        path = _cache_path()
        os.makedirs(os.path.dirname(path), exist_ok=True)
        data = {k: v for k, v in _TUNED.items() if not k.startswith('_')}
        with open(path + '.tmp', 'w') as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(path + '.tmp', path)
    except OSError:
        pass


def autotune(key, candidates, make_trial, reps=5, graph=False, block=1):
    """The fastest of ``candidates`` for problem ``key`` on this device:
    each candidate's launch sequence (make_trial(candidate) -> callable) is
    run once to warm up, then timed ``reps`` times with device events -- as
    eager launches, or (graph=True) as replays of a captured HIP graph, the
    way a plan runs it; each sample is ``block`` calls back to back (the
    steady state of repeated calls rather than one call's start-up
    latency; with block > 1 the candidates are sampled in turn); the median
    decides.  Decisions are cached in memory
    and in a JSON file (QSX_TUNE_CACHE, default
    ~/.cache/qspectra_amd/schedules.json), keyed by device name and problem
    shape."""
    _load_cache()
    full = '%s | %s' % (torch.cuda.get_device_name(), key)
    hit = _TUNED.get(full)
    if hit is not None and tuple(hit['choice']) in [tuple(c) for c in candidates]:
        return tuple(hit['choice'])

    def sample(fn):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(block):
            fn()
        b.record()
        b.synchronize()
        return a.elapsed_time(b) / block

    def prepare(cand):
        fn = make_trial(cand)
        fn()
        torch.cuda.synchronize()
        if graph:
            g = torch.cuda.CUDAGraph()
            with capture(g):
                fn()
            fn = g.replay
            fn()
            torch.cuda.synchronize()
        return fn

    samples = {}
    if block > 1:
        # whole plans (kept alive by the caller anyway): the candidates are
        # sampled in turn, so that a drift of the clocks during tuning does
        # not favour the ones timed last
        fns = {cand: prepare(cand) for cand in candidates}
        for cand in candidates:
            samples[cand] = []
        for _ in range(reps):
            for cand in candidates:
                samples[cand].append(sample(fns[cand]))
    else:
        for cand in candidates:
            fn = prepare(cand)
            samples[cand] = [sample(fn) for _ in range(reps)]
    times = {cand: float(np.median(v)) for cand, v in samples.items()}
    best = min(times, key=times.get)
    _TUNED[full] = {'choice': list(best),
                    'ms': {'%s:%s' % c: round(v, 5) for c, v in times.items()}}
    _save_cache()
    return best
