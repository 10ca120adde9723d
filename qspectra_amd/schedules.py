"""Step schedules of the propagation engine (exact classical RK4 up to
rounding; P = P4(h L) is the one-step RK4 propagator of a time-invariant L).
"""
import numpy as np
import torch

from . import _native
from .engine import rk4_schedule
from .lti import propagator_power, transposed


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


class Stage(object):
    """One propagation with its launch schedule resident on the device.

    Schedules (all exact classical RK4 up to rounding, with P = P4(hL)):
      'powers'  few save points: Y_s = P^(n_s - n_{s-1}) Y_{s-1}, one batched
                GEMM per save interval with a cached power (FP64);
      'chunks'  one right-hand side saved at every one of many equally spaced
                points: chunk starts y_{kc} = (P^c)^k y_0 by the register
                stepper with P^c, then all chunks stepped in parallel with P,
                so the sequential chain is n/c + c products instead of n;
      'stepper' the register stepper with P (or L for schedule 'rk4')."""

    def __init__(self, eom, t, t0, dt, precision, schedule, R, chunk=None):
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
        self.nsave = len(n)
        self.steps = int(n.sum())
        self.precision = _mode(precision)
        self.prec_name = precision
        dev = eom.L.device
        self.kind = 'stepper'
        if chunk is None:
            chunk = max(4, int(np.ceil(np.sqrt(self.nsave))))
        if prop and schedule in ('auto', 'blocked') and hs.size:
            if self.nsave <= 8:
                self.kind = 'powers'
            elif (R == 1 and n[0] == 0 and np.all(n[1:] == n[1]) and n[1] > 0
                  and self.nsave > 2 * chunk):
                self.kind = 'chunks'
        self.mode = _native.QSX_MODE_PROPAGATOR if prop else _native.QSX_MODE_RK4
        if self.kind == 'powers':
            self.mats = [None if k == 0 else
                         transposed(eom, ('Pn', self.h0, int(k)),
                                    propagator_power(eom, self.h0, int(k))) for k in n]
        elif self.kind == 'chunks':
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
            self.A, self.diag = stepper_operator(eom, 'propagator', self.h0, precision)
        else:
            self.A, self.diag = stepper_operator(eom, 'propagator' if prop else 'rk4',
                                                 self.h0, precision)
            self.nsub = torch.as_tensor(n.astype(np.int32)).to(dev)
            self.h = torch.as_tensor(h.astype(np.float64)).to(dev)

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
        if self.kind == 'chunks':
            B, R, M = y0.shape
            st0, _, s0 = _native.propagate_states(_native.QSX_MODE_PROPAGATOR, self.Pc, y0,
                                                   self.starts_n, self.starts_h,
                                                   steps=self.nc - 1, status=status)
            starts = st0.reshape(B, self.nc, M)
            st1, _, s1 = _native.propagate_states(self.mode, self.A, starts, self.chunk_n,
                                                   self.chunk_h, self.precision, self.diag,
                                                   steps=self.steps, status=status)
            statuses += [s0] if status is not None else [s0, s1]
            return st1.reshape(B, 1, self.nc * self.c, M)[:, :, :self.nsave]
        out, _, st = _native.propagate_states(self.mode, self.A, y0, self.nsub, self.h,
                                              self.precision, self.diag, steps=self.steps,
                                              status=status)
        statuses.append(st)
        return out
