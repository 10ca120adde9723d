"""Host side of the propagation engine: RK4 step schedules and device
buffers for the save grid.

The reference integrates with scipy's adaptive zvode and samples the
solution at the requested output times (qspectra/simulate/utils.py:16-50).
This engine integrates with fixed-step classical RK4 on the GPU.  Interval i
(from the previous output time, or t0, to t[i]) is split into n_i equal
steps of length h_i <= dt; n_i is the nearest integer to interval/dt when
that is within 1e-6 relative (so a grid that is a multiple of dt is hit
exactly), otherwise the next larger integer.  An interval of length zero
saves the current state again, which reproduces the reference's
``y[0] = save_func(y0)`` when ``t[0] == t0``.
"""
import numpy as np
import torch

from . import _native


def rk4_schedule(t, t0=None, dt=None):
    """Returns (nsub int array, h float array), one entry per output time."""
    t = np.asarray(t, dtype=float).reshape(-1)
    if t.size == 0:
        raise ValueError('empty time grid')
    if t0 is None:
        t0 = t[0]
    intervals = np.diff(np.concatenate([[float(t0)], t]))
    if np.any(intervals < 0):
        raise ValueError('output times must be non-decreasing and >= t0')
    if dt is None:
        positive = intervals[intervals > 0]
        dt = float(positive.min()) if positive.size else 1.0
    if not dt > 0:
        raise ValueError('dt must be positive')
    ratio = intervals / dt
    nearest = np.round(ratio)
    steps = np.where(np.abs(ratio - nearest) <= 1e-6 * np.maximum(ratio, 1.0),
                     nearest, np.ceil(ratio))
    steps = np.where(intervals > 0, np.maximum(steps, 1), 0).astype(np.int64)
    h = np.zeros_like(intervals)
    np.divide(intervals, steps, out=h, where=steps > 0)
    return steps, h


def schedule_tensors(t, t0=None, dt=None, dev=None):
    """rk4_schedule uploaded as (int32, float64) device tensors."""
    n, h = rk4_schedule(t, t0, dt)
    if n.max(initial=0) > np.iinfo(np.int32).max:
        raise ValueError('too many RK4 steps in one interval')
    dev = _native.device() if dev is None else dev
    return (torch.as_tensor(n.astype(np.int32)).to(dev),
            torch.as_tensor(h.astype(np.float64)).to(dev))


def total_steps(t, t0=None, dt=None):
    return int(rk4_schedule(t, t0, dt)[0].sum())
