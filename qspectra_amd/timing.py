"""Per-launch device timing of libqsx kernels with HIP events.

Inside ``with kernel_timer() as kt:``, every libqsx launch made through
qspectra_amd._native records a start and an end event on the stream it is
launched on; ``kt.summary()`` (after the work completes) gives count, total
and mean milliseconds per launch label.
"""
import contextlib

import torch

_active = []


class KernelTimes(object):
    def __init__(self):
        self.records = []

    def summary(self):
        torch.cuda.synchronize()
        out = {}
        for label, a, b in self.records:
            ms = a.elapsed_time(b)
            c, tot = out.get(label, (0, 0.0))
            out[label] = (c + 1, tot + ms)
        return {k: dict(count=c, total_ms=t, mean_ms=t / c) for k, (c, t) in out.items()}


@contextlib.contextmanager
def kernel_timer():
    kt = KernelTimes()
    _active.append(kt)
    try:
        yield kt
    finally:
        _active.remove(kt)


@contextlib.contextmanager
def launch(label, stream=None):
    if not _active:
        yield
        return
    s = torch.cuda.current_stream() if stream is None else stream
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(s)
    yield
    b.record(s)
    for kt in _active:
        kt.records.append((label, a, b))
