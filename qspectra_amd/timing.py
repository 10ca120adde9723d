"""Per-launch device timing of libqsx kernels with HIP events.

Inside ``with kernel_timer() as kt:``, every libqsx launch made through
qspectra_amd._native records a start and an end event on the stream it is
launched on; ``kt.summary()`` (after the work completes) gives count, total
and mean milliseconds per launch label.
"""
import contextlib
import gc

import torch

_active = []


class KernelTimes(object):
    def __init__(self):
        self.records = []
        self.relaunch = {}       # label -> callable repeating the last such launch

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
def launch(label, stream=None, relaunch=None):
    """Brackets one libqsx launch; ``relaunch`` (optional) repeats exactly
    that launch (same operands), for back-to-back duration measurements."""
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
        if relaunch is not None:
            kt.relaunch[label] = relaunch


@contextlib.contextmanager
def capture(graph):
    """torch.cuda.graph(graph) with Python's automatic garbage collection
    suspended while the stream captures: a collection inside the capture can
    destroy another, unreachable graph, which HIP refuses while a stream is
    capturing (the destructor then aborts the process).  torch.cuda.graph
    itself collects once before the capture begins."""
    enabled = gc.isenabled()
    gc.disable()
    try:
        with torch.cuda.graph(graph):
            yield
    finally:
        if enabled:
            gc.enable()


def back_to_back_ms(fn, n=20, reps=5):
    """Mean device time of one call of ``fn`` (a single kernel launch) when
    n calls are replayed back to back from a captured HIP graph, so that
    host launch overhead and idle gaps do not enter: median over reps."""
    fn()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with capture(g):
        for _ in range(n):
            fn()
    g.replay()
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        g.replay()
        b.record()
        b.synchronize()
        ms.append(a.elapsed_time(b) / n)
    ms.sort()
    return ms[len(ms) // 2]
