"""Batched spectra of many systems at once (static-disorder ensembles and
aggregates of uneven size).

The reference evaluates an ensemble member by member, rebuilding each
member's Liouvillian (decorators.py:51-63).  ``absorption_batch`` groups
systems by site count; each group is one batched GPU Liouvillian build
(eigensystems, C(w), Redfield blocks) and one propagation launch with every
member a batch entry.  All systems share the rotating frame ``rw_freq`` and
the step ``dt`` so their spectra share one frequency axis.
"""
import numpy as np
import torch

from . import _native
from .constants import CM_FS
from .dynamics.redfield import RedfieldModel
from .simulate.response import _batched_linear_response
from .simulate.utils import fourier_transform_device


def absorption_batch(hamiltonians, time_max, rw_freq, dt, model_cls=RedfieldModel,
                     model_kwargs=None, polarization='xx', precision='fp64', schedule='auto',
                     correlation_decay_time=None, unit_convert=CM_FS):
    """Absorption spectra of ``hamiltonians`` (any mix of site counts).

    Returns (f, X) with X [len(hamiltonians), len(f)], row i the spectrum of
    hamiltonians[i] (Re FT of -S(t), as absorption_spectra)."""
    kw = dict(hilbert_subspace='ge', unit_convert=unit_convert)
    kw.update(model_kwargs or {})
    groups = {}
    for i, h in enumerate(hamiltonians):
        groups.setdefault(h.n_sites, []).append(i)
    t = np.arange(0, time_max, dt)
    out = None
    f = None
    for n, idx in sorted(groups.items()):
        hams = [hamiltonians[i].in_rotating_frame(rw_freq) for i in idx]
        model = model_cls(hams[0], rw_freq=rw_freq, **kw).with_hamiltonians(hams)
        model.time_step_override = dt
        ticks, signal = _batched_linear_response(model, 'gg->eg->gg', time_max, None,
                                                 polarization, dt=dt, precision=precision,
                                                 schedule=schedule)
        x = torch.as_tensor(signal).to(_native.device())
        if correlation_decay_time is not None:
            x = x * torch.as_tensor(np.exp(-ticks / correlation_decay_time)).to(x.device)
        f, X = fourier_transform_device(ticks, -x, -1, rw_freq, unit_convert)
        X = X.real.cpu().numpy()
        if out is None:
            out = np.empty((len(hamiltonians), X.shape[-1]))
        out[idx] = X
    return f, out
