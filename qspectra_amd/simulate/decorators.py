"""Ensemble and isotropic-average decorators (qspectra/simulate/decorators.py).

The reference averages over static disorder with a serial Python loop that
rebuilds every member's Liouvillian (decorators.py:40-64).  Here the
decorated function is called once with a *batched* model: every member's
Liouvillian is built in one GPU launch and all members propagate together,
so wrapped functions return signals with a leading batch axis, which the
ensemble wrapper averages (or drops for a single model).
"""
from functools import wraps
import inspect

import numpy as np

from ..polarization import (check_polarizations, invariant_weights_4th_order,
                            invariant_polarizations, FOURTH_ORDER_INVARIANTS)
from ..utils import ZeroArray


def _get_call_args(func, *args, **kwargs):
    """Arguments of func(*args, **kwargs) as one keyword dict (extra **kwargs
    merged in), so the call can be repeated with modified arguments;
    functions with *args are rejected."""
    sig = inspect.signature(func)
    if any(p.kind == p.VAR_POSITIONAL for p in sig.parameters.values()):
        raise NotImplementedError('%s cannot include positional-only arguments '
                                  '(i.e., of the form *args)' % func)
    bound = sig.bind(*args, **kwargs)
    bound.apply_defaults()
    out = dict(bound.arguments)
    for name, p in sig.parameters.items():
        if p.kind == p.VAR_KEYWORD:
            out.update(out.pop(name, {}))
    return out


def _host(x):
    """numpy view of a result (device tensors are copied to the host)."""
    return x.cpu().numpy() if hasattr(x, 'cpu') else x


def optional_ensemble_average(func):
    """Adds ``ensemble_size`` / ``ensemble_random_orientations``: the first
    ensemble_size samples of the model's Hamiltonian are evaluated as one
    batch and averaged (on the device when the wrapped function returns a
    device tensor, so only the mean crosses to the host)."""
    @wraps(func)
    def wrapper(dynamical_model, *args, **kwargs):
        size = kwargs.pop('ensemble_size', None)
        random_orientations = kwargs.pop('ensemble_random_orientations', False)
        if size is None:
            ticks, signal = func(dynamical_model, *args, **kwargs)
            return ticks, _host(signal[0])
        hams = list(dynamical_model.hamiltonian.sample_ensemble(size, random_orientations))
        ticks, signal = func(dynamical_model.with_hamiltonians(hams), *args, **kwargs)
        total = signal.sum(dim=0) if hasattr(signal, 'cpu') else signal.sum(axis=0)
        return ticks, _host(total / size)
    return wrapper


def optional_4th_order_isotropic_average(func):
    """Adds ``exact_isotropic_average`` for four polarizations: the signal is
    summed over the 9-21 lab polarizations of the tensor invariants with
    non-negligible weight (Hochstrasser), each evaluated once; device
    signals are summed on the device."""
    @wraps(func)
    def wrapper(*args, **kwargs):
        if not kwargs.pop('exact_isotropic_average', False):
            return func(*args, **kwargs)
        kwargs = _get_call_args(func, *args, **kwargs)
        weights = invariant_weights_4th_order(kwargs.pop('polarization'))
        # each lab polarization once, with the summed weight of the
        # invariants it belongs to (decorators.py:67-96 evaluates each once
        # and adds it per invariant)
        pol_weight = {}
        for invariant, w in zip(FOURTH_ORDER_INVARIANTS, weights):
            if w > 1e-8:
                for p in invariant_polarizations(invariant):
                    pol_weight[p] = pol_weight.get(p, 0.0) + w
        total = ZeroArray()
        ticks = None
        for p, w in pol_weight.items():
            ticks, signal = func(polarization=p, **kwargs)
            total += w * signal
        return ticks, total
    return wrapper


def optional_2nd_order_isotropic_average(func):
    """Adds ``exact_isotropic_average`` for two polarizations:
    (xx + yy + zz) * (e1 . e2) / 3."""
    @wraps(func)
    def wrapper(*args, **kwargs):
        if not kwargs.pop('exact_isotropic_average', False):
            return func(*args, **kwargs)
        kwargs = _get_call_args(func, *args, **kwargs)
        pols = check_polarizations(kwargs.pop('polarization'), 2)
        weight = float(np.dot(*pols))
        total = ZeroArray()
        for p in ('xx', 'yy', 'zz'):
            ticks, signal = func(polarization=p, **kwargs)
            total += signal
        return ticks, total * (weight / 3.0)
    return wrapper
