"""Small host-side helpers with the semantics of qspectra/utils.py."""
import copy
import functools
import inspect

import numpy as np


class ZeroArray(object):
    """Additive identity placeholder: ``z += x`` yields ``x`` and ``z -= x``
    yields ``-x`` (qspectra/utils.py:8-36), so accumulators need no shape."""

    def __iadd__(self, other):
        return other

    def __isub__(self, other):
        return -other


def ndarray_list(arrays, length):
    """Stacks ``length`` equally shaped arrays from an iterable into one
    array (qspectra/utils.py:39-49)."""
    out = None
    for k, a in enumerate(arrays):
        a = np.asarray(a)
        if out is None:
            out = np.empty((length,) + a.shape, a.dtype)
        out[k] = a
    return out


class imemoize(object):
    """Per-instance method cache keyed on (method, args, kwargs); arguments
    must be hashable (qspectra/utils.py:52-83)."""

    def __init__(self, func):
        self.func = func
        functools.update_wrapper(self, func)

    def __get__(self, obj, objtype=None):
        if obj is None:
            return self.func
        return functools.partial(self, obj)

    def __call__(self, obj, *args, **kwargs):
        cache = obj.__dict__.setdefault('_imemoize__cache', {})
        key = (self.func, args, frozenset(kwargs.items()))
        if key not in cache:
            cache[key] = self.func(obj, *args, **kwargs)
        return cache[key]


def memoized_property(func):
    return property(imemoize(func))


def copy_with_new_cache(obj):
    """Shallow copy without the imemoize cache (qspectra/utils.py:98-108)."""
    new = copy.copy(obj)
    new.__dict__.pop('_imemoize__cache', None)
    return new


def check_random_state(seed):
    """None -> numpy's global RandomState; a RandomState is returned as is;
    anything else seeds a new RandomState (qspectra/utils.py:86-95)."""
    if seed is None:
        return np.random.mtrand._rand
    if isinstance(seed, np.random.RandomState):
        return seed
    return np.random.RandomState(seed)


def inspect_repr(obj):
    """repr built from the ``__init__`` argument names of ``obj``'s class,
    each read back as an attribute (qspectra/utils.py:115-133)."""
    names = inspect.getfullargspec(type(obj).__init__).args[1:]

    def indent(text, prefix, width=None):
        pad = '\n' + ' ' * (len(prefix) if width is None else width)
        return prefix + pad.join(text.splitlines())

    parts = []
    for name in names:
        value = getattr(obj, name)
        text = '%s(...)' % type(obj).__name__ if value is obj else repr(value)
        parts.append(indent(text, name + '='))
    return indent(',\n'.join(parts), '%s(\n    ' % type(obj).__name__, 4) + ')'
