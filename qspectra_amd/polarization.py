"""Polarization vectors and isotropic-average weights (host side), with the
semantics of qspectra/polarization.py."""
from itertools import product
from numbers import Number

import numpy as np

from .utils import check_random_state

_AXES = {'x': np.array([1, 0, 0]), 'y': np.array([0, 1, 0]), 'z': np.array([0, 0, 1])}

FOURTH_ORDER_INVARIANTS = [((0, 1), (2, 3)), ((0, 2), (1, 3)), ((0, 3), (1, 2))]

MAGIC_ANGLE = np.arccos(1 / np.sqrt(3))
"""The magic angle (about 54.7 degrees) in radians."""


def polarization_vector(p):
    """'x'/'y'/'z', an angle in the x-y plane, or a 3-vector -> 3-vector;
    ValueError for anything else."""
    if isinstance(p, str):
        if p in _AXES:
            return _AXES[p]
    elif isinstance(p, Number):
        return np.array([np.cos(p), np.sin(p), 0])
    else:
        try:
            v = np.asanyarray(p, float).reshape(-1)
        except (TypeError, ValueError):
            v = None
        if v is not None and v.size == 3:
            return v
    raise ValueError("cannot read %r as a polarization ('x', 'y', 'z', an angle in the x-y "
                     "plane or a 3-vector)" % (p,))


def check_polarizations(p, length):
    """Polarization vectors [length, 3] of a sequence of `length` items."""
    vecs = np.array([polarization_vector(x) for x in p])
    if len(vecs) != length:
        raise ValueError('expected %d polarizations, got %d' % (length, len(vecs)))
    return vecs


def invariant_weights_4th_order(polarizations):
    """Weights of the invariants <xxyy>, <xyxy>, <xyyx> giving the isotropic
    average of a fourth-rank response (Hochstrasser, Chem. Phys. 266, 273)."""
    p = check_polarizations(polarizations, 4)
    cos = p @ p.T
    prods = np.array([cos[a] * cos[b] for a, b in FOURTH_ORDER_INVARIANTS])
    return (5 * np.eye(3) - np.ones((3, 3))) @ prods / 30


def invariant_polarizations(invariant):
    """The 9 lab polarizations summed to evaluate one invariant."""
    if invariant not in FOURTH_ORDER_INVARIANTS:
        raise ValueError('unknown fourth-order invariant %r (expected one of %r)'
                         % (invariant, FOURTH_ORDER_INVARIANTS))
    return [''.join(p) for p in product('xyz', repeat=4)
            if all(p[a] == p[b] for a, b in invariant)]


def random_rotation_matrix(random_state=None):
    """Uniform random rotation (Arvo, Graphics Gems III, 117-120)."""
    x1, x2, x3 = check_random_state(random_state).rand(3)
    theta, phi = 2 * np.pi * x1, 2 * np.pi * x2
    Rz = np.array([[np.cos(theta), np.sin(theta), 0],
                   [-np.sin(theta), np.cos(theta), 0], [0, 0, 1]])
    v = np.array([np.cos(phi) * np.sqrt(x3), np.sin(phi) * np.sqrt(x3), np.sqrt(1 - x3)])
    return -(np.identity(3) - 2 * np.outer(v, v)) @ Rz
