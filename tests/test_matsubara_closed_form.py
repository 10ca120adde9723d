"""The truncated Matsubara sum of the Debye correlation function,
S(x) = sum_{k=1}^{999} nu_k / ((nu_k^2 - gam^2)(nu_k - i x)), nu_k = 2 pi k T
(bath.py:17-31 truncates the sum at 1000 terms), against its closed form in
digamma differences, which qsx_absorption_operators evaluates for
0.01 <= gam / (2 pi T) <= 32: with p = gam / (2 pi T) and q = i x / (2 pi T)
each term is k / ((k - p)(k + p)(k - q)) / (2 pi T)^2, whose partial fractions
A/(k - p) + B/(k + p) + C/(k - q) sum to digamma differences.  The direct sum
runs in long double; scipy's digamma stands in for the kernel's."""
import numpy as np
import pytest
from scipy.special import psi


def _direct(x, T, g):
    k = np.arange(1, 1000, dtype=np.longdouble)
    nu = 2 * np.pi * np.longdouble(T) * k
    t = nu / ((nu * nu - np.longdouble(g) ** 2) * (nu - 1j * np.longdouble(x)))
    return complex(np.sum(t.astype(np.clongdouble)))


def _closed(x, T, g):
    a = 2 * np.pi * T
    p, q = g / a, 1j * x / a
    A, B, C = 0.5 / (p - q), -0.5 / (p + q), q / (q * q - p * p)
    d = lambda z: psi(1000 - z) - psi(1 - z)
    return (A * d(complex(p)) + B * d(complex(-p)) + C * d(q)) / a ** 2


@pytest.mark.parametrize('p', [0.01, 0.05, 0.315, 0.93, 2.5, 7.3, 28.6, 31.7])
def test_closed_form_equals_truncated_sum(p):
    rng = np.random.RandomState(3)
    for T in (2.78, 53.5, 208.5):
        g = p * 2 * np.pi * T
        for x in np.concatenate([rng.uniform(-3, 3, 3) * T, rng.uniform(-60, 60, 3) * T,
                                 [1e-6 * T, 4000.0]]):
            d = _direct(x, T, g)
            assert abs(_closed(x, T, g) - d) <= 1e-13 * abs(d), (p, T, x)


def test_closed_form_symmetry_and_dropped_term():
    """S(-x) is the conjugate of S(x); a closed form missing the C term (the
    pole at q) or with the sign of B flipped fails the comparison."""
    T, g, x = 53.5, 106.0, 240.0
    assert abs(_closed(-x, T, g) - np.conj(_closed(x, T, g))) < 1e-15 * abs(_closed(x, T, g))
    a = 2 * np.pi * T
    p, q = g / a, 1j * x / a
    d = lambda z: psi(1000 - z) - psi(1 - z)
    no_c = (0.5 / (p - q) * d(complex(p)) - 0.5 / (p + q) * d(complex(-p))) / a ** 2
    flip_b = (0.5 / (p - q) * d(complex(p)) + 0.5 / (p + q) * d(complex(-p))
              + q / (q * q - p * p) * d(q)) / a ** 2
    ref = _direct(x, T, g)
    assert abs(no_c - ref) > 1e-3 * abs(ref) and abs(flip_b - ref) > 1e-3 * abs(ref)
