"""Time-invariant propagation helpers: the one-step RK4 propagator.

For y' = L y with constant L, one classical RK4 step of size h is exactly
y <- P y with P = I + hL + (hL)^2/2 + (hL)^3/6 + (hL)^4/24; it is built on
the GPU by Horner's rule with three complex GEMMs (qsx_zgemm):
    T = I + (h/4) L,  T = I + (h/3) L T,  T = I + (h/2) L T,  P = I + h L T.
Stepping with P replaces four Liouvillian applications per step by one and
reproduces the RK4 trajectory up to rounding.
"""
import torch

from . import _native


def rk4_propagator(L, h):
    """P4(h L) for L [B, M, M] complex128 (FP64)."""
    B, M, _ = L.shape
    eye = torch.eye(M, dtype=L.dtype, device=L.device)
    T = (L * (h / 4.0) + eye).contiguous()
    for c in (h / 3.0, h / 2.0):
        T = _native.zgemm(L, T, alpha=c, gamma=1.0)
    return _native.zgemm(L, T, alpha=h, gamma=1.0)


def rk4_increment(L, h):
    """P4(h L) - I = h L T (the MIXED-mode operator), complex128."""
    B, M, _ = L.shape
    eye = torch.eye(M, dtype=L.dtype, device=L.device)
    T = (L * (h / 4.0) + eye).contiguous()
    for c in (h / 3.0, h / 2.0):
        T = _native.zgemm(L, T, alpha=c, gamma=1.0)
    return _native.zgemm(L, T, alpha=h)


def _cache(eom):
    return eom.__dict__.setdefault('_operators', {})


def cached_propagator(eom, h):
    """P4(h L) of an equation of motion, computed once (FP64)."""
    c = _cache(eom)
    key = ('P', float(h))
    if key not in c:
        c[key] = rk4_propagator(eom.L, h)
    return c[key]


def propagator_power(eom, h, n):
    """P4(h L)^n by repeated squaring (FP64), cached; equal to n classical
    RK4 steps of size h up to rounding."""
    c = _cache(eom)
    key = ('Pn', float(h), int(n))
    if key in c:
        return c[key]
    if n < 1:
        raise ValueError('power must be positive')
    result, k = None, 0
    while (1 << k) <= n:
        sq_key = ('P2k', float(h), k)
        if sq_key not in c:
            c[sq_key] = (cached_propagator(eom, h) if k == 0
                         else _native.zgemm(c[('P2k', float(h), k - 1)],
                                            c[('P2k', float(h), k - 1)]))
        if n & (1 << k):
            result = c[sq_key] if result is None else _native.zgemm(c[sq_key], result)
        k += 1
    c[key] = result
    return result


def transposed(eom, key, A):
    """Cached contiguous transpose (for right-multiplying row states)."""
    c = _cache(eom)
    tkey = ('T',) + tuple(key)
    if tkey not in c:
        c[tkey] = A.transpose(-1, -2).contiguous()
    return c[tkey]


def split_diagonal(A):
    """(off-diagonal part as complex64, diagonal as complex128 [B, M])."""
    diag = torch.diagonal(A, dim1=-2, dim2=-1).contiguous()
    off = A.clone()
    off.diagonal(dim1=-2, dim2=-1).zero_()
    return off.to(torch.complex64).contiguous(), diag
