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


def split_diagonal(A):
    """(off-diagonal part as complex64, diagonal as complex128 [B, M])."""
    diag = torch.diagonal(A, dim1=-2, dim2=-1).contiguous()
    off = A.clone()
    off.diagonal(dim1=-2, dim2=-1).zero_()
    return off.to(torch.complex64).contiguous(), diag
