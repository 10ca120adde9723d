"""ctypes binding of libqsx.so (include/qsx.h) and device-tensor plumbing.

torch is imported first so that the HIP runtime it loads is the one libqsx
binds to (both use the soname libamdhip64.so.7); device buffers are torch
tensors and their ``data_ptr()`` values cross the C ABI.

There is no CPU fallback: every compute entry point raises
``NativeUnavailable`` when the library or a HIP device is missing.
"""
import ctypes
import os
import threading

import numpy as np
import torch

from . import timing

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libqsx.so')

QSX_FP64 = 0
QSX_MIXED = 1


class NativeUnavailable(RuntimeError):
    """libqsx.so or the HIP device it needs is not available."""


class NativeError(RuntimeError):
    """A libqsx entry point returned an error status."""


_lib = None
_lock = threading.Lock()

c_i32, c_i64, c_dbl, c_vp, c_sz = (ctypes.c_int32, ctypes.c_int64, ctypes.c_double,
                                   ctypes.c_void_p, ctypes.c_size_t)

# name -> argtypes (restype int32 unless listed in RESTYPES)
RESTYPES = {'qsx_redfield_workspace_bytes': c_i64, 'qsx_heom_workspace_bytes': c_i64,
            'qsx_adjoint_workspace_bytes': c_i64, 'qsx_zofe_workspace_bytes': c_i64,
            'qsx_redfield_dense_workspace_bytes': c_i64, 'qsx_fourier_workspace_bytes': c_i64,
            'qsx_propagator_workspace_bytes': c_i64,
            'qsx_split_propagator_workspace_bytes': c_i64,
            'qsx_redfield_vjp_workspace_bytes': c_i64}
SIGNATURES = {
    'qsx_version': [],
    'qsx_last_error': [ctypes.c_char_p, c_sz],
    'qsx_apply_dense_c128': [c_vp, c_i64, c_i32, c_vp, c_i32, c_i32, c_vp, c_vp],
    'qsx_rk4_dense': [c_i32, c_vp, c_i64, c_vp, c_i64, c_i32, c_vp, c_i32, c_i32,
                      c_vp, c_vp, c_i32, c_vp, c_i64, c_i32, c_vp, c_vp, c_vp, c_vp],
    'qsx_block_eigh': [c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp],
    'qsx_redfield_workspace_bytes': [c_i32, c_i32, c_i32, c_i32],
    'qsx_redfield_block': [c_vp, c_i32, c_i32, c_vp, c_dbl, c_dbl, c_dbl, c_i32, c_dbl,
                           c_i32, c_i32, c_vp, c_vp, c_vp],
    'qsx_unitary_block': [c_vp, c_i32, c_i32, c_vp, c_dbl, c_i32, c_i32, c_i32, c_vp, c_vp],
    'qsx_redfield_block_tabulated': [c_vp, c_i32, c_i32, c_vp, c_vp, c_vp, c_i32, c_dbl, c_i32,
                                     c_i32, c_vp, c_vp, c_vp],
    'qsx_redfield_vjp_workspace_bytes': [c_i32, c_i32, c_i32, c_i32],
    'qsx_response_buckets': [c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_vp,
                             c_vp],
    'qsx_response_bucket_weights': [c_vp, c_i32, c_vp, c_vp, c_vp, c_i32, c_vp, c_vp],
    'qsx_absorption_operators': [c_vp, c_i32, c_i32, c_vp, c_vp, c_vp, c_i32, c_dbl, c_dbl,
                                 c_i32, c_vp, c_vp, c_vp],
    'qsx_absorption_operators_max_sites': [],
    'qsx_heom_apply_hilbert': [c_i32, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                               c_i32, c_i32, c_vp, c_vp, c_vp, c_i32, c_vp, c_vp, c_vp, c_vp,
                               c_vp],
    'qsx_heom_propagate_hilbert': [c_i32, c_i32, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp,
                                   c_vp, c_vp, c_i32, c_i32, c_vp, c_vp, c_vp, c_i32, c_vp, c_vp,
                                   c_vp, c_vp, c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp],
    'qsx_heom_propagate_driven': [c_i32, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_i32, c_i32, c_vp,
                                  c_vp, c_vp, c_i32, c_vp, c_vp, c_i32, c_vp, c_i64, c_vp, c_i64,
                                  c_i64, c_vp, c_vp, c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp,
                                  c_vp],
    'qsx_response_buckets_max_states': [],
    'qsx_redfield_block_vjp': [c_vp, c_i32, c_i32, c_vp, c_dbl, c_dbl, c_dbl, c_i32, c_dbl, c_i32,
                               c_i32, c_vp, c_vp, c_vp, c_vp, c_vp],
    'qsx_propagate': [c_i32, c_i32, c_vp, c_i64, c_vp, c_i64, c_i32, c_vp, c_i32, c_i32,
                      c_vp, c_vp, c_i32, c_vp, c_vp, c_vp, c_vp],
    'qsx_propagate_chunks': [c_i32, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_i32, c_vp, c_i32,
                             c_i32, c_vp, c_vp, c_i32, c_vp, c_vp, c_vp],
    'qsx_adjoint_workspace_bytes': [c_i32, c_i32, c_i32, c_i64, c_i32],
    'qsx_propagate_adjoint': [c_vp, c_i64, c_i32, c_vp, c_i32, c_i32, c_vp, c_i32, c_vp, c_i32,
                              c_vp, c_i64, c_vp, c_vp, c_vp, c_vp],
    'qsx_zgemm': [c_i32, c_i32, c_i32, c_i32, c_i32, c_dbl, c_dbl, c_vp, c_i64, c_i64, c_vp, c_i64,
                  c_i64, c_dbl, c_dbl, c_vp, c_i64, c_i64, c_dbl, c_dbl, c_i32, c_vp],
    'qsx_heom_workspace_bytes': [c_i32, c_i32, c_i32, c_i32],
    'qsx_heom_apply': [c_i32, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_vp,
                       c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp],
    'qsx_heom_propagate': [c_i32, c_i32, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_i32, c_i32,
                           c_vp, c_vp, c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_vp,
                           c_vp, c_vp, c_vp, c_vp],
    'qsx_propagate_driven': [c_i32, c_vp, c_i64, c_vp, c_i64, c_i32, c_vp, c_vp, c_i64, c_i32,
                             c_vp, c_i64, c_i64, c_i32, c_vp, c_i32, c_i32, c_vp, c_vp, c_i32,
                             c_vp, c_vp, c_vp],
    'qsx_redfield_dense_workspace_bytes': [c_i32, c_i32, c_i32, c_i32],
    'qsx_redfield_block_dense': [c_vp, c_vp, c_i32, c_i32, c_vp, c_vp, c_i32, c_i32, c_dbl, c_dbl,
                                 c_dbl, c_i32, c_dbl, c_vp, c_vp, c_vp],
    'qsx_unitary_block_dense': [c_vp, c_vp, c_i32, c_i32, c_i32, c_dbl, c_i32, c_vp, c_vp],
    'qsx_dense_eigh': [c_vp, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp],
    'qsx_block_eigh_hermitian': [c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp],
    'qsx_dense_eigh_hermitian': [c_vp, c_i32, c_i32, c_vp, c_vp, c_vp],
    'qsx_unitary_block_hermitian': [c_vp, c_i32, c_i32, c_vp, c_dbl, c_i32, c_i32, c_i32, c_vp,
                                    c_vp],
    'qsx_unitary_block_dense_hermitian': [c_vp, c_vp, c_i32, c_i32, c_i32, c_dbl, c_i32, c_vp,
                                          c_vp],
    'qsx_fourier_workspace_bytes': [c_i32],
    'qsx_fourier_transform': [c_vp, c_i64, c_i32, c_i32, c_i32, c_i32, ctypes.c_double, c_i32,
                              c_vp, c_vp, c_vp],
    'qsx_propagator_workspace_bytes': [c_i32, c_i32],
    'qsx_rk4_propagator': [c_vp, c_i64, c_i32, c_i32, c_dbl, c_i32, c_vp, c_vp, c_i64, c_vp],
    'qsx_rk4_propagator_noalloc': [c_vp, c_i64, c_i32, c_i32, c_dbl, c_i32, c_vp, c_vp],
    'qsx_matrix_powers': [c_vp, c_i64, c_i32, c_i32, c_vp, c_i32, c_vp, c_vp, c_i64, c_vp],
    'qsx_split_increment': [c_vp, c_i32, c_i32, c_vp, c_vp, c_vp],
    'qsx_split_propagator_workspace_bytes': [c_i32, c_i32],
    'qsx_rk4_propagator_split': [c_vp, c_i64, c_i32, c_i32, c_dbl, c_i32, c_vp, c_vp, c_vp,
                                 c_i64, c_vp],
    'qsx_small_propagator_fits': [c_i32, c_i32, c_i32],
    'qsx_small_propagator': [c_vp, c_i64, c_i32, c_i32, c_dbl, c_i32, c_i64, c_vp, c_vp, c_vp,
                             c_vp, c_vp],
    'qsx_propagate_powers': [c_vp, c_i64, c_i32, c_i32, c_dbl, c_vp, c_i32, c_vp, c_i32, c_vp,
                             c_vp],
    'qsx_zofe_workspace_bytes': [c_i32, c_i32, c_i32, c_i32, c_i32],
    'qsx_zofe_propagate': [c_i32, c_i32, c_i32, c_i32, c_i32, c_vp, c_i64, c_vp, c_i64, c_vp,
                           c_vp, c_i64, c_i32, c_dbl, c_i32, c_vp, c_i64, c_vp, c_i64, c_i64,
                           c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp],
    'qsx_zofe_apply': [c_i32, c_i32, c_i32, c_i32, c_i32, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp,
                       c_i64, c_i32, c_dbl, c_vp, c_vp, c_vp, c_vp],
}


def load_library(path=LIB_PATH):
    """Loads libqsx.so (no device work); raises NativeUnavailable if absent."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise NativeUnavailable(
                    '%s is not built; run __graft_entry__.build() or '
                    'python -m qspectra_amd.csrc.build' % path)
            lib = ctypes.CDLL(path)
            for name, argtypes in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.argtypes = argtypes
                fn.restype = RESTYPES.get(name, c_i32)
            _lib = lib
    return _lib


def require_native():
    """Returns the library; raises NativeUnavailable without a HIP device."""
    lib = load_library()
    if not torch.cuda.is_available():
        raise NativeUnavailable('no HIP device visible: qspectra_amd has no CPU path')
    return lib


def device():
    require_native()
    return torch.device('cuda', torch.cuda.current_device())


def last_error():
    buf = ctypes.create_string_buffer(512)
    load_library().qsx_last_error(buf, 512)
    return buf.value.decode(errors='replace')


def check(status, what):
    if status != 0:
        raise NativeError('%s failed (status %d): %s' % (what, status, last_error()))


def stream_ptr(stream=None):
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def to_device(x, dtype=torch.complex128, dev=None):
    """numpy/array-like/tensor -> contiguous device tensor of ``dtype``."""
    dev = device() if dev is None else dev
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=dtype).contiguous()
    x = np.ascontiguousarray(x)
    if not x.flags.writeable:              # broadcast views: torch wants writable memory
        x = x.copy()
    return torch.as_tensor(x).to(device=dev, dtype=dtype).contiguous()


# --------------------------------------------------------------- wrappers
def apply_dense(L, y, stream=None):
    """out[b, r] = L[b] @ y[b, r]  (complex128; L [B|1, M, M], y [B, R, M])."""
    lib = require_native()
    B, R, M = y.shape
    assert L.dtype == torch.complex128 and y.dtype == torch.complex128
    assert L.shape[-2:] == (M, M) and L.is_contiguous() and y.is_contiguous()
    strideL = 0 if L.shape[0] == 1 else M * M
    assert L.shape[0] in (1, B)
    out = torch.empty_like(y)
    check(lib.qsx_apply_dense_c128(ptr(L), strideL, M, ptr(y), R, B, ptr(out),
                                   stream_ptr(stream)), 'qsx_apply_dense_c128')
    return out


def rk4_dense(L, y0, nsub, h, proj=None, precision=QSX_FP64, diag=None,
              want_final=False, stream=None, steps=None):
    """Fused RK4 with saves; see include/qsx.h qsx_rk4_dense.

    L     [B|1, M, M] complex128 (FP64) or complex64 off-diagonal (MIXED)
    diag  [B|1, M] complex128 diagonal (MIXED only)
    y0    [B, R, M] complex128
    nsub  int32 [nsave], h float64 [nsave] (device tensors)
    proj  [B|1, K, M] (dtype of L) or None
    Returns (out [B, R, nsave, K], y_final or None, status tensor).
    """
    lib = require_native()
    B, R, M = y0.shape
    nsave = int(nsub.shape[0])
    ldt = torch.complex128 if precision == QSX_FP64 else torch.complex64
    assert L.dtype == ldt and L.is_contiguous() and L.shape[-2:] == (M, M)
    assert y0.dtype == torch.complex128 and y0.is_contiguous()
    assert nsub.dtype == torch.int32 and h.dtype == torch.float64
    strideL = 0 if L.shape[0] == 1 else M * M
    if proj is not None:
        assert proj.dtype == ldt and proj.is_contiguous() and proj.shape[-1] == M
        K = proj.shape[-2]
        strideP = 0 if proj.shape[0] == 1 else K * M
    else:
        K, strideP = M, 0
    strideD = 0
    if precision == QSX_MIXED:
        assert diag is not None and diag.dtype == torch.complex128 and diag.is_contiguous()
        strideD = 0 if diag.shape[0] == 1 else M
    out = torch.empty((B, R, nsave, K), dtype=torch.complex128, device=y0.device)
    yfin = torch.empty_like(y0) if want_final else None
    status = torch.zeros(1, dtype=torch.int32, device=y0.device)
    label = 'rk4_dense %s M=%d R=%d B=%d K=%d steps=%s nsave=%d' % (
        'fp64' if precision == QSX_FP64 else 'mixed', M, R, B, K, steps, nsave)
    with timing.launch(label, stream):
        st = lib.qsx_rk4_dense(precision, ptr(L), strideL, ptr(diag), strideD, M, ptr(y0), R, B,
                               ptr(nsub), ptr(h), nsave, ptr(proj), strideP, K, ptr(out),
                               ptr(yfin), ptr(status), stream_ptr(stream))
    check(st, 'qsx_rk4_dense')
    return out, yfin, status


# ------------------------------------------------------ Liouvillian builders
HILBERT_BLOCK = {'g': 0, 'e': 1, 'f': 2}
QSX_SECULAR, QSX_REAL_CORR, QSX_EIGEN_BASIS, QSX_HEISENBERG = 1, 2, 4, 8


def block_size(X, n):
    return (1, n, n * (n - 1) // 2)[X]


def _workspace(n, X, Y, B, dev):
    nbytes = load_library().qsx_redfield_workspace_bytes(n, X, Y, B)
    if nbytes < 0:
        raise NativeError('invalid workspace request n=%d X=%d Y=%d B=%d' % (n, X, Y, B))
    return torch.empty((nbytes + 7) // 8, dtype=torch.float64, device=dev)


def block_eigh(H1, X, rw=None, stream=None):
    """E [B, nb], U [B, nb, nb] of Hilbert block X of H1 - X*rw (float64,
    or complex128 for complex-Hermitian H1: then U is complex128)."""
    lib = require_native()
    B, n, _ = H1.shape
    assert H1.is_contiguous()
    nb = block_size(X, n)
    if H1.dtype == torch.complex128:
        E = torch.empty((B, nb), dtype=torch.float64, device=H1.device)
        U = torch.empty((B, nb, nb), dtype=torch.complex128, device=H1.device)
        check(lib.qsx_block_eigh_hermitian(ptr(H1), n, B, X, ptr(rw), ptr(E), ptr(U),
                                           stream_ptr(stream)), 'qsx_block_eigh_hermitian')
        return E, U
    assert H1.dtype == torch.float64
    E = torch.empty((B, nb), dtype=torch.float64, device=H1.device)
    U = torch.empty((B, nb, nb), dtype=torch.float64, device=H1.device)
    ws = _workspace(n, X, X, B, H1.device)
    check(lib.qsx_block_eigh(ptr(H1), n, B, X, ptr(rw), ptr(E), ptr(U), ptr(ws),
                             stream_ptr(stream)), 'qsx_block_eigh')
    return E, U


def redfield_block(H1, rw, T, lam, gam, X, Y, flags, unit_convert, stream=None, L=None,
                   ws=None):
    """Redfield Liouvillian block XY, complex128 [B, M, M] (written into L
    and built with the workspace ws when given: redfield_workspace)."""
    lib = require_native()
    B, n, _ = H1.shape
    assert H1.dtype == torch.float64 and H1.is_contiguous()
    M = block_size(X, n) * block_size(Y, n)
    if L is None:
        L = torch.empty((B, M, M), dtype=torch.complex128, device=H1.device)
    assert L.shape == (B, M, M) and L.dtype == torch.complex128 and L.is_contiguous()
    if ws is None:
        ws = _workspace(n, X, Y, B, H1.device)

    def go():
        return lib.qsx_redfield_block(ptr(H1), n, B, ptr(rw), float(T), float(lam), float(gam),
                                      int(flags), float(unit_convert), X, Y, ptr(L), ptr(ws),
                                      stream_ptr(stream))
    with timing.launch('redfield block %s%s n=%d B=%d' % ('gef'[X], 'gef'[Y], n, B), stream, go):
        st = go()
    check(st, 'qsx_redfield_block')
    return L


def redfield_block_tabulated(H1, rw, ctabX, ctabY, X, Y, flags, unit_convert, stream=None):
    """Redfield block XY with the bath correlation function tabulated over
    each member's eigen-energy differences (ctab complex128 [B, nb, nb])."""
    lib = require_native()
    B, n, _ = H1.shape
    assert H1.dtype == torch.float64 and H1.is_contiguous()
    nX, nY = block_size(X, n), block_size(Y, n)
    for C, nb in ((ctabX, nX), (ctabY, nY)):
        assert C.dtype == torch.complex128 and C.is_contiguous() and C.shape == (B, nb, nb)
    L = torch.empty((B, nX * nY, nX * nY), dtype=torch.complex128, device=H1.device)
    ws = _workspace(n, X, Y, B, H1.device)

    def go():
        return lib.qsx_redfield_block_tabulated(ptr(H1), n, B, ptr(rw), ptr(ctabX), ptr(ctabY),
                                                int(flags), float(unit_convert), X, Y, ptr(L),
                                                ptr(ws), stream_ptr(stream))
    with timing.launch('redfield block tabulated %s%s n=%d B=%d' % ('gef'[X], 'gef'[Y], n, B),
                       stream, go):
        st = go()
    check(st, 'qsx_redfield_block_tabulated')
    return L


def redfield_block_vjp(H1, rw, T, lam, gam, X, Y, flags, unit_convert, gL, stream=None):
    """Cotangents (gH1 [B, n, n] float64, glam [B] float64) of the Redfield
    block XY that redfield_block returns for the same arguments, given its
    cotangent gL (complex128 [B, M, M])."""
    lib = require_native()
    B, n, _ = H1.shape
    assert H1.dtype == torch.float64 and H1.is_contiguous()
    M = block_size(X, n) * block_size(Y, n)
    assert gL.dtype == torch.complex128 and gL.is_contiguous() and gL.shape == (B, M, M)
    gH1 = torch.empty_like(H1)
    glam = torch.empty((B,), dtype=torch.float64, device=H1.device)
    nbytes = lib.qsx_redfield_vjp_workspace_bytes(n, X, Y, B)
    ws = torch.empty((max(int(nbytes), 8) + 7) // 8, dtype=torch.float64, device=H1.device)

    def go():
        return lib.qsx_redfield_block_vjp(ptr(H1), n, B, ptr(rw), float(T), float(lam),
                                          float(gam), int(flags), float(unit_convert), X, Y,
                                          ptr(gL), ptr(gH1), ptr(glam), ptr(ws),
                                          stream_ptr(stream))
    with timing.launch('redfield block vjp %s%s n=%d B=%d' % ('gef'[X], 'gef'[Y], n, B),
                       stream, go):
        st = go()
    check(st, 'qsx_redfield_block_vjp')
    return gH1, glam


def response_buckets_into(waves, offsets, P, y0, W, out, status, J=1, stream=None):
    """out [B, nt] <- s_k = v_b . P_b^k y0_b for problems of unequal size in
    one launch, J steps (1 or 4) per pass (qsx_response_buckets; the tables
    are device int32 [nw, 4] and int64 [B, 3] (P offset, y0 offset, output
    row); P holds P_b^J, W the projection rows (v for J = 1; see
    response_bucket_weights_into), all flat complex128)."""
    lib = require_native()
    assert waves.dtype == torch.int32 and waves.is_contiguous() and waves.shape[1] == 4
    assert offsets.dtype == torch.int64 and offsets.is_contiguous() and offsets.shape[1] == 3
    assert offsets.shape[0] == out.shape[0]
    for X in (P, y0, W, out):
        assert X.dtype == torch.complex128 and X.is_contiguous()
    assert W.numel() == J * y0.numel()
    assert status.dtype == torch.int32 and status.numel() == out.shape[0]
    nt = out.shape[1]

    def go():
        return lib.qsx_response_buckets(ptr(waves), waves.shape[0], ptr(offsets), ptr(P), ptr(y0),
                                        ptr(W), int(J), nt, ptr(out), ptr(status),
                                        stream_ptr(stream))
    with timing.launch('response buckets B=%d waves=%d nt=%d' % (out.shape[0], waves.shape[0], nt),
                       stream, go):
        st = go()
    check(st, 'qsx_response_buckets')


def absorption_operators_into(table, max_n, H1, rw, bath, flags, unit_convert, h, J, P1, PJ,
                              stream=None):
    """P1 <- P4(h L_b), PJ <- P1^J for the Heisenberg 'eg' Redfield blocks L_b
    of a ragged batch of one-exciton Hamiltonians in one launch
    (qsx_absorption_operators; table int64 [B, 3]: n <= max_n, H1 offset,
    P offset; H1, rw, bath [B, 3] float64 and P1, PJ complex128, all flat
    device tensors)."""
    lib = require_native()
    assert table.dtype == torch.int64 and table.is_contiguous() and table.shape[1] == 3
    for X in (H1, rw, bath):
        assert X.dtype == torch.float64 and X.is_contiguous()
    for X in (P1, PJ):
        assert X.dtype == torch.complex128 and X.is_contiguous()
    B = table.shape[0]
    assert rw.numel() == B and bath.numel() == 3 * B

    def go():
        return lib.qsx_absorption_operators(ptr(table), B, int(max_n), ptr(H1), ptr(rw),
                                            ptr(bath), int(flags), float(unit_convert), float(h),
                                            int(J), ptr(P1), ptr(PJ), stream_ptr(stream))
    with timing.launch('absorption operators B=%d n<=%d J=%d' % (B, max_n, J), stream, go):
        st = go()
    check(st, 'qsx_absorption_operators')


def response_bucket_weights_into(waves, offsets, P, v, J, W, stream=None):
    """W <- the projection rows (P_b^j)^T v_b, j < J, of every problem
    (qsx_response_bucket_weights; P the one-step propagators at the bucket
    offsets, v flat complex128)."""
    lib = require_native()
    for X in (P, v, W):
        assert X.dtype == torch.complex128 and X.is_contiguous()
    assert W.numel() == J * v.numel()
    check(lib.qsx_response_bucket_weights(ptr(waves), waves.shape[0], ptr(offsets), ptr(P),
                                          ptr(v), int(J), ptr(W), stream_ptr(stream)),
          'qsx_response_bucket_weights')


def redfield_workspace(n, X, Y, B, dev=None):
    """Device workspace of qsx_redfield_block for block XY of n sites."""
    return _workspace(n, X, Y, B, device() if dev is None else dev)


def unitary_block(H1, rw, unit_convert, X, Y, heisenberg=False, stream=None):
    """Unitary Liouvillian block XY, complex128 [B, M, M] (H1 float64, or
    complex128 for complex-Hermitian Hamiltonians)."""
    lib = require_native()
    B, n, _ = H1.shape
    assert H1.is_contiguous()
    M = block_size(X, n) * block_size(Y, n)
    L = torch.empty((B, M, M), dtype=torch.complex128, device=H1.device)
    if H1.dtype == torch.complex128:
        check(lib.qsx_unitary_block_hermitian(ptr(H1), n, B, ptr(rw), float(unit_convert), X, Y,
                                              int(bool(heisenberg)), ptr(L), stream_ptr(stream)),
              'qsx_unitary_block_hermitian')
        return L
    assert H1.dtype == torch.float64
    check(lib.qsx_unitary_block(ptr(H1), n, B, ptr(rw), float(unit_convert), X, Y,
                                int(bool(heisenberg)), ptr(L), stream_ptr(stream)),
          'qsx_unitary_block')
    return L


# -------------------------------------------------- register stepper / GEMM
QSX_MODE_RK4 = 0
QSX_MODE_PROPAGATOR = 1
MAX_REGISTER_BLOCK = 160


def propagate_states(mode, A, y0, nsub, h, precision=QSX_FP64, diag=None, want_final=False,
                     stream=None, steps=None, status=None):
    """qsx_propagate: full states at every save point, out [B, R, nsave, M].
    status: optional int32 device slot (set to 1 on a non-finite state; the
    caller clears it), allocated and cleared here when None.  steps: RK4
    steps of one right-hand side (recorded in the launch label)."""
    lib = require_native()
    B, R, M = y0.shape
    nsave = int(nsub.shape[0])
    adt = torch.complex128 if precision == QSX_FP64 else torch.complex64
    assert A.dtype == adt and A.is_contiguous() and A.shape[-2:] == (M, M)
    assert A.shape[0] in (1, B) and y0.dtype == torch.complex128 and y0.is_contiguous()
    assert nsub.dtype == torch.int32 and h.dtype == torch.float64
    strideA = 0 if A.shape[0] == 1 else M * M
    strideD = 0
    if precision == QSX_MIXED:
        assert diag is not None and diag.dtype == torch.complex128 and diag.is_contiguous()
        strideD = 0 if diag.shape[0] == 1 else M
    out = torch.empty((B, R, nsave, M), dtype=torch.complex128, device=y0.device)
    yfin = torch.empty_like(y0) if want_final else None
    if status is None:
        status = torch.zeros(1, dtype=torch.int32, device=y0.device)
    label = 'propagate %s %s M=%d R=%d B=%d steps=%s nsave=%d' % (
        'rk4' if mode == QSX_MODE_RK4 else 'prop', 'fp64' if precision == QSX_FP64 else 'mixed',
        M, R, B, steps, nsave)
    def go():
        return lib.qsx_propagate(mode, precision, ptr(A), strideA, ptr(diag), strideD, M,
                                 ptr(y0), R, B, ptr(nsub), ptr(h), nsave, ptr(out), ptr(yfin),
                                 ptr(status), stream_ptr(stream))
    with timing.launch(label, stream, go):
        st = go()
    check(st, 'qsx_propagate')
    return out, yfin, status


def propagate_chunks(Pc, A, y0, nc, nsub, h, precision=QSX_FP64, diag=None, stream=None,
                     steps=None, status=None):
    """qsx_propagate_chunks: trajectory of y0 [B, 1, M] in nc chunks of
    nsave saves each, out [B, nc, nsave, M]; chunk j starts from
    Pc^j y0 (Pc complex128 [B|1, M, M]) and is stepped with A over the
    intervals nsub/h.  status as in propagate_states; steps: products of
    one chunk (recorded in the launch label)."""
    lib = require_native()
    B, R, M = y0.shape
    assert R == 1 and nc >= 1
    nsave = int(nsub.shape[0])
    adt = torch.complex128 if precision == QSX_FP64 else torch.complex64
    assert A.dtype == adt and A.is_contiguous() and A.shape[-2:] == (M, M)
    assert Pc.dtype == torch.complex128 and Pc.is_contiguous() and Pc.shape[-2:] == (M, M)
    assert A.shape[0] in (1, B) and Pc.shape[0] in (1, B)
    assert y0.dtype == torch.complex128 and y0.is_contiguous()
    assert nsub.dtype == torch.int32 and h.dtype == torch.float64 and h.shape[0] == nsave
    strideA = 0 if A.shape[0] == 1 else M * M
    strideP = 0 if Pc.shape[0] == 1 else M * M
    strideD = 0
    if precision == QSX_MIXED:
        assert diag is not None and diag.dtype == torch.complex128 and diag.is_contiguous()
        assert diag.shape[-1] == M and diag.shape[0] in (1, B)
        strideD = 0 if diag.shape[0] == 1 else M
    out = torch.empty((B, nc, nsave, M), dtype=torch.complex128, device=y0.device)
    if status is None:
        status = torch.zeros(1, dtype=torch.int32, device=y0.device)
    label = 'chunks prop %s M=%d nc=%d B=%d steps=%s nsave=%d' % (
        'fp64' if precision == QSX_FP64 else 'mixed', M, nc, B, steps, nsave)

    def go():
        return lib.qsx_propagate_chunks(precision, ptr(Pc), strideP, ptr(A), strideA, ptr(diag),
                                        strideD, M, ptr(y0), nc, B, ptr(nsub), ptr(h), nsave,
                                        ptr(out), ptr(status), stream_ptr(stream))
    with timing.launch(label, stream, go):
        st = go()
    check(st, 'qsx_propagate_chunks')
    return out, status


MAX_DRIVEN_BLOCK = 1536
MAX_FIELD_TERMS = 4


def ell_pack(C):
    """ELL form of field operators C [B|1, P, M, M] (numpy) over the union of
    their sparsity patterns (all terms and batch entries): (col int32
    [B|1, M, K], coef complex128 [B|1, M, K, P]); K is the largest number of
    nonzero columns of a row (at least 1), padding slots point at column 0
    with zero coefficients."""
    C = np.asarray(C)
    Bc, P, M, _ = C.shape
    mask = np.any(C != 0, axis=(0, 1))                          # [M, M]
    nnz = mask.sum(axis=1)
    K = max(int(nnz.max()), 1)
    col = np.zeros((M, K), dtype=np.int32)
    for i in range(M):
        cols = np.flatnonzero(mask[i])
        col[i, :cols.size] = cols
    valid = np.arange(K)[None, :] < nnz[:, None]                # [M, K]
    coef = np.take_along_axis(C, col[None, None].repeat(Bc, 0).repeat(P, 1), axis=3)
    coef = np.where(valid[None, None], coef, 0).transpose(0, 2, 3, 1)   # [B|1, M, K, P]
    return np.ascontiguousarray(col[None]), np.ascontiguousarray(coef)


def propagate_driven(A, col, coef, F, y0, nsub, h, precision=QSX_FP64, diag=None, stream=None,
                     status=None):
    """qsx_propagate_driven: RK4 of y' = A y + sum_i f_i(t) C_i y with the
    field operators in ELL form (col int32 [1, M, K], coef [B|1, M, K, P],
    see ell_pack) and the field values F [B|1, nF, 3, P] (complex128, at
    the three stage times of every step); returns (out [B, R, nsave, M],
    status)."""
    lib = require_native()
    B, R, M = y0.shape
    nsave = int(nsub.shape[0])
    adt = torch.complex128 if precision == QSX_FP64 else torch.complex64
    K, P = int(coef.shape[2]), int(coef.shape[3])
    if A.dtype != adt or coef.dtype != adt or F.dtype != torch.complex128 or \
            col.dtype != torch.int32:
        raise TypeError('propagate_driven: operand dtypes do not match the precision')
    if (A.shape[-2:] != (M, M) or A.shape[0] not in (1, B) or tuple(col.shape) != (1, M, K)
            or coef.dim() != 4 or coef.shape[0] not in (1, B) or coef.shape[1] != M
            or F.dim() != 4 or F.shape[0] not in (1, B) or F.shape[2] != 3 or F.shape[3] != P):
        raise ValueError('propagate_driven: shapes A %s, col %s, coef %s, F %s do not fit '
                         'states %s' % (tuple(A.shape), tuple(col.shape), tuple(coef.shape),
                                        tuple(F.shape), tuple(y0.shape)))
    if not 1 <= P <= MAX_FIELD_TERMS:
        raise ValueError('propagate_driven: %d field terms (1..%d)' % (P, MAX_FIELD_TERMS))
    if M > MAX_DRIVEN_BLOCK:
        raise ValueError('propagate_driven: block of %d states exceeds %d' % (M, MAX_DRIVEN_BLOCK))
    assert nsub.dtype == torch.int32 and h.dtype == torch.float64
    A, col, coef, F, y0 = (x.contiguous() for x in (A, col, coef, F, y0))
    if col.numel() and (int(col.min()) < 0 or int(col.max()) >= M):
        raise ValueError('propagate_driven: ELL column index outside 0..%d' % (M - 1))
    nF = int(F.shape[1])
    strideA = 0 if A.shape[0] == 1 else M * M
    strideE = 0 if coef.shape[0] == 1 else M * K
    strideF = 0 if F.shape[0] == 1 else nF * 3 * P
    strideD = 0
    if precision == QSX_MIXED:
        assert diag is not None and diag.dtype == torch.complex128
        diag = diag.contiguous()
        strideD = 0 if diag.shape[0] == 1 else M
    out = torch.empty((B, R, nsave, M), dtype=torch.complex128, device=y0.device)
    if status is None:
        status = torch.zeros(1, dtype=torch.int32, device=y0.device)
    label = 'driven %s M=%d K=%d P=%d R=%d B=%d steps=%d nsave=%d' % (
        'fp64' if precision == QSX_FP64 else 'mixed', M, K, P, R, B, nF, nsave)
    def go():
        return lib.qsx_propagate_driven(precision, ptr(A), strideA, ptr(diag), strideD, K,
                                        ptr(col), ptr(coef), strideE, P, ptr(F), strideF, nF, M,
                                        ptr(y0), R, B, ptr(nsub), ptr(h), nsave, ptr(out),
                                        ptr(status), stream_ptr(stream))
    with timing.launch(label, stream, go):
        st = go()
    check(st, 'qsx_propagate_driven')
    return out, status


def propagate_adjoint(A, y0, nsub, total_steps, gout, ckpt=64, stream=None):
    """qsx_propagate_adjoint: (gA [B, R, M, M], gy0 [B, R, M], status) for
    the propagator schedule y <- A y (A = P, complex128) with step counts
    nsub (device int32, summing to total_steps) and saved-state cotangents
    gout [B, R, nsave, M]."""
    lib = require_native()
    B, R, M = y0.shape
    nsave = int(nsub.shape[0])
    if A.dtype != torch.complex128 or y0.dtype != torch.complex128:
        raise TypeError('propagate_adjoint needs complex128 operands')
    if A.shape[-2:] != (M, M) or A.shape[0] not in (1, B):
        raise ValueError('operator shape %s does not match states %s' % (tuple(A.shape),
                                                                          tuple(y0.shape)))
    if tuple(gout.shape) != (B, R, nsave, M):
        raise ValueError('cotangent shape %s, expected %s' % (tuple(gout.shape),
                                                              (B, R, nsave, M)))
    strideA = 0 if A.shape[0] == 1 else M * M
    A, y0, gout = A.contiguous(), y0.contiguous(), gout.contiguous()
    nbytes = lib.qsx_adjoint_workspace_bytes(M, R, B, int(total_steps), int(ckpt))
    if nbytes < 0:
        raise NativeError('invalid adjoint workspace request')
    ws = torch.empty((nbytes + 15) // 16, dtype=torch.complex128, device=y0.device)
    gA = torch.empty((B, R, M, M), dtype=torch.complex128, device=y0.device)
    gy0 = torch.empty((B, R, M), dtype=torch.complex128, device=y0.device)
    status = torch.zeros(1, dtype=torch.int32, device=y0.device)
    with timing.launch('adjoint M=%d R=%d B=%d steps=%d ckpt=%d' % (M, R, B, total_steps, ckpt),
                       stream):
        st = lib.qsx_propagate_adjoint(ptr(A), strideA, M, ptr(y0), R, B, ptr(nsub), nsave,
                                       ptr(gout), int(ckpt),
                                       ptr(ws), ws.numel() * 16, ptr(gA), ptr(gy0), ptr(status),
                                       stream_ptr(stream))
    check(st, 'qsx_propagate_adjoint')
    return gA, gy0, status


def _mat_strides(X):
    """(leading dimension, batch stride) of a [b, rows, cols] operand whose
    last axis is contiguous (views such as column slices are accepted)."""
    assert X.dim() == 3 and (X.shape[2] == 1 or X.stride(2) == 1), \
        'operand rows must be contiguous'
    ld = X.stride(1) if X.shape[1] > 1 else X.shape[2]
    assert ld >= X.shape[2]
    return ld, (X.stride(0) if X.shape[0] > 1 else 0)


def zgemm(A, B, C=None, alpha=1.0, beta=0.0, gamma=0.0, precision=QSX_FP64, stream=None,
          transa=False, transb=False):
    """C = alpha op(A) op(B) + beta C + gamma I over a leading batch axis
    (1 = shared); op(X) = X^T when trans is set.

    A [b|1, m, k] ([b|1, k, m] when transa), B [b|1, k, n] ([b|1, n, k] when
    transb), complex128 (or complex64 for MIXED); C [b, m, n] complex128
    (allocated when None).  Operands may be strided views whose last axis is
    contiguous (for example column slices of a wider matrix)."""
    lib = require_native()
    assert A.dim() == 3 and B.dim() == 3
    bA = A.shape[0]
    m, k = (A.shape[2], A.shape[1]) if transa else (A.shape[1], A.shape[2])
    bB = B.shape[0]
    k2, n = (B.shape[2], B.shape[1]) if transb else (B.shape[1], B.shape[2])
    assert k == k2, (tuple(A.shape), tuple(B.shape), transa, transb)
    batch = max(bA, bB) if C is None else C.shape[0]
    if C is None:
        C = torch.empty((batch, m, n), dtype=torch.complex128, device=A.device)
        beta = 0.0
    assert C.dtype == torch.complex128 and C.shape == (batch, m, n)
    assert bA in (1, batch) and bB in (1, batch)
    dt = torch.complex128 if precision == QSX_FP64 else torch.complex64
    assert A.dtype == dt and B.dtype == dt
    if m == 0 or n == 0:
        return C
    if k == 0:              # empty contraction: only the beta and identity terms
        if complex(beta) == 0:
            C.zero_()
        else:
            C.mul_(complex(beta))
        if complex(gamma) != 0:
            C.diagonal(dim1=1, dim2=2).add_(complex(gamma))
        return C
    lda, sA = _mat_strides(A)
    ldb, sB = _mat_strides(B)
    ldc, sC = _mat_strides(C)
    if batch > 1:
        assert sC != 0
    al, be, ga = complex(alpha), complex(beta), complex(gamma)
    trans = (1 if transa else 0) | (2 if transb else 0)
    def go():
        return lib.qsx_zgemm(precision, trans, m, n, k, al.real, al.imag, ptr(A), lda, sA,
                             ptr(B), ldb, sB, be.real, be.imag, ptr(C), ldc, sC, ga.real,
                             ga.imag, batch, stream_ptr(stream))
    with timing.launch('zgemm %s m=%d n=%d k=%d batch=%d%s' % (
            'fp64' if precision == QSX_FP64 else 'mixed', m, n, k, batch,
            ' trans=%d' % trans if trans else ''), stream, go):
        st = go()
    check(st, 'qsx_zgemm')
    return C


# ------------------------------------------------- propagation operators
def propagator_workspace(M, B, dev=None):
    """Device workspace of qsx_rk4_propagator / qsx_matrix_powers."""
    n = load_library().qsx_propagator_workspace_bytes(M, B)
    if n < 0:
        raise NativeError('qsx_propagator_workspace_bytes(%d, %d) failed' % (M, B))
    dev = device() if dev is None else dev
    return torch.empty(n // 16, dtype=torch.complex128, device=dev)


def rk4_propagator_into(L, h, q, P, workspace, stream=None):
    """P [B, M, M] <- P4(h L)^q on the device (L [B|1, M, M] complex128)."""
    lib = require_native()
    B, M, _ = P.shape
    assert L.dtype == torch.complex128 and L.is_contiguous() and L.shape[-2:] == (M, M)
    assert L.shape[0] in (1, B) and P.dtype == torch.complex128 and P.is_contiguous()
    assert workspace.dtype == torch.complex128
    strideL = 0 if L.shape[0] == 1 else M * M

    def go():
        return lib.qsx_rk4_propagator(ptr(L), strideL, M, B, float(h), int(q), ptr(P),
                                      ptr(workspace), workspace.numel() * 16, stream_ptr(stream))
    with timing.launch('rk4 propagator M=%d B=%d q=%d' % (M, B, q), stream, go):
        st = go()
    check(st, 'qsx_rk4_propagator')
    return P


def matrix_powers_into(A, exps, out, workspace, stream=None):
    """out [B, ns, M, M] <- A^exps[s] (A [B|1, M, M] complex128)."""
    lib = require_native()
    B, ns, M, _ = out.shape
    assert A.dtype == torch.complex128 and A.is_contiguous() and A.shape[-2:] == (M, M)
    assert A.shape[0] in (1, B) and out.is_contiguous() and len(exps) == ns
    e = (ctypes.c_int64 * ns)(*[int(x) for x in exps])
    strideA = 0 if A.shape[0] == 1 else M * M

    def go():
        return lib.qsx_matrix_powers(ptr(A), strideA, M, B, e, ns, ptr(out), ptr(workspace),
                                     workspace.numel() * 16, stream_ptr(stream))
    with timing.launch('matrix powers M=%d B=%d e=%s' % (M, B, ','.join(map(str, exps))),
                       stream, go):
        st = go()
    check(st, 'qsx_matrix_powers')
    return out


def split_increment_into(P, off, diag, stream=None):
    """off [B, M, M] complex64 <- P with zero diagonal; diag [B, M] <- diag(P) - 1."""
    lib = require_native()
    B, M, _ = P.shape
    assert P.dtype == torch.complex128 and P.is_contiguous()
    assert off.dtype == torch.complex64 and off.shape == P.shape and off.is_contiguous()
    assert diag.dtype == torch.complex128 and diag.shape == (B, M) and diag.is_contiguous()

    def go():
        return lib.qsx_split_increment(ptr(P), M, B, ptr(off), ptr(diag), stream_ptr(stream))
    with timing.launch('split increment M=%d B=%d' % (M, B), stream, go):
        st = go()
    check(st, 'qsx_split_increment')
    return off, diag


def split_propagator_workspace(M, B, dev=None):
    """Device workspace of qsx_rk4_propagator_split."""
    n = load_library().qsx_split_propagator_workspace_bytes(M, B)
    if n < 0:
        raise NativeError('qsx_split_propagator_workspace_bytes(%d, %d) failed' % (M, B))
    dev = device() if dev is None else dev
    return torch.empty(-(-n // 16), dtype=torch.complex128, device=dev)


def rk4_propagator_split_into(L, h, q, off, diag, workspace, stream=None):
    """MIXED operands of P4(h L)^q built on the FP32 matrix cores: off
    [B, M, M] complex64, diag [B, M] complex128 (q a power of two)."""
    lib = require_native()
    B, M, _ = off.shape
    assert L.dtype == torch.complex128 and L.is_contiguous() and L.shape[-2:] == (M, M)
    assert L.shape[0] in (1, B) and off.dtype == torch.complex64 and off.is_contiguous()
    assert diag.dtype == torch.complex128 and diag.shape == (B, M) and diag.is_contiguous()
    strideL = 0 if L.shape[0] == 1 else M * M

    def go():
        return lib.qsx_rk4_propagator_split(ptr(L), strideL, M, B, float(h), int(q), ptr(off),
                                            ptr(diag), ptr(workspace), workspace.numel() * 16,
                                            stream_ptr(stream))
    with timing.launch('split propagator M=%d B=%d q=%d' % (M, B, q), stream, go):
        st = go()
    check(st, 'qsx_rk4_propagator_split')
    return off, diag


def small_fits(M, R=0, ns=0):
    """True when a block of M states fits the one-work-group kernels
    (qsx_small_propagator with R = 0, qsx_propagate_powers otherwise)."""
    return bool(load_library().qsx_small_propagator_fits(M, R, ns))


def small_propagator_into(L, h, q, c=0, P=None, Pc=None, off=None, diag=None, stream=None):
    """One launch for blocks of M <= 64: P <- P4(h L)^q, Pc <- P^c, and the
    MIXED split (off complex64, diag complex128) of P; outputs may be None."""
    lib = require_native()
    Bl, M, _ = L.shape
    assert L.dtype == torch.complex128 and L.is_contiguous()
    for X, dt, shp in ((P, torch.complex128, (Bl, M, M)), (Pc, torch.complex128, (Bl, M, M)),
                       (off, torch.complex64, (Bl, M, M)), (diag, torch.complex128, (Bl, M))):
        if X is not None:
            assert X.dtype == dt and X.is_contiguous() and X.numel() == int(np.prod(shp))

    def go():
        return lib.qsx_small_propagator(ptr(L), M * M, M, Bl, float(h), int(q), int(c), ptr(P),
                                        ptr(Pc), ptr(off), ptr(diag), stream_ptr(stream))
    with timing.launch('small propagator M=%d B=%d q=%d c=%d' % (M, Bl, q, c), stream, go):
        st = go()
    check(st, 'qsx_small_propagator')


def propagate_powers_into(L, h, exps, y0, out, stream=None):
    """out [B, ns, R, M] <- P4(h L)^exps[s] y0 [B, R, M] (blocks of M <= 64)."""
    lib = require_native()
    B, R, M = y0.shape
    ns = len(exps)
    assert L.dtype == torch.complex128 and L.is_contiguous() and L.shape[-2:] == (M, M)
    assert L.shape[0] in (1, B) and y0.dtype == torch.complex128 and y0.is_contiguous()
    assert out.dtype == torch.complex128 and out.is_contiguous() and out.numel() == B * ns * R * M
    e = (ctypes.c_int64 * ns)(*[int(x) for x in exps])
    strideL = 0 if L.shape[0] == 1 else M * M

    def go():
        return lib.qsx_propagate_powers(ptr(L), strideL, M, B, float(h), e, ns, ptr(y0), R,
                                        ptr(out), stream_ptr(stream))
    with timing.launch('propagate powers M=%d B=%d R=%d e=%s' % (
            M, B, R, ','.join(map(str, exps))), stream, go):
        st = go()
    check(st, 'qsx_propagate_powers')
    return out


# ------------------------------------------------------------------- HEOM
def _heom_ptrs(eom):
    return (eom.n_ado, eom.M, eom.n_comm, ptr(eom.Hop), ptr(eom.shift), ptr(eom.dvec),
            eom.nbins, eom.K1, ptr(eom.bsrc), ptr(eom.bcoef), ptr(eom.ck), eom.nrel,
            ptr(eom.rel_site), ptr(eom.rel_occ))


def _heom_entry(lib, eom, name):
    """(entry point, leading arguments) of qsx_heom_<name> for the form of
    eom: the Hilbert-commutator entries take the state tables in place of
    the commutator block size."""
    hb = getattr(eom, 'hilbert', None)
    if hb is None:
        return getattr(lib, 'qsx_heom_' + name), _heom_ptrs(eom)
    rest = _heom_ptrs(eom)[3:]
    return (getattr(lib, 'qsx_heom_%s_hilbert' % name),
            (eom.n_ado, eom.M, hb['N'], ptr(hb['est']), ptr(hb['pos']), ptr(hb['blk'])) + rest)


def heom_apply(eom, y, stream=None):
    """F(y) for hierarchy states y [B, n_ado * M] (complex128)."""
    lib = require_native()
    B = eom.Hop.shape[0]
    assert y.shape == (B, eom.n_ado * eom.M) and y.dtype == torch.complex128
    y = y.contiguous()
    out = torch.empty_like(y)
    fn, args = _heom_entry(lib, eom, 'apply')
    check(fn(B, *args, ptr(y), ptr(out), stream_ptr(stream)),
          'qsx_heom_apply')
    return out


def heom_propagate(eom, y0, nsub, h, precision=QSX_FP64, want_final=False, stream=None):
    """RK4 over the hierarchy; nsub, h are host arrays.  Returns ADO 0 at
    every save point, out [B, nsave, M], the final states and the status."""
    lib = require_native()
    B = eom.Hop.shape[0]
    S = eom.n_ado * eom.M
    assert y0.shape == (B, S) and y0.dtype == torch.complex128
    y0 = y0.contiguous()
    nsub = np.ascontiguousarray(nsub, dtype=np.int32)
    h = np.ascontiguousarray(h, dtype=np.float64)
    nsave = int(nsub.size)
    nbytes = lib.qsx_heom_workspace_bytes(precision, B, eom.n_ado, eom.M)
    if nbytes < 0:
        raise NativeError('invalid HEOM workspace request')
    ws = torch.empty((nbytes + 15) // 16, dtype=torch.complex128, device=y0.device)
    out = torch.empty((B, nsave, eom.M), dtype=torch.complex128, device=y0.device)
    yfin = torch.empty_like(y0) if want_final else None
    status = torch.zeros(1, dtype=torch.int32, device=y0.device)
    label = 'heom %s B=%d n_ado=%d M=%d steps=%d' % (
        'fp64' if precision == QSX_FP64 else 'mixed', B, eom.n_ado, eom.M, int(nsub.sum()))
    with timing.launch(label, stream):
        fn, args = _heom_entry(lib, eom, 'propagate')
        st = fn(
            precision, B, *args, ptr(y0), nsub.ctypes.data_as(c_vp),
            h.ctypes.data_as(c_vp), nsave, ptr(out), ptr(yfin), ptr(ws), ptr(status),
            stream_ptr(stream))
    check(st, 'qsx_heom_propagate')
    return out, yfin, status


def heom_propagate_driven(eom, C, F, y0, nsub, h, precision=QSX_FP64, want_final=False,
                          stream=None):
    """qsx_heom_propagate under classical fields (dense hierarchy blocks):
    C [B|1, P, M, M] the dipole commutators, F [B|1, nF, 3, P] the field
    samples (complex128 device tensors); nsub, h host arrays.  Returns
    (ADO 0 at every save point [B, nsave, M], final states or None, status)."""
    lib = require_native()
    if eom.n_comm != 0:
        raise ValueError('fields need the dense hierarchy form (several Liouville blocks, '
                         'M <= 64)')
    if getattr(eom, 'hilbert', None) is not None:
        raise NotImplementedError('fields on hierarchies of more than 64 elements per ADO')
    B = eom.Hop.shape[0]
    M, S = eom.M, eom.n_ado * eom.M
    assert y0.shape == (B, S) and y0.dtype == torch.complex128
    assert C.dtype == torch.complex128 and C.is_contiguous() and C.shape[0] in (1, B)
    assert F.dtype == torch.complex128 and F.is_contiguous() and F.shape[0] in (1, B)
    P, nF = C.shape[1], F.shape[1]
    assert C.shape[2:] == (M, M) and F.shape[2:] == (3, P)
    y0 = y0.contiguous()
    nsub = np.ascontiguousarray(nsub, dtype=np.int32)
    h = np.ascontiguousarray(h, dtype=np.float64)
    nsave = int(nsub.size)
    nbytes = lib.qsx_heom_workspace_bytes(precision, B, eom.n_ado, M)
    if nbytes < 0:
        raise NativeError('invalid HEOM workspace request')
    ws = torch.empty((nbytes + 15) // 16, dtype=torch.complex128, device=y0.device)
    Hwork = torch.empty((B, M, M), dtype=torch.complex128, device=y0.device)
    out = torch.empty((B, nsave, M), dtype=torch.complex128, device=y0.device)
    yfin = torch.empty_like(y0) if want_final else None
    status = torch.zeros(1, dtype=torch.int32, device=y0.device)
    n_ado, M_, _, *rest = _heom_ptrs(eom)
    label = 'heom driven %s B=%d n_ado=%d M=%d P=%d steps=%d' % (
        'fp64' if precision == QSX_FP64 else 'mixed', B, eom.n_ado, M, P, int(nsub.sum()))
    with timing.launch(label, stream):
        st = lib.qsx_heom_propagate_driven(
            precision, B, n_ado, M_, *rest, P, ptr(C), 0 if C.shape[0] == 1 else P * M * M,
            ptr(F), 0 if F.shape[0] == 1 else nF * 3 * P, nF, ptr(y0),
            nsub.ctypes.data_as(c_vp), h.ctypes.data_as(c_vp), nsave, ptr(out), ptr(yfin),
            ptr(Hwork), ptr(ws), ptr(status), stream_ptr(stream))
    check(st, 'qsx_heom_propagate_driven')
    return out, yfin, status


def _zofe_ws(lib, B, R, eom, dev):
    nbytes = lib.qsx_zofe_workspace_bytes(B, R, eom.N, eom.S, eom.P)
    if nbytes < 0:
        raise NativeError('invalid ZOFE workspace request')
    return torch.empty((nbytes + 15) // 16, dtype=torch.complex128, device=dev)


def _zofe_common(eom, B):
    strideH = 0 if eom.H.shape[0] == 1 else eom.N * eom.N
    strideL = 0 if eom.L.shape[0] == 1 else eom.S * eom.N * eom.N
    strideG = 0 if eom.Gamma.shape[0] == 1 else eom.P * eom.S
    return (ptr(eom.H), strideH, ptr(eom.L), strideL, ptr(eom.Gamma), ptr(eom.w), strideG,
            eom.flags, float(eom.unit_convert))


def zofe_apply(eom, y, stream=None):
    """qsx_zofe_apply: eom(t, y) for y [B, R, size] complex128 (reference
    layout); returns the derivatives in the same layout."""
    lib = require_native()
    B, R, size = y.shape
    y = y.contiguous()
    out = torch.empty_like(y)
    ws = _zofe_ws(lib, B, R, eom, y.device)
    st = lib.qsx_zofe_apply(B, R, eom.N, eom.S, eom.P, *_zofe_common(eom, B), ptr(y), ptr(out),
                            ptr(ws), stream_ptr(stream))
    check(st, 'qsx_zofe_apply')
    return out


def zofe_propagate(eom, y0, nsub, h, proj=None, V=None, F=None, stream=None):
    """qsx_zofe_propagate: y0 [B, R, size] (reference layout), device
    schedule (nsub int32, h float64); proj [B|1, K, N*N] rows on vec(rho)
    or None (full states); field terms V [B|1, P, N, N] with samples
    F [B|1, nF, 3, P].  Returns (out [B, R, nsave, K or size], status)."""
    lib = require_native()
    B, R, size = y0.shape
    if size != eom.size:
        raise ValueError('ZOFE states have %d elements, the model %d' % (eom.size, size))
    nsave = int(nsub.shape[0])
    y0 = y0.contiguous()
    K, strideP = 0, 0
    if proj is not None:
        proj = proj.contiguous()
        K = int(proj.shape[1])
        strideP = 0 if proj.shape[0] == 1 else K * eom.N * eom.N
    nfield, strideV, strideF, nF = 0, 0, 0, 0
    if V is not None:
        V, F = V.contiguous(), F.contiguous()
        nfield = int(V.shape[1])
        strideV = 0 if V.shape[0] == 1 else nfield * eom.N * eom.N
        nF = int(F.shape[1])
        strideF = 0 if F.shape[0] == 1 else nF * 3 * nfield
    out = torch.empty((B, R, nsave, K if K else size), dtype=torch.complex128, device=y0.device)
    ws = _zofe_ws(lib, B, R, eom, y0.device)
    status = torch.zeros(1, dtype=torch.int32, device=y0.device)
    label = 'zofe N=%d S=%d P=%d R=%d B=%d fields=%d nsave=%d' % (eom.N, eom.S, eom.P, R, B,
                                                                   nfield, nsave)
    def go():
        return lib.qsx_zofe_propagate(B, R, eom.N, eom.S, eom.P, *_zofe_common(eom, B), nfield,
                                      ptr(V), strideV, ptr(F), strideF, nF, ptr(y0), ptr(nsub),
                                      ptr(h), nsave, K, ptr(proj), strideP, ptr(out), ptr(ws),
                                      ptr(status), stream_ptr(stream))
    with timing.launch(label, stream, go):
        st = go()
    check(st, 'qsx_zofe_propagate')
    return out, status


def dense_eigh(H, stream=None):
    """qsx_dense_eigh: E [B, nb], U [B, nb, nb] of real symmetric H [B, nb, nb]
    (float64 device tensor); complex Hermitian H (complex128) through
    qsx_dense_eigh_hermitian, U then complex128."""
    lib = require_native()
    if H.is_complex():
        H = H.to(torch.complex128).contiguous()
        B, nb, _ = H.shape
        E = torch.empty((B, nb), dtype=torch.float64, device=H.device)
        U = torch.empty((B, nb, nb), dtype=torch.complex128, device=H.device)
        check(lib.qsx_dense_eigh_hermitian(ptr(H), nb, B, ptr(E), ptr(U), stream_ptr(stream)),
              'qsx_dense_eigh_hermitian')
        return E, U
    H = H.to(torch.float64).contiguous()
    B, nb, _ = H.shape
    nbytes = lib.qsx_redfield_dense_workspace_bytes(1, nb, nb, B)
    ws = torch.empty((max(nbytes, 8) + 7) // 8, dtype=torch.float64, device=H.device)
    E = torch.empty((B, nb), dtype=torch.float64, device=H.device)
    U = torch.empty((B, nb, nb), dtype=torch.float64, device=H.device)
    check(lib.qsx_dense_eigh(ptr(H), nb, B, ptr(E), ptr(U), ptr(ws), stream_ptr(stream)),
          'qsx_dense_eigh')
    return E, U


def redfield_block_dense(HX, HY, occX, occY, T, lam, gam, flags, unit_convert, stream=None):
    """qsx_redfield_block_dense: L [B, nX*nY, nX*nY] complex128 from dense
    rotating-frame blocks HX [B, nX, nX], HY [B, nY, nY] (float64; HY may be
    HX) and coupling diagonals occX [n, nX], occY [n, nY]."""
    lib = require_native()
    B, nX, _ = HX.shape
    nY = HY.shape[-1]
    n = occX.shape[0]
    M = nX * nY
    nbytes = lib.qsx_redfield_dense_workspace_bytes(n, nX, nY, B)
    ws = torch.empty((nbytes + 7) // 8, dtype=torch.float64, device=HX.device)
    L = torch.empty((B, M, M), dtype=torch.complex128, device=HX.device)
    check(lib.qsx_redfield_block_dense(ptr(HX), ptr(HY), nX, nY, ptr(occX), ptr(occY), n, B,
                                       float(T), float(lam), float(gam), flags,
                                       float(unit_convert), ptr(L), ptr(ws),
                                       stream_ptr(stream)), 'qsx_redfield_block_dense')
    return L


def unitary_block_dense(HX, HY, unit_convert, heisenberg=False, stream=None):
    """qsx_unitary_block_dense: -i unit [H, .] on the XY block (complex128
    blocks through qsx_unitary_block_dense_hermitian)."""
    lib = require_native()
    B, nX, _ = HX.shape
    nY = HY.shape[-1]
    M = nX * nY
    L = torch.empty((B, M, M), dtype=torch.complex128, device=HX.device)
    if HX.is_complex() or HY.is_complex():
        HX = HX.to(torch.complex128).contiguous()
        HY = HY.to(torch.complex128).contiguous()
        check(lib.qsx_unitary_block_dense_hermitian(ptr(HX), ptr(HY), nX, nY, B,
                                                    float(unit_convert), int(bool(heisenberg)),
                                                    ptr(L), stream_ptr(stream)),
              'qsx_unitary_block_dense_hermitian')
        return L
    check(lib.qsx_unitary_block_dense(ptr(HX), ptr(HY), nX, nY, B, float(unit_convert),
                                      int(bool(heisenberg)), ptr(L), stream_ptr(stream)),
          'qsx_unitary_block_dense')
    return L


def _fourier(x, n, N, lo, sign, dt, adjoint, stream):
    lib = require_native()
    x = x.to(torch.complex128).contiguous()
    lead = tuple(x.shape[:-1])
    if x.shape[-1] != (N if adjoint else n):
        raise ValueError('fourier transform: last axis has %d entries, expected %d'
                         % (x.shape[-1], N if adjoint else n))
    out = torch.empty(lead + ((n if adjoint else N),), dtype=torch.complex128, device=x.device)
    batch = int(np.prod(lead)) if lead else 1
    if batch == 0:
        return out
    nbytes = lib.qsx_fourier_workspace_bytes(int(N))
    ws = torch.empty((nbytes + 15) // 16, dtype=torch.complex128, device=x.device)
    check(lib.qsx_fourier_transform(ptr(x), batch, int(n), int(N), int(lo), int(sign),
                                    float(dt), int(adjoint), ptr(ws), ptr(out),
                                    stream_ptr(stream)), 'qsx_fourier_transform')
    return out


def fourier_transform(x, N, lo, sign, dt, stream=None):
    """qsx_fourier_transform along the last axis of a device tensor:
    [..., n] samples -> [..., N] on the zero-padded symmetric grid."""
    return _fourier(x, x.shape[-1], N, lo, sign, dt, False, stream)


def fourier_transform_adjoint(g, n, lo, sign, dt, stream=None):
    """The conjugate transpose of fourier_transform: [..., N] -> [..., n]."""
    return _fourier(g, n, g.shape[-1], lo, sign, dt, True, stream)
