"""qsx_rk4_propagator_noalloc (the allocation-free entry point, workspace in
the library's device array) against qsx_rk4_propagator with a caller
workspace: identical results, on two calls in a row, the second smaller;
problems beyond the array are rejected before any launch."""
import numpy as np
import pytest
import torch

from qspectra_amd import _native

pytestmark = pytest.mark.gpu


def _noalloc(L, h, q, P):
    lib = _native.require_native()
    B, M, _ = P.shape
    strideL = 0 if L.shape[0] == 1 else M * M
    return lib.qsx_rk4_propagator_noalloc(_native.ptr(L), strideL, M, B, float(h), int(q),
                                          _native.ptr(P), _native.stream_ptr())


def test_noalloc_propagator_matches_workspace_entry(gpu):
    rs = np.random.RandomState(21)
    for M, B, h, q in [(147, 4, 1.0, 2), (49, 2, 0.5, 3)]:      # the second call smaller
        L = ((rs.randn(B, M, M) + 1j * rs.randn(B, M, M)) * (0.05 / np.sqrt(M))
             - (0.08 + 0.3j) * np.eye(M))
        Ld = torch.as_tensor(L).cuda()
        ref = torch.empty((B, M, M), dtype=torch.complex128, device='cuda')
        _native.rk4_propagator_into(Ld, h, q, ref, _native.propagator_workspace(M, B))
        out = torch.full((B, M, M), complex('nan'), dtype=torch.complex128, device='cuda')
        assert _noalloc(Ld, h, q, out) == 0, _native.last_error()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (M, B)
    # 4 * 25 * 147^2 complex128 = 34.6 MB exceeds the 32 MiB array
    big = torch.zeros((25, 147, 147), dtype=torch.complex128, device='cuda')
    Pbig = torch.empty_like(big)
    assert _noalloc(big, 1.0, 1, Pbig) == -1
    assert 'scratch array holds' in _native.last_error()
