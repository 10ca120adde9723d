"""qsx_fourier_transform (the reference's zero-padded symmetric FFT,
qspectra/simulate/utils.py:128-219, on the GPU: a direct sum for rows
shorter than 2048 samples or points, the chirp-z form with FFTs in LDS for
longer ones) against the oracle's numpy FFT pipeline: grids starting at,
before and after t = 0 (the zero padding on either side), both signs,
batched rows along any axis, and the adjoint against the conjugate-transposed
transform matrix, through autograd, and (long rows) against numpy's inverse
FFT run through the transposed pipeline.  Tolerance 1e-13 relative
(max-norm)."""
import numpy as np
import pytest
import torch

import qspectra_amd as qa
from qspectra_amd import _native
from qspectra_amd.simulate.utils import fourier_transform_device
from oracle import propagate as P
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('n,t0,dt', [(2, 0.0, 1.0), (7, 0.0, 0.7), (64, -20.0, 0.5),
                                     (300, 11.0, 0.25), (1989, 0.0, 5.0297804513411934),
                                     (4096, -4096 * 0.5, 0.5), (5000, 7.0, 2.0)])
def test_matches_numpy_fft_pipeline(gpu, n, t0, dt):
    rs = np.random.RandomState(n)
    _check_against_oracle(rs, n, t0, dt)


def _check_against_oracle(rs, n, t0, dt):
    t = t0 + dt * np.arange(n)
    x = rs.randn(3, n) + 1j * rs.randn(3, n)
    for sign in (1, -1):
        for conv in ('angular', 'linear'):
            fo, Xo = P.fourier_transform(t, x, axis=-1, rw_freq=123.0, unit_convert=0.3,
                                         sign=sign, convention=conv)
            f, X = qa.fourier_transform(t, x, axis=-1, rw_freq=123.0, unit_convert=0.3,
                                        sign=sign, convention=conv)
            assert np.array_equal(f, fo)
            assert rel_err(X, Xo) < 1e-13, (n, sign)


def test_axes_and_adjoint(gpu):
    rs = np.random.RandomState(1)
    t1 = 0.5 * np.arange(40)
    t3 = 0.5 * np.arange(33)
    R = rs.randn(40, 5, 33) + 1j * rs.randn(40, 5, 33)
    f1, A = P.fourier_transform(t1, R, axis=0, sign=-1)
    f3, B = P.fourier_transform(t3, A, axis=2)
    _, Ag = qa.fourier_transform(t1, R, axis=0, sign=-1)
    _, Bg = qa.fourier_transform(t3, Ag, axis=2)
    assert rel_err(Ag, A) < 1e-13 and rel_err(Bg, B) < 1e-13
    # adjoint: <F x, g> = <x, F^H g>, and the autograd backward is F^H g
    t = -3.0 + 0.25 * np.arange(21)
    x = torch.as_tensor(rs.randn(21) + 1j * rs.randn(21)).cuda().requires_grad_(True)
    _, X = fourier_transform_device(t, x, -1, 0.0, 1.0, sign=1)
    g = torch.as_tensor(rs.randn(X.shape[-1]) + 1j * rs.randn(X.shape[-1])).cuda()
    (gx,) = torch.autograd.grad((X * g.conj()).sum().real, x)
    F = np.stack([P.fourier_transform(t, e, sign=1)[1] for e in np.eye(21)], axis=1)
    assert rel_err(gx.cpu().numpy(), F.conj().T @ g.cpu().numpy()) < 1e-13
    ga = _native.fourier_transform_adjoint(g[None], 21, 0, 1, 0.25)[0].cpu().numpy()
    assert rel_err(ga, F.conj().T @ g.cpu().numpy()) < 1e-13


def _adjoint_numpy(t, g, sign):
    """The oracle's pipeline transposed: flip (sign +1), ifftshift (the
    transpose of fftshift), N * ifft (the conjugate transpose of fft),
    fftshift (the transpose of ifftshift), the sample positions of the
    zero-padded grid, times dt."""
    t_sym, _ = P.symmetrize(t, np.zeros(t.size))
    lo = int(np.searchsorted(t_sym, t[0]))
    N = g.shape[-1]
    h = g[..., ::-1] if sign == 1 else g
    y = np.fft.fftshift(N * np.fft.ifft(np.fft.ifftshift(h, axes=-1), axis=-1), axes=-1)
    return (t[1] - t[0]) * y[..., lo:lo + t.size], lo


@pytest.mark.parametrize('n,t0,dt,rows', [(4096, 0.0, 1.0, 37), (2048, -3.0, 0.5, 5),
                                          (6000, 11.0, 0.25, 3)])
def test_long_rows_forward_and_adjoint(gpu, n, t0, dt, rows):
    """Rows of 2048 samples and more (the chirp-z path: one piece and two
    output chunks for 4096 samples from t = 0, two pieces and three chunks
    for 6000) against the numpy FFT pipeline, with a scale; the adjoint
    against the transposed numpy pipeline."""
    rs = np.random.RandomState(n + rows)
    t = t0 + dt * np.arange(n)
    x = rs.randn(rows, n) + 1j * rs.randn(rows, n)
    for sign in (1, -1):
        fo, Xo = P.fourier_transform(t, x, axis=-1, sign=sign)
        f, X = fourier_transform_device(t, torch.as_tensor(x).cuda(), -1, sign=sign,
                                        scale=-2.5)
        assert np.array_equal(f, fo)
        assert rel_err(X.cpu().numpy(), -2.5 * Xo) < 1e-13, (n, sign)
        g = rs.randn(rows, Xo.shape[-1]) + 1j * rs.randn(rows, Xo.shape[-1])
        xa_ref, lo = _adjoint_numpy(t, g, sign)
        xa = _native.fourier_transform_adjoint(torch.as_tensor(g).cuda(), n, lo, sign, dt)
        assert rel_err(xa.cpu().numpy(), xa_ref) < 1e-13, (n, sign)
