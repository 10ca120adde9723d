"""Gradients of third-order responses and 2-D spectra (gradients.
RedfieldThirdOrder): the differentiable forward pass against the product's
pathway-by-pathway third_order_response and two_dimensional_spectra of the
same disorder members (1e-10), and the gradient of a 2-D spectral loss with
respect to every site energy, coupling and lambda against central finite
differences of the FP64 forward pass (1e-5)."""
import numpy as np
import pytest
import torch

import qspectra_amd as qa
from qspectra_amd import gradients as G
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

BATH = (qa.CM_K * 77, 35.0, 106.0)
RW, DT, TMAX, T2 = 12500.0, 1.0, 48.0, [0.0, 20.0, 50.0]
FWHM = 60.0


def _trimer():
    H = np.array([[12410., 85., -20.], [85., 12505., 55.], [-20., 55., 12590.]])
    d = np.array([[1.0, 0.2, 0.0], [0.3, -0.8, 0.1], [-0.5, 0.1, 0.9]])
    return H, d


def _host(x):
    return x.cpu().numpy() if hasattr(x, 'cpu') else np.asarray(x)


def _theta(H):
    iu = np.triu_indices(len(H), 1)
    mk = lambda a: torch.tensor(np.array(a, dtype=float), device='cuda', requires_grad=True)
    return mk(np.diag(H)), mk(H[iu]), mk(BATH[1])


def _product_members(H, d, size):
    ham = qa.ElectronicHamiltonian(H, bath=qa.DebyeBath(*BATH), dipoles=d, disorder=FWHM)
    models = []
    for h in ham.sample_ensemble(size):
        m = qa.RedfieldModel(h, rw_freq=RW, hilbert_subspace='gef', unit_convert=qa.CM_FS,
                             secular=False)
        m.time_step_override = DT
        models.append(m)
    return models


@pytest.mark.parametrize('geometry', ['-++', '+-+', '++-'])
def test_third_order_forward_matches_product(gpu, geometry):
    H, d = _trimer()
    tds = G.RedfieldThirdOrder(d, BATH[2], BATH[0], RW, TMAX, T2, DT, geometry=geometry,
                               disorder=FWHM, ensemble_size=3)
    E, J, lam = _theta(H)
    with torch.no_grad():
        R = tds.response(E, J, lam).cpu().numpy()
        (f1, t2, f3), X = tds(E, J, lam)
    ref, refX = 0, 0
    for m in _product_members(H, d, 3):
        (t1, t2p, t3), r = qa.third_order_response(m, TMAX, population_times=T2,
                                                   geometry=geometry, dt=DT, plan='pathways',
                                                   precision='fp64')
        ref = ref + _host(r) / 3
        (g1, _, g3), x = qa.two_dimensional_spectra(m, TMAX, population_times=T2,
                                                    geometry=geometry, dt=DT, plan='pathways',
                                                    precision='fp64')
        refX = refX + _host(x) / 3
    assert R.shape == ref.shape == (len(tds.t1), 3, len(tds.t3))
    assert np.allclose(t1, tds.t1) and np.allclose(t2p, T2)
    assert rel_err(R, ref) < 1e-10
    assert np.allclose(f1, g1) and np.allclose(f3, g3)
    assert rel_err(X.cpu().numpy(), refX) < 1e-10


def test_two_dimensional_gradient_matches_finite_differences(gpu):
    """All 7 parameters (3 E, 3 J, lam) of a 2-member ensemble; the loss is
    the squared distance of the real 2-D spectrum from that of perturbed
    parameters; central differences with h = 1e-3 cm^-1."""
    H, d = _trimer()
    tds = G.RedfieldThirdOrder(d, BATH[2], BATH[0], RW, 32.0, [0.0, 30.0], DT,
                               disorder=FWHM, ensemble_size=2, ckpt=16)
    rng = np.random.RandomState(5)
    Ht = H.copy()
    Ht[np.diag_indices(3)] += rng.uniform(-30, 30, 3)
    Et, Jt, lt = _theta(Ht)
    with torch.no_grad():
        target = tds(Et, Jt, lt + 4.0)[1].real

    def loss(E, J, lam):
        return ((tds(E, J, lam)[1].real - target) ** 2).sum()

    E, J, lam = _theta(H)
    loss(E, J, lam).backward()
    grad = torch.cat([E.grad, J.grad, lam.grad[None]]).cpu().numpy()
    theta0 = [E.detach(), J.detach(), lam.detach()]
    fd, h = [], 1e-3
    with torch.no_grad():
        for which, t in enumerate(theta0):
            for i in range(t.numel()):
                val = []
                for sgn in (1, -1):
                    th = [x.clone() for x in theta0]
                    th[which].reshape(-1)[i] += sgn * h
                    val.append(float(loss(*th)))
                fd.append((val[0] - val[1]) / (2 * h))
    fd = np.array(fd)
    err = np.abs(grad - fd).max() / np.abs(fd).max()
    assert err < 1e-5, (err, grad, fd)


def test_seven_site_gradient_matches_finite_differences(gpu):
    """The 7-site aggregate ('fe' block of 147 elements through the
    large-block adjoint): all 29 gradients of a 2-D spectral loss against
    central differences of the FP64 forward pass (1e-5)."""
    from qspectra_amd import workloads as W
    H, d, _ = W.aggregate7_hamiltonian()
    tds = G.RedfieldThirdOrder(d, BATH[2], BATH[0], 12450.0, 16.0, [0.0, 10.0], DT, ckpt=8)
    rng = np.random.RandomState(7)
    Ht = H + np.diag(rng.uniform(-30, 30, 7))
    Et, Jt, lt = _theta(Ht)
    with torch.no_grad():
        target = tds(Et, Jt, lt + 3.0)[1].real

    def loss(E, J, lam):
        return ((tds(E, J, lam)[1].real - target) ** 2).sum()

    E, J, lam = _theta(H)
    loss(E, J, lam).backward()
    grad = torch.cat([E.grad, J.grad, lam.grad[None]]).cpu().numpy()
    theta0 = [E.detach(), J.detach(), lam.detach()]
    fd, h = [], 1e-3
    with torch.no_grad():
        for which, t in enumerate(theta0):
            for i in range(t.numel()):
                val = []
                for sgn in (1, -1):
                    th = [x.clone() for x in theta0]
                    th[which].reshape(-1)[i] += sgn * h
                    val.append(float(loss(*th)))
                fd.append((val[0] - val[1]) / (2 * h))
    fd = np.array(fd)
    err = np.abs(grad - fd).max() / np.abs(fd).max()
    assert err < 1e-5, (err, grad, fd)
