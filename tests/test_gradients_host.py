"""Host-side checks of the differentiable Redfield restatement used by the
GPU reverse mode's tests: its blocks against the oracle (redfield.py:9-104),
its derivatives against finite differences (torch.autograd.gradcheck), and
the ensemble/Hamiltonian helpers of the gradient path against the model
classes."""
import numpy as np
import pytest
import torch

import qspectra_amd as qa
from qspectra_amd import gradients as G
from tests import torch_redfield as TR
from oracle import system as S
from oracle import liouville as Lv
from tests.conftest import rel_err

BATH = (S.CM_K * 77, 35.0, 106.0)


def _random_h(n, seed):
    rng = np.random.RandomState(seed)
    J = rng.normal(0.0, 60.0, (n, n))
    H = (J + J.T) / 2
    H[np.diag_indices(n)] = rng.uniform(12200, 12700, n)
    return H, rng.normal(size=(n, 3))


@pytest.mark.parametrize('secular', [False, True])
@pytest.mark.parametrize('real_corr', [False, True])
def test_torch_redfield_blocks_match_oracle(secular, real_corr):
    H, d = _random_h(3, 11)
    rw = 12450.0
    om = Lv.LiouvilleModel(S.ElectronicHamiltonian(H, bath=S.DebyeBath(*BATH), dipoles=d),
                           rw_freq=rw, hilbert_subspace='ge', unit_convert=S.CM_FS,
                           secular=secular, discard_imag_corr=real_corr)
    Ht = torch.as_tensor(H)[None]
    for X, Y in [('e', 'g'), ('g', 'e'), ('e', 'e'), ('g', 'g')]:
        L = TR.redfield_block(Ht, BATH[1], BATH[2], BATH[0], rw, X, Y, secular, real_corr,
                             S.CM_FS)[0].numpy()
        ref = om.evolve_matrix(X + Y)
        assert L.shape == ref.shape
        assert rel_err(L, ref) < 1e-12, (X, Y)


def test_debye_correlation_matches_oracle():
    bath = S.DebyeBath(*BATH)
    x = np.array([0.0, -350.0, -20.5, 1e-3, 17.0, 480.0])
    Cc = TR.debye_correlation(torch.as_tensor(x), BATH[1], BATH[2], BATH[0]).numpy()
    Cr = TR.debye_correlation(torch.as_tensor(x), BATH[1], BATH[2], BATH[0], real=True).numpy()
    for i, xi in enumerate(x):
        assert abs(Cc[i] - bath.corr_func_complex(xi)) <= 1e-13 * abs(bath.corr_func_complex(xi))
        assert abs(Cr[i] - bath.corr_func_real(xi)) <= 1e-13 * abs(bath.corr_func_real(xi))


@pytest.mark.parametrize('X,Y', [('e', 'g'), ('e', 'e')])
def test_torch_redfield_derivatives_gradcheck(X, Y):
    """dL/d(E, J, lam) of the restatement against central differences."""
    H, _ = _random_h(3, 5)
    E = torch.tensor(np.diag(H).copy(), requires_grad=True)
    J = torch.tensor(H[np.triu_indices(3, 1)].copy(), requires_grad=True)
    lam = torch.tensor(BATH[1], dtype=torch.float64, requires_grad=True)

    def f(E, J, lam):
        Hs = G.site_hamiltonian(E, J)[None]
        L = TR.redfield_block(Hs, lam, BATH[2], BATH[0], 12450.0, X, Y, False, False, S.CM_FS)
        return torch.view_as_real(L)

    assert torch.autograd.gradcheck(f, (E, J, lam), eps=1e-4, atol=1e-9, rtol=1e-5)


def test_site_hamiltonian_and_disorder_shifts():
    H, d = _random_h(4, 2)
    iu = np.triu_indices(4, 1)
    Ht = G.site_hamiltonian(torch.as_tensor(np.diag(H).copy()), torch.as_tensor(H[iu].copy()))
    assert np.array_equal(Ht.numpy(), H)
    ham = qa.ElectronicHamiltonian(H, bath=qa.DebyeBath(*BATH), dipoles=d, disorder=100,
                                   random_seed=7)
    shifts = G.disorder_shifts(4, 100, 5, random_seed=7)
    for n, member in enumerate(ham.sample_ensemble(5)):
        assert np.allclose(member.H_1exc, H + np.diag(shifts[n]), rtol=0, atol=1e-12)
