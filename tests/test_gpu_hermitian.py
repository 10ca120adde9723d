"""Complex-Hermitian Hamiltonians (the reference accepts any Hermitian
matrix: check_hermitian, qspectra/hamiltonian.py:18-22, and diagonalises it
with eigh, :310-328).

- qsx_block_eigh_hermitian / qsx_dense_eigh_hermitian: eigenvalues against
  numpy's eigvalsh of the oracle's blocks (1e-12 relative), residual
  |H U - U E| and |U^H U - 1| (1e-12), the phase convention (largest
  component real and positive), and on a real H the real solver's (E, U).
- UnitaryModel: the Liouvillian -i [H, .] against the oracle's (1e-13: the
  frame shift of 12000 cm^-1 is subtracted in another order),
  Schroedinger and Heisenberg blocks, electronic and vibronic; an RK4
  trajectory against the oracle's RK4 on that matrix (1e-11).
- thermal_state against the oracle's expm(-H/T) / tr (1e-12).
- HEOMModel: the hierarchy right-hand side against the oracle's sparse
  generator built from the same complex H (1e-13); ZOFEModel: the ZOFE
  right-hand side against the oracle's (1e-12).
- RedfieldModel raises NotImplementedError: the reference's kron(U, U)
  basis change makes its result depend on the eigenvector phases.
"""
import numpy as np
import pytest
import torch

import qspectra_amd as qa
from qspectra_amd import _native
from oracle import system as O, liouville as OL, propagate as P
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _complex_H(n, seed):
    rs = np.random.RandomState(seed)
    A = 80.0 * (rs.randn(n, n) + 1j * rs.randn(n, n))
    return (A + A.conj().T) / 2 + np.diag(12000.0 + 300.0 * rs.randn(n))


def _check_eigensystem(Hb, E, U, tol=1e-12):
    scale = max(np.abs(np.linalg.eigvalsh(Hb)).max(), 1.0)
    assert np.abs(E - np.linalg.eigvalsh(Hb)).max() < tol * scale
    assert np.abs(Hb @ U - U * E).max() < tol * scale
    assert np.abs(U.conj().T @ U - np.eye(len(E))).max() < tol
    big = np.argmax(np.abs(U), axis=0)
    lead = U[big, np.arange(len(E))]
    assert np.all(lead.imag == 0) and np.all(lead.real > 0)


def test_block_eigh_hermitian_matches_numpy(gpu):
    n, B = 6, 3
    Hs = [_complex_H(n, s) for s in range(B)]
    rw = np.array([0.0, 12000.0, 11500.0])
    H1 = torch.as_tensor(np.stack(Hs)).cuda()
    rwt = torch.as_tensor(rw).cuda()
    for X, sub in enumerate('gef'):
        E, U = _native.block_eigh(H1, X, rwt)
        assert U.dtype == torch.complex128
        E, U = E.cpu().numpy(), U.cpu().numpy()
        for b in range(B):
            Hb = O.ElectronicHamiltonian(Hs[b]).H(sub, rotating=False)
            Hb = Hb - X * rw[b] * np.eye(len(Hb))
            _check_eigensystem(Hb, E[b], U[b])


def test_hermitian_solver_on_real_input_equals_real_solver(gpu):
    rs = np.random.RandomState(5)
    A = rs.randn(3, 7, 7) * 90.0
    H = (A + A.transpose(0, 2, 1)) / 2 + np.diag(12000.0 + 200.0 * rs.randn(7))
    Hr = torch.as_tensor(H).cuda()
    Hc = torch.as_tensor(H.astype(complex)).cuda()
    for X in (1, 2):
        Er, Ur = _native.block_eigh(Hr, X)
        Ec, Uc = _native.block_eigh(Hc, X)
        assert rel_err(Ec.cpu().numpy(), Er.cpu().numpy()) < 1e-13
        assert np.abs(Uc.cpu().numpy() - Ur.cpu().numpy()).max() < 1e-11


@pytest.mark.parametrize('nb', [1, 2, 17, 64])
def test_dense_eigh_hermitian(gpu, nb):
    rs = np.random.RandomState(nb)
    A = rs.randn(2, nb, nb) + 1j * rs.randn(2, nb, nb)
    H = (A + A.conj().transpose(0, 2, 1)) / 2
    E, U = _native.dense_eigh(torch.as_tensor(H).cuda())
    for b in range(2):
        _check_eigensystem(H[b], E[b].cpu().numpy(), U[b].cpu().numpy())


def _models(H, rw=12000.0, subspace='gef', T=qa.CM_K * 300):
    ham = qa.ElectronicHamiltonian(H, bath=qa.DebyeBath(T, 35.0, 106.0),
                                   dipoles=[[1.0, 0, 0]] * len(H))
    oham = O.ElectronicHamiltonian(H, bath=O.DebyeBath(T, 35.0, 106.0))
    return ham, oham


def test_unitary_liouvillian_and_dynamics(gpu):
    H = _complex_H(4, 11)
    ham, oham = _models(H)
    m = qa.UnitaryModel(ham, rw_freq=12000.0, hilbert_subspace='gef', unit_convert=qa.CM_FS)
    om = OL.LiouvilleModel(oham, 12000.0, 'gef', qa.CM_FS, kind='unitary')
    assert rel_err(m.evolution_super_operator, om.evolution_super_operator) < 1e-13
    for s in ('eg', 'gg,ge,eg,ee', 'fe'):
        for heis in (False, True):
            L = m.liouvillian(s, heisenberg_picture=heis)[0].cpu().numpy()
            assert rel_err(L, om.evolve_matrix(s, heis)) < 1e-13, (s, heis)
    # a trajectory through the device stepper against the oracle's RK4
    s = 'gg,ge,eg,ee'
    L = om.evolve_matrix(s)
    rs = np.random.RandomState(2)
    y0 = rs.randn(L.shape[0]) + 1j * rs.randn(L.shape[0])
    t = np.arange(0, 40.0, 2.0)
    got = qa.integrate(m.equation_of_motion(s), y0, t, dt=0.25)
    ref = P.integrate_rk4(L, y0, t, dt=0.25)
    assert rel_err(got, ref) < 1e-11


def test_thermal_state_and_eig(gpu):
    H = _complex_H(5, 3)
    ham, oham = _models(H)
    for sub in ('g', 'e', 'ge', 'gef'):
        assert rel_err(ham.thermal_state(sub), oham.thermal_state(sub)) < 1e-12, sub
        assert rel_err(ham.E(sub), oham.E(sub)) < 1e-13, sub
        E, U = ham.eig(sub)
        Hs = oham.H(sub, rotating=False)
        assert np.abs(Hs @ U - U * E).max() <= 1e-12 * max(np.abs(E).max(), 1.0)
    hr = ham.in_rotating_frame(12000.0)
    assert rel_err(hr.E('e'), oham.E('e') - 12000.0) < 1e-12


def test_vibronic_complex_unitary(gpu):
    H = _complex_H(2, 7)
    el = qa.ElectronicHamiltonian(H, bath=qa.DebyeBath(qa.CM_K * 77, 35.0, 106.0))
    oel = O.ElectronicHamiltonian(H, bath=O.DebyeBath(qa.CM_K * 77, 35.0, 106.0))
    levels, energies, c = [2, 3], [600.0, 900.0], [[200.0, 0.0], [0.0, 300.0]]
    vib = qa.VibronicHamiltonian(el, levels, energies, c)
    ovib = O.VibronicHamiltonian(oel, levels, energies, c)
    m = qa.UnitaryModel(vib, rw_freq=12000.0, hilbert_subspace='ge', unit_convert=qa.CM_FS)
    om = OL.LiouvilleModel(ovib, 12000.0, 'ge', qa.CM_FS, kind='unitary')
    for s in ('eg', 'ee'):
        L = m.liouvillian(s)[0].cpu().numpy()
        assert rel_err(L, om.evolve_matrix(s)) < 1e-13, s
    assert rel_err(vib.E('ge'), ovib.E('ge')) < 1e-12


def test_heom_rhs_complex_hamiltonian(gpu):
    H = _complex_H(2, 13)
    ham, oham = _models(H, subspace='ge')
    m = qa.HEOMModel(ham, rw_freq=12000.0, hilbert_subspace='ge', unit_convert=qa.CM_FS,
                     level_cutoff=3, K=1)
    oh = oham.in_rotating_frame(12000.0)
    rs = np.random.RandomState(4)
    for s in ('ee', 'gg,ge,eg,ee', 'eg'):
        L = OL.heom_evolve_matrix(oh, 'ge', s, qa.CM_FS, 1, 3)
        y = rs.randn(L.shape[0]) + 1j * rs.randn(L.shape[0])
        for heis in (False, True):
            eom = m.equation_of_motion(s, heisenberg_picture=heis)
            ref = (L.T if heis else L) @ y
            assert rel_err(eom(0, y), ref) < 1e-13, (s, heis)


def test_redfield_rejects_complex_hamiltonian(gpu):
    ham, _ = _models(_complex_H(3, 1))
    m = qa.RedfieldModel(ham, rw_freq=12000.0, hilbert_subspace='ge', unit_convert=qa.CM_FS)
    with pytest.raises(NotImplementedError, match='kron'):
        m.liouvillian('eg')


def test_zofe_rhs_complex_hamiltonian(gpu):
    from oracle import zofe as Z
    from tests.conftest import golden
    g = golden('zofe')
    n = 2
    on = np.ones(n, complex)
    pm = [np.outer(g[k], on) for k in ('pm_Omega', 'pm_gamma', 'pm_huang')]
    H = _complex_H(n, 17)
    hp = qa.ElectronicHamiltonian(H, bath=qa.PseudomodeBath(16, *pm), dipoles=g['in_dipoles'])
    ho = O.ElectronicHamiltonian(H, bath=Z.PseudomodeBath(16, *pm), dipoles=g['in_dipoles'])
    mp = qa.ZOFEModel(hp, rw_freq=12000.0, hilbert_subspace='ge', unit_convert=qa.CM_FS)
    mo = Z.ZOFEOracleModel(ho, rw_freq=12000.0, hilbert_subspace='ge', unit_convert=O.CM_FS)
    eom = mp.equation_of_motion('ge')
    rs = np.random.RandomState(6)
    y = (rs.randn(3, mp.state_size) + 1j * rs.randn(3, mp.state_size)) * 0.3
    got = _native.zofe_apply(eom, _native.to_device(y[None])).cpu().numpy()[0]
    assert rel_err(got, np.array([mo.rhs(v) for v in y])) < 1e-12
