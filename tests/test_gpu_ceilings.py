"""Blocks beyond the register-resident and dense forms: field-driven
propagation of Liouville subspaces above 64 elements (the LDS-resident driven
kernel) and hierarchies on multi-block subspaces above 64 elements (the
Hilbert-commutator form of the HEOM stage kernel), against the oracle."""
import numpy as np
import pytest

import qspectra_amd as qa
from qspectra_amd import _native
from oracle import system as S
from oracle import liouville as Lv
from oracle import propagate as P
from tests.conftest import golden, rel_err

pytestmark = pytest.mark.gpu


def _aggregate(n, seed):
    rng = np.random.RandomState(seed)
    J = rng.normal(0.0, 60.0, (n, n))
    H = np.triu(J, 1)
    H = H + H.T + np.diag(rng.uniform(12300, 12600, n))
    d = rng.normal(size=(n, 3))
    return H, d / np.linalg.norm(d, axis=1, keepdims=True)


def test_driven_blocks_above_64(gpu):
    """An 8-site pump in 'gg,ge,eg,ee' (M = 81): FP64 against the oracle's
    driven RK4 on the same schedule (1e-10), MIXED 1e-6."""
    H, d = _aggregate(8, 3)
    T = S.CM_K * 77
    mp = qa.RedfieldModel(qa.ElectronicHamiltonian(H, bath=qa.DebyeBath(T, 35, 106), dipoles=d),
                          hilbert_subspace='ge', unit_convert=qa.CM_FS, secular=False)
    mo = Lv.LiouvilleModel(S.ElectronicHamiltonian(H, bath=S.DebyeBath(T, 35, 106), dipoles=d),
                           hilbert_subspace='ge', unit_convert=S.CM_FS, secular=False)
    pp = qa.GaussianPulse(12450., 20., t_peak=0., scale=0.01, freq_convert=qa.CM_FS)
    po = S.GaussianPulse(12450., 20., t_peak=0., scale=0.01, freq_convert=S.CM_FS)
    dt = 0.5
    t, y = qa.simulate_pump(mp, pp, 'x', time_extra=20., dt=dt)
    to, yo = P.simulate_pump(mo, po, 'x', time_extra=20., dt=dt)
    np.testing.assert_allclose(t, to)
    assert np.asarray(y).shape[-1] == 81
    assert rel_err(np.asarray(y), yo) < 1e-10
    _, ym = qa.simulate_pump(mp, pp, 'x', time_extra=20., dt=dt, precision='mixed')
    assert rel_err(np.asarray(ym), yo) < 1e-6


def _heom_pair(n, seed, **kw):
    H, d = _aggregate(n, seed)
    T = S.CM_K * 300
    hp = qa.ElectronicHamiltonian(H, bath=qa.DebyeBath(T, 35, 106), dipoles=d)
    ho = S.ElectronicHamiltonian(H, bath=S.DebyeBath(T, 35, 106), dipoles=d)
    return (qa.HEOMModel(hp, hilbert_subspace='gef', unit_convert=qa.CM_FS, **kw),
            Lv.HEOMOracleModel(ho, hilbert_subspace='gef', unit_convert=S.CM_FS, **kw))


@pytest.mark.parametrize('subspace', ['gg,ge,eg,ee,fe,ef,ff', 'gg,ge,eg,ee,gf,fg,ef,fe,ff'])
def test_heom_multiblock_above_64(gpu, subspace):
    """4 sites, 'gef' blocks (M = 109 / 121): the hierarchy's right-hand side
    in both pictures against the oracle's generator (1e-13) and RK4
    trajectories against the oracle's RK4 on the same schedule (1e-10)."""
    mp, mo = _heom_pair(4, 5, level_cutoff=2, K=1)
    rng = np.random.RandomState(1)
    for heis in (False, True):
        eom = mp.equation_of_motion(subspace, heisenberg_picture=heis)
        assert eom.M > 64 and eom.hilbert is not None
        Lo = mo.evolve_matrix(subspace, heisenberg_picture=heis)
        y = rng.normal(size=(1, eom.size)) + 1j * rng.normal(size=(1, eom.size))
        got = eom(0.0, y)
        ref = Lo.dot(y[0])
        assert rel_err(got[0], ref) < 1e-13, heis
    eom = mp.equation_of_motion(subspace)
    Lo = mo.evolve_matrix(subspace)
    y0 = np.zeros(eom.size, complex)
    rho = np.zeros((11, 11), complex)
    rho[1, 1] = 0.6
    rho[2, 2] = 0.4
    rho[1, 2] = rho[2, 1] = 0.3
    full = rho.reshape(-1, order='F')
    idx = mo.lspace.liouville_subspace_index(subspace)
    y0[:idx.size] = full[idx]
    t = np.arange(0, 20.0, 2.0)
    for prec, tol in (('fp64', 1e-10), ('mixed', 1e-6)):
        got = qa.integrate(eom, y0, t, dt=0.5, precision=prec)
        ref = P.integrate_rk4(Lo, y0, t, dt=0.5)
        assert rel_err(got, ref) < tol, prec


def test_heom_hilbert_form_equals_dense_form(gpu, monkeypatch):
    """The Hilbert-commutator form (forced on the dimer's 'gg,ge,eg,ee')
    equals the dense form."""
    from qspectra_amd.dynamics import heom as Hm
    g = golden('redfield_dimer')
    ham = qa.ElectronicHamiltonian(g['in_H1'], bath=qa.DebyeBath(qa.CM_K * 300, 35, 106),
                                   dipoles=g['in_dipoles'], energy_spread_extra=float(g['extra']))
    dense = qa.HEOMModel(ham, hilbert_subspace='gef', unit_convert=qa.CM_FS, level_cutoff=3, K=1)
    e1 = dense.equation_of_motion('gg,ge,eg,ee', heisenberg_picture=True)
    assert e1.hilbert is None
    monkeypatch.setattr(Hm, 'DENSE_BLOCK_MAX', 0)
    forced = qa.HEOMModel(ham, hilbert_subspace='gef', unit_convert=qa.CM_FS, level_cutoff=3, K=1)
    e2 = forced.equation_of_motion('gg,ge,eg,ee', heisenberg_picture=True)
    assert e2.hilbert is not None
    rng = np.random.RandomState(2)
    y = rng.normal(size=(1, e1.size)) + 1j * rng.normal(size=(1, e1.size))
    assert rel_err(e2(0.0, y), e1(0.0, y)) < 1e-13
