"""Pins the oracle's Liouville layer: index maps, Redfield and HEOM operators.

Against the reference's known answers (tests/test_liouville_space.py of
qspectra), the golden Liouvillians written by the reference, and closed forms
a dropped term, wrong sign or transposed index would break.
"""
import numpy as np
from numpy.testing import assert_allclose, assert_equal
from scipy.special import comb

from oracle import system as S
from oracle import liouville as Lv
from tests.conftest import golden, golden_csr, rel_err


def test_index_maps_known_answers():
    # qspectra tests/test_liouville_space.py:11-37
    assert_equal(Lv.liouville_subspace_index('eg,ge', 'ge', 2), [1, 2, 3, 6])
    assert_equal(Lv.liouville_subspace_index('eg,fe', 'gef', 2), [1, 2, 7, 11])
    assert_equal(Lv.liouville_subspace_index('gg,ee,ff', 'gef', 2), [0, 5, 6, 9, 10, 15])
    assert Lv.all_liouville_subspaces('gef') == 'gg,ge,gf,eg,ee,ef,fg,fe,ff'


def test_liouville_operator_known_answers():
    # qspectra tests/test_liouville_space.py:124-138
    class M(object):
        def liouville_subspace_index(self, s):
            return Lv.liouville_subspace_index(s, 'ge', 1)
    X = np.array([[1, 2], [3, 4]])
    op = Lv.LiouvilleSpaceOperator(X, 'gg,eg,ge,ee->gg', M())
    v = np.array([1, 10, 100, 1000])
    assert_allclose(op.left_multiply(v), [21])
    assert_allclose(op.right_multiply(v), [301])
    assert_allclose(op.commutator(v), [-280])
    assert_allclose(op.expectation_value(v), [21])
    op = Lv.LiouvilleSpaceOperator(X, 'ee->gg,ee', M())
    assert_allclose(op.left_multiply([1]), [0, 4])
    assert_allclose(op.expectation_value([1]), 4)


def _dimer(extra=None, T=S.CM_K * 77):
    g = golden('redfield_dimer')
    return S.ElectronicHamiltonian(g['in_H1'], bath=S.DebyeBath(T, 35, 106),
                                   dipoles=g['in_dipoles'],
                                   energy_spread_extra=float(g['extra']) if extra is None else extra)


def test_redfield_dimer_blocks_match_reference():
    g = golden('redfield_dimer')
    ham = _dimer()
    for sec in (True, False):
        for dic in (True, False):
            m = Lv.LiouvilleModel(ham, unit_convert=S.CM_FS, secular=sec, discard_imag_corr=dic)
            for s in ['gg', 'eg', 'ge', 'ee', 'fg', 'fe', 'gg,ge,eg,ee']:
                ref = g['L_%s_%d%d' % (s.replace(',', '_'), sec, dic)]
                assert rel_err(m.evolve_matrix(s), ref) < 1e-12, (s, sec, dic)
    m = Lv.LiouvilleModel(ham, unit_convert=S.CM_FS, secular=False, evolve_basis='eigen')
    assert rel_err(m.evolve_matrix('ee'), g['Leig_ee']) < 1e-12
    assert float(g['rw_freq']) == m.rw_freq and float(g['time_step']) == m.time_step
    R = Lv.redfield_tensor(m.hamiltonian, 'ge', secular=False)
    assert rel_err(R, g['R_ge']) < 1e-12


def test_redfield_7site_blocks_match_reference():
    g = golden('redfield_7site')
    ham = S.ElectronicHamiltonian(g['in_H1'], bath=S.DebyeBath(S.CM_K * 77, 35, 106),
                                  dipoles=g['in_dipoles'], energy_spread_extra=float(g['extra']))
    m = Lv.LiouvilleModel(ham, unit_convert=S.CM_FS, secular=False)
    for s in ['ge', 'eg', 'ee', 'fe', 'fg']:
        assert rel_err(m.evolve_matrix(s), g['L_' + s]) < 1e-12, s
    ms = Lv.LiouvilleModel(ham, unit_convert=S.CM_FS, secular=True)
    assert rel_err(ms.evolve_matrix('ee'), g['Lsec_ee']) < 1e-12


def test_redfield_closed_forms():
    ham = _dimer()
    m = Lv.LiouvilleModel(ham, unit_convert=1.0, secular=False)
    R = Lv.redfield_tensor(m.hamiltonian, 'gef', secular=False)
    # trace preservation: sum_a R_aacd = 0 (redfield.py:65-69)
    assert np.max(np.abs(np.einsum('aacd->cd', R))) < 1e-11 * np.max(np.abs(R))
    # non-secular 'eg' block in the eigenbasis is -i*Omega - Gs restricted to e
    E, U = m.hamiltonian.eig('gef')
    me = Lv.LiouvilleModel(ham, unit_convert=1.0, secular=False, evolve_basis='eigen')
    Leg = me.evolve_matrix('eg')
    K = [U.T @ c @ U for c in m.hamiltonian.system_bath_couplings('gef')]
    C = np.array([[ham.bath.corr_func_complex(a - b) for b in E] for a in E])
    Gam = np.einsum('iab,icd,dc->abcd', K, K, C)
    Gs = np.einsum('abbc->ac', Gam)[1:3, 1:3]
    expect = -1j * np.diag(E[1:3] - E[0]) - Gs
    assert rel_err(Leg, expect) < 1e-12
    # monomer with real C: 'eg' decays at rate 2*lam*T/gamma (Lorentzian HWHM)
    mono = S.ElectronicHamiltonian([[12500.]], bath=S.DebyeBath(S.CM_K * 77, 35, 106),
                                   dipoles=[[1, 0, 0]])
    mm = Lv.LiouvilleModel(mono, unit_convert=1.0, discard_imag_corr=True)
    T = S.CM_K * 77
    assert_allclose(mm.evolve_matrix('eg'), [[-2 * 35 * T / 106]], rtol=1e-13)
    # with complex C the line shifts by -lambda (imag part of C(0) = -lambda)
    mc = Lv.LiouvilleModel(mono, unit_convert=1.0, discard_imag_corr=False)
    assert_allclose(mc.evolve_matrix('eg'), [[-35 * (2 * T / 106 - 1j)]], rtol=1e-13)


def test_heom_matches_reference():
    g = golden('heom')
    ham = _dimer(T=S.CM_K * 300)
    hr = ham.in_rotating_frame()
    for name, kw in [('dimer', dict(level_cutoff=3, K=1)),
                     ('dimer_mod', dict(level_cutoff=3, K=1, modified_HEOM=True)),
                     ('dimer_nolt', dict(level_cutoff=3, K=2, low_temp_corr=False))]:
        for s in ('ee', 'gg,ge,eg,ee', 'eg'):
            L = Lv.heom_evolve_matrix(hr, 'ge', s, S.CM_FS, **kw)
            ref = golden_csr(g, '%s_%s' % (name, s.replace(',', '_')))
            assert abs(L - ref).max() < 1e-13 * abs(ref).max(), (name, s)
        ados, _ = Lv.ado_mappings(2, kw['K'], 3)
        assert_equal(np.array([a.flatten() for a in ados]), g[name + '_ados'])
    g7 = golden('redfield_7site')
    h7 = S.ElectronicHamiltonian(g7['in_H1'], bath=S.DebyeBath(S.CM_K * 300, 35, 106),
                                 energy_spread_extra=float(g7['extra'])).in_rotating_frame()
    L = Lv.heom_evolve_matrix(h7, 'ge', 'ee', S.CM_FS, K=1, level_cutoff=3)
    ref = golden_csr(g, 'agg7_ee')
    assert abs(L - ref).max() < 1e-13 * abs(ref).max()


def test_heom_closed_forms():
    # ADO count C(N(K+1) + L, L) with level_cutoff = L + 1 (exclusive cutoff)
    for N, K, L in [(2, 1, 2), (7, 1, 4), (3, 2, 3)]:
        ados, _ = Lv.ado_mappings(N, K, L + 1)
        assert len(ados) == comb(N * (K + 1) + L, L, exact=True)
    ham = _dimer(T=S.CM_K * 300).in_rotating_frame()
    # level_cutoff = 1: -i L_H - Delta * sum_j (P^L + P^R - 2 P^L P^R)
    L1 = Lv.heom_evolve_matrix(ham, 'ge', 'ee', 1.0, K=1, level_cutoff=1).toarray()
    nu = Lv.matsubara_frequencies(5001, 106, S.CM_K * 300)
    c = Lv.corr_func_coeffs(5001, 106, S.CM_K * 300, 35, nu)
    Delta = np.sum((c / nu)[2:])
    idx = Lv.liouville_subspace_index('ee', 'ge', 2)
    mesh = np.ix_(idx, idx)
    H = ham.H('ge')
    expect = -1j * Lv.super_commutator_matrix(H)[mesh]
    for p in ham.system_bath_couplings('ge'):
        pl, pr = Lv.super_left_matrix(p)[mesh], Lv.super_right_matrix(p)[mesh]
        expect = expect - Delta * (pl + pr - 2 * pl @ pr)
    assert rel_err(L1, expect) < 1e-13
    # lambda = 0: c_k = 0 and Delta = 0, so no ADO is fed from ADO 0 and ADO 0
    # evolves unitarily
    h0 = S.ElectronicHamiltonian(ham.H_1exc, bath=S.DebyeBath(S.CM_K * 300, 0.0, 106),
                                 energy_spread_extra=1.0).in_rotating_frame()
    L0 = Lv.heom_evolve_matrix(h0, 'ge', 'ee', 1.0, K=1, level_cutoff=3).toarray()
    assert_allclose(L0[:4, :4], -1j * Lv.super_commutator_matrix(h0.H('ge'))[mesh], atol=1e-14)
    assert np.max(np.abs(L0[4:, :4])) == 0
