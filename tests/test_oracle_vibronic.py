"""The oracle's vibronic Hamiltonian (qspectra/hamiltonian.py:633-807) and the
Redfield / unitary paths on it, pinned by the reference's known-answer tests
(tests/test_hamiltonian.py:192-231), closed forms of the displaced harmonic
oscillator and the reference's golden outputs (tests/golden/vibronic.npz),
plus the product's host-side vibronic API (no GPU needed)."""
import numpy as np
import scipy.linalg
from numpy.testing import assert_allclose

from oracle import system as S
from oracle import liouville as Lv
from oracle import propagate as P
from tests.conftest import golden, rel_err

TIGHT = dict(rtol=1e-12, atol=1e-14, nsteps=10 ** 7)


class _Bath2(object):
    temperature = 2


def _test_monomer():
    """tests/test_hamiltonian.py:193-197."""
    el = S.ElectronicHamiltonian([[1.0]], bath=_Bath2(), dipoles=[[1, 0, 0]], disorder=0)
    return S.VibronicHamiltonian(el, [2], [10], [[5]])


def _jonas(g, n=2):
    bath = S.DebyeBath(float(g['in_T']), float(g['in_lam']), float(g['in_gamma']))
    el = S.ElectronicHamiltonian(g['in_H_el'][:n, :n], bath=bath, dipoles=g['in_dipoles'][:n])
    return S.VibronicHamiltonian(el, g['in_levels'][:n], g['in_vib'][:n],
                                 g['in_couplings'][:n, :n])


def test_reference_known_answers():
    h = _test_monomer()
    assert h.n_sites == 1 and h.n_vibrational_states == 2 and h.n_states('gef') == 4
    assert_allclose(h.H('ge'), [[0, 0, 0, 0], [0, 10, 0, 0], [0, 0, 1, 5], [0, 0, 5, 11]])
    assert abs(h.transition_energy - 6) < 1e-12
    assert_allclose(h.thermal_state('g'), 1 / (1 + np.exp(-5)) * np.diag([1, np.exp(-5)]),
                    atol=1e-15)
    assert_allclose(h.system_bath_couplings('ge'),
                    [[[0, 0, 0, 0], [0, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]]])
    # rotating frame and idempotence (tests/test_hamiltonian.py:20-37)
    assert_allclose(h.in_rotating_frame(2).E('e') + 2, h.E('e'))
    assert_allclose(h.in_rotating_frame(2).in_rotating_frame(3).E('e') + 3, h.E('e'))


def test_displaced_oscillator_closed_form():
    """e block of a monomer: eps + w n + c (b + b^+) is a displaced
    oscillator with levels eps + w k - c^2 / w; with many levels the low
    eigenvalues converge to them (a dropped sqrt, a wrong neighbour index or
    a sign error in the coupling breaks this)."""
    eps, w, c = 100.0, 10.0, 7.0
    el = S.ElectronicHamiltonian([[eps]], dipoles=[[1, 0, 0]])
    h = S.VibronicHamiltonian(el, [40], [w], [[c]])
    E = scipy.linalg.eigvalsh(h.H('e'))
    assert_allclose(E[:5], eps + w * np.arange(5) - c ** 2 / w, atol=1e-9)
    # ground block: bare oscillator; two uncoupled modes: sums of levels
    h2 = S.VibronicHamiltonian(el, [3, 2], [10., 3.], [[0., 0.]])
    Eg = np.sort(np.add.outer(10. * np.arange(3), 3. * np.arange(2)).ravel())
    assert_allclose(np.sort(np.diag(h2.H('g'))), Eg)
    assert np.all(h2.H('g') == np.diag(np.diag(h2.H('g'))))
    # product basis order: the last mode is the fastest index
    assert_allclose(np.diag(h2.H('g')), [0, 3, 10, 13, 20, 23])


def test_jonas_dimer_matches_reference():
    g = golden('vibronic')
    ham = _jonas(g)
    assert_allclose(ham.H('gef'), g['jonas_H_gef'], atol=1e-12)
    m = Lv.LiouvilleModel(ham, hilbert_subspace='gef', unit_convert=S.CM_FS,
                          discard_imag_corr=True)
    assert abs(m.rw_freq - float(g['jonas_rw'])) < 1e-9
    assert abs(m.time_step - float(g['jonas_time_step'])) < 1e-12 * m.time_step
    assert_allclose(m.hamiltonian.E('gef'), g['jonas_E_gef'], atol=1e-9)
    assert_allclose(m.hamiltonian.thermal_state('g'), g['jonas_thermal_gg'], atol=1e-14)
    for s in ['gg', 'ee', 'eg', 'fe', 'ff']:
        L = m.evolve_matrix(s)
        ref = g['jonas_L_' + s]
        assert abs(L - ref).max() < 1e-12 * abs(ref).max(), s
    e0 = np.zeros(8)
    e0[0] = 1
    t, rho = P.simulate_dynamics(m, e0, 1000., integrator='zvode', **TIGHT)
    assert_allclose(t, g['jonas_dyn_t'])
    assert rel_err(rho, g['jonas_dyn_rho']) < 1e-9
    f, X = P.absorption_spectra(m, float(g['jonas_abs_T']), integrator='zvode', **TIGHT)
    assert_allclose(f, g['jonas_abs_f'])
    assert rel_err(X, g['jonas_abs_X']) < 1e-9


def test_jonas_third_order_gsb_isotropic_matches_reference():
    g = golden('vibronic')
    m = Lv.LiouvilleModel(_jonas(g), hilbert_subspace='gef', unit_convert=S.CM_FS,
                          discard_imag_corr=True)
    t, R = P.isotropic_average_4th(
        lambda polarization: P.third_order_response(
            m, 200.5, population_times=g['jonas_R3_t2'], geometry='-++',
            polarization=polarization, include_signal='GSB', integrator='zvode', **TIGHT),
        'xxxx')
    assert_allclose(t[0], g['jonas_R3_t1'])
    assert rel_err(R, g['jonas_R3']) < 1e-8


def test_vibronic_monomers_match_reference():
    g = golden('vibronic')
    mm = Lv.LiouvilleModel(_jonas(g, 1), hilbert_subspace='ge', unit_convert=S.CM_FS)
    f, X = P.absorption_spectra(mm, float(g['jmono_abs_T']), integrator='zvode', **TIGHT)
    assert_allclose(f, g['jmono_abs_f'])
    assert rel_err(X, g['jmono_abs_X']) < 1e-9
    el = S.ElectronicHamiltonian([[11500.]], dipoles=[[1., 0., 0.]])
    uh = S.VibronicHamiltonian(el, [5], [200.], [[-100.]])
    assert_allclose(uh.H('ge'), g['umono_H_ge'], atol=1e-12)
    um = Lv.LiouvilleModel(uh, hilbert_subspace='ge', unit_convert=S.CM_FS, kind='unitary')
    assert abs(um.time_step - float(g['umono_time_step'])) < 1e-12 * um.time_step
    e0 = np.zeros(5)
    e0[0] = 1
    t, rho = P.simulate_dynamics(um, e0, 1000., integrator='zvode', **TIGHT)
    assert rel_err(rho, g['umono_dyn_rho']) < 1e-9
    f, X = P.absorption_spectra(um, float(g['umono_abs_T']), correlation_decay_time=1000.,
                                integrator='zvode', **TIGHT)
    assert rel_err(X, g['umono_abs_X']) < 1e-9


def test_product_vibronic_host_api():
    """The product's VibronicHamiltonian on the host: matrices, labels and
    operators (tests/test_hamiltonian.py:199-231) and agreement with the
    oracle's element-by-element construction."""
    import qspectra_amd as qa
    el = qa.ElectronicHamiltonian([[1.0]], bath=_Bath2(), dipoles=[[1, 0, 0]], disorder=0)
    h = qa.VibronicHamiltonian(el, [2], [10], [[5]])
    assert h.n_sites == 1 and h.n_vibrational_states == 2 and h.n_states('gef') == 4
    assert_allclose(h.H('ge'), [[0, 0, 0, 0], [0, 10, 0, 0], [0, 0, 1, 5], [0, 0, 5, 11]])
    assert_allclose(h.system_bath_couplings('ge'),
                    [[[0, 0, 0, 0], [0, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]]])
    assert h.basis_labels('gef', braket=1) == ['|0>|0>', '|0>|1>', '|1>|0>', '|1>|1>']
    assert h.basis_labels('gef') == [('0', '0'), ('0', '1'), ('1', '0'), ('1', '1')]
    el1 = qa.ElectronicHamiltonian([[1.0]], bath=None, dipoles=[[1, 0, 0]], disorder=0,
                                   site_labels=['one'])
    hl = qa.VibronicHamiltonian(el1, [2], [10], [[5]])
    assert hl.basis_labels('gef', braket=1) == ['|g>|0>', '|g>|1>', '|one>|0>', '|one>|1>']
    assert hl.basis_labels('gef') == [('g', '0'), ('g', '1'), ('one', '0'), ('one', '1')]
    g = golden('vibronic')
    elj = qa.ElectronicHamiltonian(g['in_H_el'], dipoles=g['in_dipoles'],
                                   bath=qa.DebyeBath(float(g['in_T']), float(g['in_lam']),
                                                     float(g['in_gamma'])))
    hj = qa.VibronicHamiltonian(elj, g['in_levels'], g['in_vib'], g['in_couplings'])
    oj = _jonas(g)
    assert_allclose(hj.H('gef'), g['jonas_H_gef'], atol=1e-12)
    for s in ('g', 'e', 'f', 'gef'):
        assert_allclose(hj.dipole_operator(s, 'x'), oj.dipole_operator(s, 'x'))
        assert_allclose(hj.system_bath_couplings(s), oj.system_bath_couplings(s))
    assert hj == qa.VibronicHamiltonian(elj, g['in_levels'], g['in_vib'], g['in_couplings'])
    blk = hj.dense_block('e', 'gef')
    assert_allclose(blk, g['jonas_H_gef'][4:12, 4:12], atol=1e-12)
