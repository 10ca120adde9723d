"""Pins the oracle's integrators, Fourier transform and response drivers.

RK4 against the textbook degree-4 Taylor propagator and its expm limit; the
zvode restatement and end-to-end spectra against trajectories and spectra the
reference wrote (golden fixtures); FFT against the closed forms the
reference's tests/test_simulate_utils.py uses.
"""
import numpy as np
import scipy.linalg
from numpy.testing import assert_allclose

from oracle import system as S
from oracle import liouville as Lv
from oracle import propagate as P
from tests.conftest import golden, golden_csr, rel_err

TIGHT = dict(rtol=1e-12, atol=1e-14, nsteps=10 ** 7)


def _agg7(T=S.CM_K * 77, **kw):
    g = golden('redfield_7site')
    return S.ElectronicHamiltonian(g['in_H1'], bath=S.DebyeBath(T, 35, 106),
                                   dipoles=g['in_dipoles'], energy_spread_extra=float(g['extra']),
                                   **kw)


def _dimer():
    g = golden('redfield_dimer')
    return S.ElectronicHamiltonian(g['in_H1'], bath=S.DebyeBath(S.CM_K * 77, 35, 106),
                                   dipoles=g['in_dipoles'], energy_spread_extra=float(g['extra']))


def test_rk4_is_taylor_propagator_and_converges_to_expm():
    rs = np.random.RandomState(0)
    L = rs.randn(6, 6) + 1j * rs.randn(6, 6)
    y0 = rs.randn(6) + 1j * rs.randn(6)
    h = 0.05
    hL = h * L
    P4 = np.eye(6) + hL + hL @ hL / 2 + hL @ hL @ hL / 6 + hL @ hL @ hL @ hL / 24
    assert rel_err(P.rk4_step(L, y0, h), P4 @ y0) < 1e-14
    t = np.arange(0, 2.0001, 0.25)
    exact = np.array([scipy.linalg.expm(ti * L) @ y0 for ti in t])
    errs = [rel_err(P.integrate_rk4(L, y0, t, dt=0.25 / s), exact) for s in (4, 8)]
    assert errs[1] < errs[0] / 14  # fourth order: 2x steps -> ~16x smaller error


def test_fft_closed_forms_and_symmetrize():
    # qspectra tests/test_simulate_utils.py:26-81
    a = (1.0 + 0.3j) / 100
    t = np.linspace(0, 1000, num=10000)
    f = (t > 0) * np.exp(-a * t)
    v, X = P.fourier_transform(t, f)
    F = 1.0 / (a - 1j * v)
    assert_allclose(X, F, atol=0.002 * abs(1 / a))
    v, X = P.fourier_transform(t, f, sign=-1, rw_freq=5)
    assert_allclose(X, 1.0 / (a - 1j * (-(v - 5))), atol=0.002 * abs(1 / a))
    assert_allclose(P.symmetrize([0, 1, 2], [1, 1, 1]), ([-2, -1, 0, 1, 2], [0, 0, 1, 1, 1]))
    assert_allclose(P.symmetrize([-1.25, -0.25], [1, 1]), ([-1.25, -0.25, 0.75], [1, 1, 0]))
    assert_allclose(P.symmetrize([0.5, 1.5], [1, 1]), ([-1.5, -0.5, 0.5, 1.5], [0, 0, 1, 1]))


def test_zvode_driver_matches_reference_trajectories():
    g = golden('dynamics')
    m = Lv.LiouvilleModel(_agg7(), unit_convert=S.CM_FS, secular=False)
    t, rho = P.simulate_dynamics(m, [1, 0, 0, 0, 0, 0, 0], 200, integrator='zvode', **TIGHT)
    assert_allclose(t, g['t'])
    assert rel_err(rho, g['rho']) < 1e-9
    mu = Lv.LiouvilleModel(_agg7(), unit_convert=S.CM_FS, kind='unitary')
    t, rho = P.simulate_dynamics(mu, [0, 1, 0, 0, 0, 0, 0], 100, integrator='zvode', **TIGHT)
    assert rel_err(rho, g['unitary_rho']) < 1e-9


def test_heom_zvode_dynamics_matches_reference():
    g = golden('heom')
    gd = golden('redfield_dimer')
    ham = S.ElectronicHamiltonian(gd['in_H1'], bath=S.DebyeBath(S.CM_K * 300, 35, 106),
                                  energy_spread_extra=float(gd['extra'])).in_rotating_frame()
    L = golden_csr(g, 'dimer_ee')
    Lo = Lv.heom_evolve_matrix(ham, 'ge', 'ee', S.CM_FS, K=1, level_cutoff=3)
    assert abs(L - Lo).max() < 1e-13 * abs(L).max()
    y0 = np.zeros(L.shape[0], complex)
    y0[0] = 1
    y = P.integrate_zvode(lambda t, v: Lo.dot(v), y0, g['dyn_t'], **TIGHT)
    assert rel_err(y, g['dyn_y']) < 1e-9
    # RK4 with a fine step agrees to its own discretisation error
    yr = P.integrate_rk4(Lo, y0, g['dyn_t'], dt=float(g['dyn_time_step']) / 8)
    assert rel_err(yr, g['dyn_y']) < 1e-6


def test_linear_spectra_match_reference():
    g = golden('spectra_linear')
    md = Lv.LiouvilleModel(_dimer(), unit_convert=S.CM_FS)
    t, Sg = P.linear_response(md, 'gg->eg->gg', 511.5, integrator='zvode', **TIGHT)
    assert rel_err(Sg, g['dimer_S']) < 1e-9
    f, X = P.absorption_spectra(md, 511.5, integrator='zvode', **TIGHT)
    assert_allclose(f, g['dimer_f'])
    assert rel_err(X, g['dimer_X']) < 1e-9
    # the fixed-step RK4 oracle at dt = 1/8 fs agrees to its truncation error
    f, X = P.absorption_spectra(md, 511.5, integrator='rk4', dt=1.0 / 8)
    assert rel_err(X, g['dimer_X']) < 1e-8
    iso = P.isotropic_average_2nd(
        lambda polarization: P.absorption_spectra(md, 511.5, polarization=polarization,
                                                  integrator='zvode', **TIGHT),
        polarization='xx')
    assert rel_err(iso[1], g['dimer_X_iso']) < 1e-9
    m7 = Lv.LiouvilleModel(_agg7(disorder=100 / S.GAUSSIAN_SD_FWHM), unit_convert=S.CM_FS,
                           secular=False)
    f, X = P.absorption_spectra(m7, 2 * 255.5, integrator='zvode', **TIGHT)
    assert rel_err(X, g['agg7_X']) < 1e-9
    f, X = P.ensemble_average(
        lambda mm: P.absorption_spectra(mm, 2 * 127.5, integrator='zvode', **TIGHT), m7, 8)
    assert rel_err(X, g['agg7_ens8_X']) < 1e-9
    f, X = P.absorption_spectra(m7, 2 * 127.5, correlation_decay_time=300.0,
                                polarization=[0.3, 1.1], integrator='zvode', **TIGHT)
    assert rel_err(X, g['agg7_decay_X']) < 1e-9


def test_third_order_matches_reference():
    g = golden('spectra_third_order')
    m7 = Lv.LiouvilleModel(_agg7(), unit_convert=S.CM_FS, secular=False)
    for geom in ('-++', '+-+', '++-'):
        (t1, t2, t3), X = P.third_order_response(m7, 2 * 15.5, population_times=g['t2'],
                                                 geometry=geom, integrator='zvode', **TIGHT)
        key = 'R3_' + geom.replace('-', 'm').replace('+', 'p')
        assert rel_err(X, g[key]) < 1e-8, geom
    (f1, t2, f3), X = P.two_dimensional_spectra(m7, 2 * 15.5, population_times=g['t2'],
                                                include_signal='GSB,ESE', integrator='rk4',
                                                dt=2.0 / 8)
    assert_allclose(f1, g['S2d_f1'])
    assert_allclose(f3, g['S2d_f3'])
    assert rel_err(X, g['S2d']) < 1e-6
