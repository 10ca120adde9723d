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


def test_impulsive_probe_closed_relations():
    """impulsive_probe (response.py:157-231): zero for the thermal state;
    linear in the state; a ground-population change c alone gives c times the
    'gg->eg->gg' linear response (GSB pathway); a single-exciton population
    enters through ESE and ESA only."""
    g = golden('redfield_dimer')
    ham = S.ElectronicHamiltonian(g['in_H1'], bath=S.DebyeBath(S.CM_K * 77, 35, 106),
                                  dipoles=g['in_dipoles'], energy_spread_extra=float(g['extra']))
    m = Lv.LiouvilleModel(ham, hilbert_subspace='gef', unit_convert=S.CM_FS)
    sub = 'gg,ge,eg,ee'
    rho0 = m.thermal_state(sub)
    f, X0 = P.impulsive_probe(m, rho0, 64.0)
    assert np.abs(X0).max() == 0
    gg = m.liouville_subspace_index('gg')
    full = m.liouville_subspace_index(sub)
    d_g = np.zeros(len(full), complex)
    d_g[np.flatnonzero(np.isin(full, gg))] = 0.3
    _, Xg = P.impulsive_probe(m, rho0 + d_g, 64.0)
    t, S_gsb = P.linear_response(m, 'gg->eg->gg', 64.0, 'xx', np.array([0.3]))
    _, ref = P.fourier_transform(t, S_gsb, rw_freq=m.rw_freq, unit_convert=m.unit_convert)
    assert rel_err(Xg, ref) < 1e-14
    rs = np.random.RandomState(4)
    a = rs.randn(len(full)) + 1j * rs.randn(len(full))
    b = rs.randn(len(full)) + 1j * rs.randn(len(full))
    _, Xa = P.impulsive_probe(m, rho0 + a, 64.0)
    _, Xb = P.impulsive_probe(m, rho0 + b, 64.0)
    _, Xab = P.impulsive_probe(m, rho0 + 2 * a - b, 64.0)
    assert rel_err(Xab, 2 * Xa - Xb) < 1e-12
    _, Xs = P.impulsive_probe(m, np.stack([rho0 + a, rho0 + b]), 64.0)
    assert rel_err(Xs[0], Xa) < 1e-13 and rel_err(Xs[1], Xb) < 1e-13
    ee = m.liouville_subspace_index('ee')
    d_e = np.zeros(len(full), complex)
    d_e[np.flatnonzero(np.isin(full, ee))[0]] = 0.5
    _, Xe = P.impulsive_probe(m, rho0 + d_e, 64.0)
    _, Xe_gsb = P.impulsive_probe(m, rho0 + d_e, 64.0, include_signal='GSB')
    assert np.abs(Xe_gsb).max() == 0 and np.abs(Xe).max() > 0


def test_pulses_known_answers():
    # qspectra tests/test_pulse.py:15-38
    pump = S.CustomPulse(0, 1, lambda x, r: x + r)
    assert (pump.t_init, pump.t_final, pump(1, 2)) == (0, 1, 3)
    pump = S.GaussianPulse(carrier_freq=10, fwhm=1, t_peak=5, scale=100, freq_convert=1,
                           t_limits_multiple=2)
    assert_allclose(pump(np.array([5, 5.5]), 10), [100, 50])
    assert_allclose(pump(np.array([5, 5.5]), 5), [100, 50 * np.exp(2.5j)])
    assert pump.t_init == 5 - 2 * S.GAUSSIAN_SD_FWHM and pump.t_final == 5 + 2 * S.GAUSSIAN_SD_FWHM
    pump = S.GaussianPulse(carrier_freq=10, fwhm=2, t_peak=0, scale=100, freq_convert=2)
    assert_allclose(pump(np.array([0, 1]), 10.0), [100, 50])
    assert_allclose(pump(1, 0.0), 50 * np.exp(20.0j))


def _two_level():
    """Monomer at the rotating-wave frequency (H = 0 in the rotating frame),
    unit dipole along x, no bath: the fields alone drive it."""
    ham = S.ElectronicHamiltonian(np.array([[12500.]]), dipoles=[[1., 0, 0]])
    return Lv.LiouvilleModel(ham, hilbert_subspace='ge', kind='unitary')


def test_driven_rk4_rabi_closed_form():
    """A resonant constant field E acting through '-' and '+' is the
    Hamiltonian |E| (e^{i phi}|g><e| + h.c.): from the ground state the
    excited population is sin^2(|E| t) whatever the phase phi, and the
    coherence magnitude |sin cos| (Rabi oscillation)."""
    m = _two_level()
    sub = 'gg,ge,eg,ee'
    T = 6.0
    for E in (0.7, 0.7 * np.exp(1.1j)):
        pulse = S.CustomPulse(0.0, T, lambda t, rw, E=E: E * np.ones_like(np.asarray(t, float)))
        t, y = P.simulate_with_fields(m, [pulse, pulse], '-+', 'xx', times=np.linspace(-T, 0, 13),
                                      liouville_subspace=sub, dt=T / 4000)
        assert_allclose(t, np.linspace(0, T, 13))
        ee = np.flatnonzero(np.isin(m.liouville_subspace_index(sub),
                                    m.liouville_subspace_index('ee')))[0]
        eg = np.flatnonzero(np.isin(m.liouville_subspace_index(sub),
                                    m.liouville_subspace_index('eg')))[0]
        assert np.max(np.abs(y[:, ee] - np.sin(abs(E) * t) ** 2)) < 1e-11
        assert np.max(np.abs(np.abs(y[:, eg]) - np.abs(np.sin(abs(E) * t) * np.cos(abs(E) * t)))) < 1e-11
        # a field through '-' only has no Hermitian partner: the ground state
        # is its fixed point ([|g><e|, |g><g|] acts on ge/eg only via '+')
        t, y1 = P.simulate_with_fields(m, [pulse], '+', 'x', times=np.linspace(-T, 0, 5),
                                       liouville_subspace=sub, dt=T / 400)
        assert np.abs(y1[:, ee]).max() < 1e-14


def test_driven_rk4_fourth_order():
    """Fixed-step RK4 for the time-dependent equation converges at fourth
    order to the tight zvode solution (a dropped or misplaced stage time
    degrades the order)."""
    md = Lv.LiouvilleModel(_dimer(), hilbert_subspace='gef', unit_convert=S.CM_FS)
    pump = S.GaussianPulse(12750., 20., scale=0.02, freq_convert=S.CM_FS)
    times = np.arange(0., 20., 4.)
    _, ref = P.simulate_pump(md, pump, 'x', times=times, integrator='zvode', **TIGHT)
    errs = [rel_err(P.simulate_pump(md, pump, 'x', times=times, dt=dt)[1], ref)
            for dt in (1.0, 0.5)]
    assert 12 < errs[0] / errs[1] < 20


def test_fields_match_reference():
    """simulate_with_fields / simulate_pump / impulsive_probe against the
    reference's own zvode outputs (tests/golden/fields.npz)."""
    g = golden('fields')
    md = Lv.LiouvilleModel(_dimer(), hilbert_subspace='gef', unit_convert=S.CM_FS)
    pump = S.GaussianPulse(12750., 20., t_peak=0., scale=0.02, freq_convert=S.CM_FS)
    t, y = P.simulate_pump(md, pump, 'x', time_extra=40., integrator='zvode', **TIGHT)
    assert_allclose(t, g['pump_t'])
    assert rel_err(y, g['pump_y']) < 1e-9
    t, y = P.simulate_pump(md, pump, 'x', time_extra=40., dt=md.time_step / 4)
    assert rel_err(y, g['pump_y']) < 1e-8
    t, y = P.simulate_pump(md, pump, [1., 0.5, 0.], times=np.arange(0., 60., 3.),
                           exact_isotropic_average=True, integrator='zvode', **TIGHT)
    assert_allclose(t, g['iso_t'])
    assert rel_err(y, g['iso_y']) < 1e-9
    p1 = S.GaussianPulse(12700., 15., t_peak=-10., scale=0.015, freq_convert=S.CM_FS)
    p2 = S.GaussianPulse(12900., 25., t_peak=5., scale=0.01 * np.exp(0.6j),
                         freq_convert=S.CM_FS, t_limits_multiple=2.5)
    t, y = P.simulate_with_fields(md, [p1, p2], '-+', 'xy', time_extra=20., integrator='zvode',
                                  **TIGHT)
    assert_allclose(t, g['two_t'])
    assert rel_err(y, g['two_y']) < 1e-9
    t, y = P.simulate_with_fields(md, [p1, p2, p1], '-+-', ['x', 'y', 0.7], times=[0., 5., 12.5],
                                  integrator='zvode', **TIGHT)
    assert_allclose(t, g['three_t'])
    assert rel_err(y, g['three_y']) < 1e-9
    mu = Lv.LiouvilleModel(_dimer(), hilbert_subspace='gef', unit_convert=S.CM_FS, kind='unitary')
    t, y = P.simulate_pump(mu, pump, 'y', time_extra=30., integrator='zvode', **TIGHT)
    assert rel_err(y, g['unitary_y']) < 1e-9
    m7 = Lv.LiouvilleModel(_agg7(disorder=100 / S.GAUSSIAN_SD_FWHM), unit_convert=S.CM_FS,
                           secular=False)
    pump7 = S.GaussianPulse(12450., 20., t_peak=0., scale=0.01, freq_convert=S.CM_FS)
    t, y = P.simulate_pump(m7, pump7, 'x', time_extra=20., integrator='zvode', **TIGHT)
    assert_allclose(t, g['agg7_t'])
    assert rel_err(y, g['agg7_y']) < 1e-9
    t, y = P.ensemble_average(
        lambda mm: P.simulate_pump(mm, pump7, 'z', times=np.array([0., 10., 20.]),
                                   integrator='zvode', **TIGHT), m7, 2)
    assert rel_err(y, g['agg7_ens2_y']) < 1e-9
    f, X = P.impulsive_probe(md, g['probe_states'], 300.5, integrator='zvode', **TIGHT)
    assert_allclose(f, g['probe_f'])
    assert rel_err(X, g['probe_X']) < 1e-9
    f, X = P.impulsive_probe(md, g['probe_states'][0], 300.5, polarization='xy',
                             include_signal='ESE,ESA', integrator='zvode', **TIGHT)
    assert rel_err(X, g['probe1_X']) < 1e-9


def test_heom_spectra_match_reference():
    """Linear and third-order responses of HEOM models (the hierarchy
    carried through every interval, dipoles per ADO, bra on ADO 0) against
    the reference's tight zvode outputs (tests/golden/heom_spectra.npz)."""
    g = golden('heom_spectra')
    for name, kw in [('lt', dict(level_cutoff=3, K=1)),
                     ('nolt', dict(level_cutoff=3, K=1, low_temp_corr=False))]:
        hm = Lv.HEOMOracleModel(_dimer(), hilbert_subspace='gef', unit_convert=S.CM_FS, **kw)
        t, Sg = P.linear_response(hm, 'gg->eg->gg', 150.5, integrator='zvode', **TIGHT)
        assert_allclose(t, g[name + '_lin_t'])
        assert rel_err(Sg, g[name + '_lin_S']) < 1e-9
        f, X = P.absorption_spectra(hm, 150.5, integrator='zvode', **TIGHT)
        assert rel_err(X, g[name + '_abs_X']) < 1e-9
        for geom in ('-++', '++-'):
            (t1, t2, t3), R = P.third_order_response(hm, 8.5, population_times=[0., 3., 7.],
                                                     geometry=geom, integrator='zvode', **TIGHT)
            key = '%s_R3_%s' % (name, geom.replace('-', 'm').replace('+', 'p'))
            assert rel_err(R, g[key]) < 1e-9, key
            (t1, t2, t3), R = P.third_order_response(hm, 8.5, population_times=[0., 3., 7.],
                                                     geometry=geom, dt=g['time_step'] / 8)
            assert rel_err(R, g[key]) < 1e-7, key


def test_heom_fields_match_reference():
    """Field-driven HEOM trajectories (the dipole commutators applied to
    every ADO) against the reference's tight zvode outputs
    (tests/golden/heom_fields.npz), full hierarchy states."""
    g = golden('heom_fields')
    hm = Lv.HEOMOracleModel(_dimer(), hilbert_subspace='gef', unit_convert=S.CM_FS,
                            level_cutoff=3, K=1)
    assert hm.ado_count == int(g['ado_count'])
    pump = S.GaussianPulse(12750., 20., t_peak=0., scale=0.02, freq_convert=S.CM_FS)
    t, y = P.simulate_pump(hm, pump, 'x', time_extra=30., integrator='zvode', **TIGHT)
    assert_allclose(t, g['pump_t'])
    assert rel_err(y, g['pump_y']) < 1e-9
    t, y = P.simulate_pump(hm, pump, 'x', time_extra=30., dt=hm.time_step / 8)
    assert rel_err(y, g['pump_y']) < 1e-7
    p1 = S.GaussianPulse(12700., 15., t_peak=-10., scale=0.015, freq_convert=S.CM_FS)
    p2 = S.GaussianPulse(12900., 25., t_peak=5., scale=0.01 * np.exp(0.6j),
                         freq_convert=S.CM_FS, t_limits_multiple=2.5)
    t, y = P.simulate_with_fields(hm, [p1, p2], '-+', 'xy', times=[0., 4., 10.],
                                  integrator='zvode', **TIGHT)
    assert_allclose(t, g['two_t'])
    assert rel_err(y, g['two_y']) < 1e-9


def _pm_arrays(g, n_sites):
    """The 16 pseudomodes of the reference's ZOFE examples, per site."""
    on = np.ones(n_sites, complex)
    return [np.outer(g[k], on) for k in ('pm_Omega', 'pm_gamma', 'pm_huang')]


def test_zofe_matches_reference():
    """ZOFE dynamics (three Hilbert spaces, the ham_hermit / rho_hermit
    forms), absorption (Schroedinger picture with the expectation value),
    pump and third-order response against the reference's tight zvode
    outputs (tests/golden/zofe.npz), by zvode and by the fixed-step RK4."""
    from oracle import zofe as Z
    g = golden('zofe')
    H1, d2 = g['in_H1'], g['in_dipoles']

    def ham(extra, H=H1, d=d2, n=2):
        return S.ElectronicHamiltonian(H, bath=Z.PseudomodeBath(16, *_pm_arrays(g, n)),
                                       dipoles=d, energy_spread_extra=extra)
    extra = float(g['extra_dimer'])
    m = Z.ZOFEOracleModel(ham(extra), hilbert_subspace='e', unit_convert=S.CM_FS)
    assert abs(m.time_step - float(g['time_step'])) < 1e-12
    t, rho = P.simulate_dynamics(m, [1, 0], 100.5, integrator='zvode', **TIGHT)
    assert_allclose(t, g['dyn_e_t'])
    assert rel_err(rho, g['dyn_e_rho']) < 1e-9
    t, rho = P.simulate_dynamics(m, [1, 0], 100.5, dt=0.125)
    assert rel_err(rho, g['dyn_e_rho']) < 1e-8
    for hh, rh in [(True, False), (False, True), (True, True)]:
        mf = Z.ZOFEOracleModel(ham(extra), hilbert_subspace='e', unit_convert=S.CM_FS,
                               ham_hermit=hh, rho_hermit=rh)
        t, rho = P.simulate_dynamics(mf, [1, 0], 50.5, integrator='zvode', **TIGHT)
        assert rel_err(rho, g['dyn_e_flags%d%d_rho' % (hh, rh)]) < 1e-9
    mg = Z.ZOFEOracleModel(ham(extra), hilbert_subspace='gef', unit_convert=S.CM_FS)
    t, rho = P.simulate_dynamics(mg, np.array([0.6, 0.8j, 0., 0.]), 60.5, integrator='zvode',
                                 **TIGHT)
    assert rel_err(rho, g['dyn_gef_rho']) < 1e-9
    mge = Z.ZOFEOracleModel(ham(extra), hilbert_subspace='ge', unit_convert=S.CM_FS)
    f, X = P.absorption_spectra(mge, 300.5, integrator='zvode', **TIGHT)
    assert_allclose(f, g['abs_f'])
    assert rel_err(X, g['abs_X']) < 1e-9
    pump = S.GaussianPulse(12800., 30., scale=2e-3, freq_convert=S.CM_FS)
    t, y = P.simulate_pump(mge, pump, 'x', time_extra=40., integrator='zvode', **TIGHT)
    assert_allclose(t, g['pump_t'])
    assert rel_err(y, g['pump_y']) < 1e-9
    (t1, t2, t3), R = P.third_order_response(mg, 6.5, population_times=[0., 4.],
                                             integrator='zvode', **TIGHT)
    assert rel_err(R, g['R3']) < 1e-9
    (t1, t2, t3), R = P.third_order_response(mg, 6.5, population_times=[0., 4.], dt=0.125)
    assert rel_err(R, g['R3']) < 1e-7
    m7 = Z.ZOFEOracleModel(ham(float(g['extra_agg7']), golden('redfield_7site')['in_H1'],
                               golden('redfield_7site')['in_dipoles'], 7),
                           hilbert_subspace='e', unit_convert=S.CM_FS)
    t, rho = P.simulate_dynamics(m7, np.eye(7)[0], 41., integrator='zvode', **TIGHT)
    assert rel_err(rho, g['agg7_rho']) < 1e-9
