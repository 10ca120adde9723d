"""Writes the golden fixtures under tests/golden/ by running the reference.

Run only in the build container, where the reference sources are present:

    python tests/golden/make_golden.py /root/reference            # every file
    python tests/golden/make_golden.py /root/reference fields     # fields.npz only
    python tests/golden/make_golden.py /root/reference heom_spectra
    python tests/golden/make_golden.py /root/reference heom_fields
    python tests/golden/make_golden.py /root/reference zofe
    python tests/golden/make_golden.py /root/reference vibronic

It imports qspectra from the given path and stores plain arrays (inputs and
outputs) as .npz files.  Nothing here is imported by the tests or by the
product; the fixtures are data.  Only outputs that no closed form or
known-answer test pins are written:

* Redfield Liouvillian blocks (non-secular, complex C(w), multi-site),
* HEOM sparse Liouvillians (ADO order, low-temperature correction),
* zvode-integrated trajectories and end-to-end spectra (absorption,
  third-order response, 2-D spectra, isotropic and ensemble averages),
* linear and third-order responses of HEOM models (dimer, 'gef'),
* field-driven HEOM trajectories (pump, two pulses; full hierarchy states),
* ZOFE (pseudomode) dynamics, absorption, pump and third-order outputs,
* vibronic Hamiltonians: Redfield and unitary Liouvillian blocks,
  dynamics, absorption and third-order GSB responses,
* field-driven trajectories (simulate_with_fields, simulate_pump: Gaussian
  pulses, complex amplitudes, three field terms, unitary / 7-site /
  ensemble models) and the impulsive probe of a pumped state,
* disorder samples and the complex bath correlation function.
"""
import os
import sys
import warnings

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TIGHT = dict(rtol=1e-12, atol=1e-14, nsteps=10 ** 7)


def dimer_inputs():
    th = np.deg2rad(30.0)
    return dict(H1=np.array([[12881., 120.], [120., 12719.]]),
                dipoles=np.array([[1., 0, 0], [np.cos(th), np.sin(th), 0]]),
                T=0.69503476 * 77, lam=35., gamma=106.)


def aggregate7_inputs(seed=1):
    rs = np.random.RandomState(seed)
    E = rs.uniform(12200, 12700, 7)
    J = rs.normal(0, 60, (7, 7))
    H1 = np.triu(J, 1)
    H1 = H1 + H1.T + np.diag(E)
    d = np.array([[3.019, 3.442, 0.797, 3.213, 2.969, 0.547, 1.983],
                  [2.284, -2.023, -3.871, 2.145, -2.642, 3.562, 2.837],
                  [1.506, 0.431, 0.853, 1.112, -0.661, -1.851, 2.015]]).T
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    return dict(H1=H1, dipoles=d, T=0.69503476 * 77, lam=35., gamma=106.)


def extra_for_dt(qs, H1, dt):
    """energy_spread_extra giving the model time step ``dt`` (fs)."""
    ham = qs.ElectronicHamiltonian(H1).in_rotating_frame()
    E = ham.E('gef')
    return 1.0 / (2 * dt * qs.CM_FS) - max(E.max(), -E.min())


def write_fields(qs, ham, ham7):
    """fields.npz: simulate_with_fields / simulate_pump / impulsive_probe."""
    out = {}
    md = qs.RedfieldModel(ham, hilbert_subspace='gef', unit_convert=qs.CM_FS)
    pump = qs.GaussianPulse(12750., 20., t_peak=0., scale=0.02, freq_convert=qs.CM_FS)
    t, y = qs.simulate_pump(md, pump, 'x', time_extra=40., **TIGHT)
    out['pump_t'], out['pump_y'] = t, y
    t, y = qs.simulate_pump(md, pump, [1., 0.5, 0.], times=np.arange(0., 60., 3.),
                            exact_isotropic_average=True, **TIGHT)
    out['iso_t'], out['iso_y'] = t, y
    p1 = qs.GaussianPulse(12700., 15., t_peak=-10., scale=0.015, freq_convert=qs.CM_FS)
    p2 = qs.GaussianPulse(12900., 25., t_peak=5., scale=0.01 * np.exp(0.6j),
                          freq_convert=qs.CM_FS, t_limits_multiple=2.5)
    t, y = qs.simulate_with_fields(md, [p1, p2], '-+', 'xy', time_extra=20., **TIGHT)
    out['two_t'], out['two_y'] = t, y
    t, y = qs.simulate_with_fields(md, [p1, p2, p1], '-+-', ['x', 'y', 0.7], times=[0., 5., 12.5],
                                   **TIGHT)
    out['three_t'], out['three_y'] = t, y
    mu = qs.UnitaryModel(ham, hilbert_subspace='gef', unit_convert=qs.CM_FS)
    t, y = qs.simulate_pump(mu, pump, 'y', time_extra=30., **TIGHT)
    out['unitary_t'], out['unitary_y'] = t, y
    m7 = qs.RedfieldModel(ham7, hilbert_subspace='gef', unit_convert=qs.CM_FS, secular=False)
    pump7 = qs.GaussianPulse(12450., 20., t_peak=0., scale=0.01, freq_convert=qs.CM_FS)
    t, y = qs.simulate_pump(m7, pump7, 'x', time_extra=20., **TIGHT)
    out['agg7_t'], out['agg7_y'] = t, y
    t, y = qs.simulate_pump(m7, pump7, 'z', times=np.array([0., 10., 20.]), ensemble_size=2,
                            **TIGHT)
    out['agg7_ens2_t'], out['agg7_ens2_y'] = t, y
    states = out['pump_y'][[-1, -10, -20]]
    f, X = qs.impulsive_probe(md, states, 300.5, **TIGHT)
    out['probe_states'], out['probe_f'], out['probe_X'] = states, f, X
    f, X = qs.impulsive_probe(md, states[0], 300.5, polarization='xy', include_signal='ESE,ESA',
                              **TIGHT)
    out['probe1_f'], out['probe1_X'] = f, X
    np.savez_compressed(os.path.join(HERE, 'fields.npz'), **out)


def write_heom_spectra(qs, HEOMModel, dim, extra):
    """heom_spectra.npz: linear and third-order responses of HEOM models
    (the hierarchy states carried through every interval)."""
    out = {}
    ham = qs.ElectronicHamiltonian(dim['H1'], bath=qs.DebyeBath(dim['T'], dim['lam'],
                                                              dim['gamma']),
                                   dipoles=dim['dipoles'], energy_spread_extra=extra)
    for name, kw in [('lt', dict(level_cutoff=3, K=1)),
                     ('nolt', dict(level_cutoff=3, K=1, low_temp_corr=False))]:
        hm = HEOMModel(ham, hilbert_subspace='gef', unit_convert=qs.CM_FS, **kw)
        t, S = qs.linear_response(hm, 'gg->eg->gg', 150.5, **TIGHT)
        out[name + '_lin_t'], out[name + '_lin_S'] = t, S
        f, X = qs.absorption_spectra(hm, 150.5, **TIGHT)
        out[name + '_abs_f'], out[name + '_abs_X'] = f, X
        for geom in ('-++', '++-'):
            (t1, t2, t3), R = qs.third_order_response(hm, 8.5, population_times=[0., 3., 7.],
                                                      geometry=geom, **TIGHT)
            g = geom.replace('-', 'm').replace('+', 'p')
            out['%s_R3_%s' % (name, g)] = R
            out['t1'], out['t2'], out['t3'] = t1, t2, t3
    out['time_step'] = hm.time_step
    np.savez_compressed(os.path.join(HERE, 'heom_spectra.npz'), **out)


def write_heom_fields(qs, HEOMModel, dim, extra):
    """heom_fields.npz: field-driven HEOM trajectories (simulate_pump and
    simulate_with_fields of a HEOMModel: the dipole commutators act on every
    ADO, eom.py:79-103 with heom.py:22-58); full hierarchy states."""
    out = {}
    ham = qs.ElectronicHamiltonian(dim['H1'], bath=qs.DebyeBath(dim['T'], dim['lam'],
                                                              dim['gamma']),
                                   dipoles=dim['dipoles'], energy_spread_extra=extra)
    hm = HEOMModel(ham, hilbert_subspace='gef', unit_convert=qs.CM_FS, level_cutoff=3, K=1)
    pump = qs.GaussianPulse(12750., 20., t_peak=0., scale=0.02, freq_convert=qs.CM_FS)
    t, y = qs.simulate_pump(hm, pump, 'x', time_extra=30., **TIGHT)
    out['pump_t'], out['pump_y'] = t, y
    p1 = qs.GaussianPulse(12700., 15., t_peak=-10., scale=0.015, freq_convert=qs.CM_FS)
    p2 = qs.GaussianPulse(12900., 25., t_peak=5., scale=0.01 * np.exp(0.6j),
                          freq_convert=qs.CM_FS, t_limits_multiple=2.5)
    t, y = qs.simulate_with_fields(hm, [p1, p2], '-+', 'xy', times=[0., 4., 10.], **TIGHT)
    out['two_t'], out['two_y'] = t, y
    out['ado_count'] = hm.ado_count
    out['time_step'] = hm.time_step
    np.savez_compressed(os.path.join(HERE, 'heom_fields.npz'), **out)


PM_OMEGA = [-500., -200., -90., 1., 21., 60., 80., 130., 200., 300., 400., 500., 600., 800.,
            1100., 1500.]
PM_HUANG = [-2.5133e-03, -7.5398e-03, -2.5133e-02, 5.0265e+01, 2.2619e+00, 4.5239e-02,
            2.7646e-01, 9.2991e-03, 2.2619e-02, 1.5080e-02, 3.0159e-03, 3.5186e-03, 2.8274e-04,
            1.7593e-03, 4.3982e-04, 4.3982e-04]
PM_GAMMA = [500., 100., 50., 50., 50., 50., 80., 40., 80., 150., 200., 200., 80., 250., 200.,
            300.]


def pseudomode_arrays(n_sites):
    """The 16-pseudomode fit of the reference's ZOFE examples, per site."""
    on = np.ones(n_sites, complex)
    return [np.array([v[pm] * on for pm in range(len(v))])
            for v in (PM_OMEGA, PM_GAMMA, PM_HUANG)]


def write_zofe(qs, agg):
    """zofe.npz: ZOFE dynamics, absorption, pump and third-order outputs."""
    from qspectra.dynamics.zofe import ZOFEModel
    out = {'pm_Omega': np.array(PM_OMEGA), 'pm_gamma': np.array(PM_GAMMA),
           'pm_huang': np.array(PM_HUANG)}
    H1 = np.array([[12881., 120.], [120., 12719.]])
    d2 = [[1., 0., 0.], [2. * np.cos(.3), 2. * np.sin(.3), 0.]]
    out['in_H1'], out['in_dipoles'] = H1, np.array(d2)
    out['extra_dimer'] = extra_for_dt(qs, H1, 1.0)
    out['extra_agg7'] = extra_for_dt(qs, agg['H1'], 2.0)
    Om, ga, hu = pseudomode_arrays(2)
    ham = qs.ElectronicHamiltonian(H1, bath=qs.PseudomodeBath(16, Om, ga, hu), dipoles=d2,
                                   energy_spread_extra=extra_for_dt(qs, H1, 1.0))
    m = ZOFEModel(ham, hilbert_subspace='e', unit_convert=qs.CM_FS)
    t, rho = qs.simulate_dynamics(m, [1, 0], 100.5, **TIGHT)
    out['dyn_e_t'], out['dyn_e_rho'] = t, rho
    out['time_step'] = m.time_step
    for flags in [(True, False), (False, True), (True, True)]:
        mf = ZOFEModel(ham, hilbert_subspace='e', unit_convert=qs.CM_FS, ham_hermit=flags[0],
                       rho_hermit=flags[1])
        t, rho = qs.simulate_dynamics(mf, [1, 0], 50.5, **TIGHT)
        out['dyn_e_flags%d%d_rho' % flags] = rho
    mg = ZOFEModel(ham, hilbert_subspace='gef', unit_convert=qs.CM_FS)
    psi = np.array([0.6, 0.8j, 0., 0.])
    t, rho = qs.simulate_dynamics(mg, psi, 60.5, **TIGHT)
    out['dyn_gef_t'], out['dyn_gef_rho'] = t, rho
    mge = ZOFEModel(ham, hilbert_subspace='ge', unit_convert=qs.CM_FS)
    f, X = qs.absorption_spectra(mge, 300.5, **TIGHT)
    out['abs_f'], out['abs_X'] = f, X
    pump = qs.GaussianPulse(12800., 30., scale=2e-3, freq_convert=qs.CM_FS)
    t, y = qs.simulate_pump(mge, pump, 'x', time_extra=40., **TIGHT)
    out['pump_t'], out['pump_y'] = t, y
    (t1, t2, t3), R = qs.third_order_response(mg, 6.5, population_times=[0., 4.], **TIGHT)
    out['R3_t'] = np.array([len(t1), len(t2), len(t3)])
    out['R3'], out['t1'], out['t2'], out['t3'] = R, t1, t2, t3
    Om7, ga7, hu7 = pseudomode_arrays(7)
    ham7 = qs.ElectronicHamiltonian(agg['H1'], bath=qs.PseudomodeBath(16, Om7, ga7, hu7),
                                    dipoles=agg['dipoles'],
                                    energy_spread_extra=extra_for_dt(qs, agg['H1'], 2.0))
    m7 = ZOFEModel(ham7, hilbert_subspace='e', unit_convert=qs.CM_FS)
    t, rho = qs.simulate_dynamics(m7, [1, 0, 0, 0, 0, 0, 0], 41., **TIGHT)
    out['agg7_t'], out['agg7_rho'] = t, rho
    np.savez_compressed(os.path.join(HERE, 'zofe.npz'), **out)


JONAS = dict(H_el=np.array([[11500., 66.], [66., 11650.]]), levels=np.array([2, 2]),
             vib=np.array([200., 200.]), couplings=-32. * np.eye(2),
             dipoles=np.array([[1., 0., 0.], [0., 1., 0.]]), lam=1.5 * 35., gamma=106.)


def frac_ok(duration, dt):
    """time grids end away from a multiple of the step, so arange lengths
    do not depend on the last bits of the step"""
    f = (duration / dt) % 1.0
    assert 0.02 < f < 0.98, (duration, dt, f)
    return duration


def pow2_grid(k, dt):
    """a duration whose grid arange(0, T, dt) ends at index k = 2^m: then
    t[-1] / dt == k exactly, and the zero-padding count of
    fourier_transform (int(t[-1] / dt), utils.py:128-151) does not depend on
    the last bits of dt"""
    assert k & (k - 1) == 0
    return frac_ok(float(np.ceil((k + 0.3) * dt)), dt)


def write_vibronic(qs):
    """vibronic.npz: vibronic Hamiltonians (hamiltonian.py:633-807) of the
    reference's two vibronic examples: the Jonas dimer with Redfield
    dynamics (real C(w)), its monomer (complex C(w)) and the unitary
    five-level vibronic monomer."""
    out = {'in_' + k: np.asarray(v) for k, v in JONAS.items()}
    out['in_T'] = qs.CM_K * 77
    bath = qs.DebyeBath(qs.CM_K * 77, JONAS['lam'], JONAS['gamma'])
    el = qs.ElectronicHamiltonian(JONAS['H_el'], dipoles=JONAS['dipoles'], bath=bath)
    ham = qs.VibronicHamiltonian(el, JONAS['levels'], JONAS['vib'], JONAS['couplings'])
    dimer = qs.RedfieldModel(ham, hilbert_subspace='gef', unit_convert=qs.CM_FS,
                             discard_imag_corr=True)
    dt = dimer.time_step
    out['jonas_time_step'], out['jonas_rw'] = dt, dimer.rw_freq
    out['jonas_H_gef'] = ham.H('gef')
    out['jonas_E_gef'] = dimer.hamiltonian.E('gef')
    out['jonas_thermal_gg'] = dimer.hamiltonian.thermal_state('g')
    for s in ['gg', 'ee', 'eg', 'fe', 'ff']:
        idx = dimer.liouville_subspace_index(s)
        out['jonas_L_' + s] = dimer.evolution_super_operator[np.ix_(idx, idx)]
    t, rho = qs.simulate_dynamics(dimer, qs.unit_vec(0, 8), frac_ok(1000., dt), **TIGHT)
    out['jonas_dyn_t'], out['jonas_dyn_rho'] = t, rho
    T = pow2_grid(2048, dt)
    f, X = qs.absorption_spectra(dimer, T, **TIGHT)
    out['jonas_abs_T'] = T
    out['jonas_abs_f'], out['jonas_abs_X'] = f, X
    (t1, t2, t3), R = qs.third_order_response(dimer, coherence_time_max=frac_ok(200.5, dt),
                                              population_times=[0., 250., 500.],
                                              geometry='-++', polarization='xxxx',
                                              include_signal='GSB',
                                              exact_isotropic_average=True, **TIGHT)
    out['jonas_R3_t1'], out['jonas_R3_t2'], out['jonas_R3_t3'] = t1, t2, t3
    out['jonas_R3'] = R
    mono = qs.VibronicHamiltonian(
        qs.ElectronicHamiltonian(JONAS['H_el'][:1, :1], dipoles=JONAS['dipoles'][:1],
                                 bath=bath),
        JONAS['levels'][:1], JONAS['vib'][:1], JONAS['couplings'][:1, :1])
    mm = qs.RedfieldModel(mono, hilbert_subspace='ge', unit_convert=qs.CM_FS)
    T = pow2_grid(1024, mm.time_step)
    f, X = qs.absorption_spectra(mm, T, **TIGHT)
    out['jmono_time_step'], out['jmono_abs_f'], out['jmono_abs_X'] = mm.time_step, f, X
    out['jmono_abs_T'] = T
    uham = qs.VibronicHamiltonian(
        qs.ElectronicHamiltonian([[11500.]], dipoles=[[1., 0., 0.]], bath=None),
        n_vibrational_levels=[5], vib_energies=[200.], elec_vib_couplings=[[-100.]])
    um = qs.UnitaryModel(uham, hilbert_subspace='ge', unit_convert=qs.CM_FS)
    out['umono_time_step'], out['umono_H_ge'] = um.time_step, uham.H('ge')
    t, rho = qs.simulate_dynamics(um, qs.unit_vec(0, 5), frac_ok(1000., um.time_step), **TIGHT)
    out['umono_dyn_t'], out['umono_dyn_rho'] = t, rho
    T = pow2_grid(2048, um.time_step)
    f, X = qs.absorption_spectra(um, T, correlation_decay_time=1000., **TIGHT)
    out['umono_abs_f'], out['umono_abs_X'], out['umono_abs_T'] = f, X, T
    np.savez_compressed(os.path.join(HERE, 'vibronic.npz'), **out)


def write_full_2d(qs, ham_of, agg):
    """spectra_2d_full.npz: the 2-D workload of the 7-site aggregate on its
    full grid (t1 = t3 = 128 points at 2 fs, t2 in {0, 100, 200, 500} fs,
    GSB + ESE + ESA, '-++', non-secular Redfield g/e/f) at tight zvode
    tolerance: the time-domain response in complex128 and the 2-D spectrum
    in complex64 (its rounding, 6e-8 of each entry, is far below the 1e-6
    tolerance it is compared at).  The response at the reference's default
    tolerance is stored as its max-norm distance from the tight one."""
    extra7 = extra_for_dt(qs, agg['H1'], 2.0)
    m7 = qs.RedfieldModel(ham_of(agg, extra7), hilbert_subspace='gef', unit_convert=qs.CM_FS,
                          secular=False)
    kw = dict(population_times=[0., 100., 200., 500.], geometry='-++',
              include_signal='GSB,ESE,ESA')
    (t1, t2, t3), R = qs.third_order_response(m7, 2 * 127.5, **dict(kw, **TIGHT))
    (f1, _, f3), S = qs.two_dimensional_spectra(m7, 2 * 127.5, **dict(kw, **TIGHT))
    _, Rd = qs.third_order_response(m7, 2 * 127.5, **kw)
    np.savez_compressed(os.path.join(HERE, 'spectra_2d_full.npz'), t1=t1, t2=t2, t3=t3, R3=R,
                        f1=f1, f3=f3, S2d=S.astype(np.complex64), rw_freq=m7.rw_freq,
                        time_step=m7.time_step,
                        default_tol_rel_err=np.abs(Rd - R).max() / np.abs(R).max())


def write_manifest(ref_path):
    with open(os.path.join(HERE, 'MANIFEST.txt'), 'w') as fh:
        fh.write('written by tests/golden/make_golden.py from the reference at %s\n' % ref_path)
        fh.write('numpy %s scipy %s python %s\n' % (np.__version__, __import__('scipy').__version__,
                                                  sys.version.split()[0]))
        fh.write('tight zvode: %r\n' % (TIGHT,))
        for fn in sorted(os.listdir(HERE)):
            if fn.endswith('.npz'):
                fh.write('%s %d bytes\n' % (fn, os.path.getsize(os.path.join(HERE, fn))))


def main(ref_path, only=None):
    sys.path.insert(0, ref_path)
    import qspectra as qs
    from qspectra.dynamics.redfield import redfield_tensor
    from qspectra.dynamics.heom import HEOMModel
    from qspectra.simulate.utils import integrate
    warnings.simplefilter('ignore')

    def ham_of(inp, extra=None, disorder=None, seed=0):
        return qs.ElectronicHamiltonian(
            inp['H1'], bath=qs.DebyeBath(inp['T'], inp['lam'], inp['gamma']),
            dipoles=inp['dipoles'], disorder=disorder, random_seed=seed,
            energy_spread_extra=extra)

    if only == 'vibronic':
        write_vibronic(qs)
        write_manifest(ref_path)
        return
    if only == 'full_2d':
        write_full_2d(qs, ham_of, aggregate7_inputs())
        write_manifest(ref_path)
        return
    if only == 'zofe':
        write_zofe(qs, aggregate7_inputs())
        write_manifest(ref_path)
        return
    if only == 'heom_spectra':
        from qspectra.dynamics.heom import HEOMModel
        dim = dimer_inputs()
        write_heom_spectra(qs, HEOMModel, dim, extra_for_dt(qs, dim['H1'], 1.0))
        write_manifest(ref_path)
        return
    if only == 'heom_fields':
        from qspectra.dynamics.heom import HEOMModel
        dim = dimer_inputs()
        write_heom_fields(qs, HEOMModel, dim, extra_for_dt(qs, dim['H1'], 1.0))
        write_manifest(ref_path)
        return
    if only == 'fields':
        dim, agg = dimer_inputs(), aggregate7_inputs()
        ham = ham_of(dim, extra_for_dt(qs, dim['H1'], 1.0))
        ham7 = ham_of(agg, extra_for_dt(qs, agg['H1'], 2.0), disorder=100 / qs.GAUSSIAN_SD_FWHM)
        write_fields(qs, ham, ham7)
        write_manifest(ref_path)
        return

    # ---------------------------------------------------------------- bath
    b = qs.DebyeBath(0.69503476 * 77, 35., 106.)
    x = np.concatenate([np.linspace(-900, 900, 61), [0.0, 1e-3, -1e-3]])
    b300 = qs.DebyeBath(0.69503476 * 300, 35., 106.)
    np.savez_compressed(os.path.join(HERE, 'bath.npz'), x=x,
                        c77=np.array([b.corr_func_complex(v) for v in x]),
                        r77=np.array([b.corr_func_real(v) for v in x]),
                        c300=np.array([b300.corr_func_complex(v) for v in x]))

    # ------------------------------------------------- Redfield dimer blocks
    dim = dimer_inputs()
    extra = extra_for_dt(qs, dim['H1'], 1.0)
    ham = ham_of(dim, extra)
    out = dict(extra=extra, **{'in_' + k: np.asarray(v) for k, v in dim.items()})
    subspaces = ['gg', 'eg', 'ge', 'ee', 'fg', 'fe', 'gg,ge,eg,ee']
    for sec in (True, False):
        for dic in (True, False):
            m = qs.RedfieldModel(ham, hilbert_subspace='gef', unit_convert=qs.CM_FS,
                                 secular=sec, discard_imag_corr=dic)
            for s in subspaces:
                idx = m.liouville_subspace_index(s)
                key = 'L_%s_%d%d' % (s.replace(',', '_'), sec, dic)
                out[key] = m.evolution_super_operator[np.ix_(idx, idx)]
    m = qs.RedfieldModel(ham, hilbert_subspace='gef', unit_convert=qs.CM_FS,
                         secular=False, evolve_basis='eigen')
    out['Leig_ee'] = m.evolution_super_operator[np.ix_(*[m.liouville_subspace_index('ee')] * 2)]
    out['E_gef'], out['U_gef'] = m.hamiltonian.E('gef'), m.hamiltonian.U('gef')
    out['rw_freq'] = m.rw_freq
    out['time_step'] = m.time_step
    out['R_ge'] = redfield_tensor(m.hamiltonian, 'ge', secular=False)
    np.savez_compressed(os.path.join(HERE, 'redfield_dimer.npz'), **out)

    # ---------------------------------------------- Redfield 7-site blocks
    agg = aggregate7_inputs()
    extra7 = extra_for_dt(qs, agg['H1'], 2.0)
    ham7 = ham_of(agg, extra7, disorder=100 / qs.GAUSSIAN_SD_FWHM)
    out = dict(extra=extra7, **{'in_' + k: np.asarray(v) for k, v in agg.items()})
    m7 = qs.RedfieldModel(ham7, hilbert_subspace='gef', unit_convert=qs.CM_FS,
                          secular=False)
    for s in ['ge', 'eg', 'ee', 'fe', 'fg']:
        idx = m7.liouville_subspace_index(s)
        out['L_' + s] = m7.evolution_super_operator[np.ix_(idx, idx)]
    m7s = qs.RedfieldModel(ham7, hilbert_subspace='gef', unit_convert=qs.CM_FS,
                           secular=True)
    idx = m7s.liouville_subspace_index('ee')
    out['Lsec_ee'] = m7s.evolution_super_operator[np.ix_(idx, idx)]
    out['rw_freq'] = m7.rw_freq
    out['time_step'] = m7.time_step
    for n in range(4):
        out['sample_H1_%d' % n] = ham7.sample(n).H_1exc
    hs = ham7.in_rotating_frame().sample(2)
    out['sample2_rot_E_gef'] = hs.E('gef')
    out['sample2_time_step'] = hs.time_step
    np.savez_compressed(os.path.join(HERE, 'redfield_7site.npz'), **out)

    # ------------------------------------------------------ spectra (zvode)
    out = {}
    mono = qs.ElectronicHamiltonian(np.array([[12500.]]), bath=qs.DebyeBath(qs.CM_K * 77, 35, 106),
                                    dipoles=[[1., 0, 0]])
    for dic in (True, False):
        mm = qs.RedfieldModel(mono, discard_imag_corr=dic, unit_convert=qs.CM_FS)
        f, X = qs.absorption_spectra(mm, 2000, **TIGHT)
        out['mono_f_%d' % dic], out['mono_X_%d' % dic] = f, X
        out['mono_dt_%d' % dic] = mm.time_step
    md = qs.RedfieldModel(ham, hilbert_subspace='gef', unit_convert=qs.CM_FS)
    t, S = qs.linear_response(md, 'gg->eg->gg', 511.5, **TIGHT)
    out['dimer_t'], out['dimer_S'] = t, S
    f, X = qs.absorption_spectra(md, 511.5, **TIGHT)
    out['dimer_f'], out['dimer_X'] = f, X
    f, X = qs.absorption_spectra(md, 511.5)
    out['dimer_X_default_tol'] = X
    f, X = qs.absorption_spectra(md, 511.5, exact_isotropic_average=True, **TIGHT)
    out['dimer_X_iso'] = X
    f, X = qs.absorption_spectra(m7, 2 * 255.5, **TIGHT)
    out['agg7_f'], out['agg7_X'] = f, X
    f, X = qs.absorption_spectra(m7, 2 * 127.5, ensemble_size=8, **TIGHT)
    out['agg7_ens8_f'], out['agg7_ens8_X'] = f, X
    f, X = qs.absorption_spectra(m7, 2 * 127.5, correlation_decay_time=300.0,
                                 polarization=[0.3, 1.1], **TIGHT)
    out['agg7_decay_X'] = X
    np.savez_compressed(os.path.join(HERE, 'spectra_linear.npz'), **out)

    # --------------------------------------- third-order (reduced 7-site grid)
    out = {}
    t2s = np.array([0., 20., 40., 100.])
    for geom in ('-++', '+-+', '++-'):
        (t1, t2, t3), X = qs.third_order_response(m7, 2 * 15.5, population_times=t2s,
                                                  geometry=geom, **TIGHT)
        g = geom.replace('-', 'm').replace('+', 'p')
        out['R3_' + g] = X
        out['t1'], out['t2'], out['t3'] = t1, t2, t3
    (f1, t2, f3), X = qs.two_dimensional_spectra(m7, 2 * 15.5, population_times=t2s,
                                                 include_signal='GSB,ESE', **TIGHT)
    out['S2d_f1'], out['S2d_f3'], out['S2d'] = f1, f3, X
    (t1, t2, t3), X = qs.third_order_response(md, 8.5, population_times=[0., 3., 7.],
                                              exact_isotropic_average=True,
                                              polarization=[0, 0, qs.MAGIC_ANGLE, qs.MAGIC_ANGLE],
                                              **TIGHT)
    out['dimer_iso_R3'] = X
    np.savez_compressed(os.path.join(HERE, 'spectra_third_order.npz'), **out)

    # ------------------------------------------------------------------ HEOM
    out = {}
    dim300 = dict(dim, T=0.69503476 * 300)
    hamh = ham_of(dim300, extra)
    for name, kw in [('dimer', dict(level_cutoff=3, K=1)),
                     ('dimer_mod', dict(level_cutoff=3, K=1, modified_HEOM=True)),
                     ('dimer_nolt', dict(level_cutoff=3, K=2, low_temp_corr=False))]:
        hm = HEOMModel(hamh, hilbert_subspace='ge', unit_convert=qs.CM_FS, **kw)
        for s in ('ee', 'gg,ge,eg,ee', 'eg'):
            L = (qs.CM_FS * hm.HEOM_tensor(s)).tocsr()
            key = '%s_%s' % (name, s.replace(',', '_'))
            out[key + '_data'], out[key + '_indices'], out[key + '_indptr'] = \
                L.data, L.indices, L.indptr
            out[key + '_shape'] = np.array(L.shape)
        out[name + '_ados'] = np.array([a.flatten() for a in hm.ado_indices])
    agg300 = dict(agg, T=0.69503476 * 300)
    ham7h = ham_of(agg300, extra7)
    hm7 = HEOMModel(ham7h, hilbert_subspace='ge', unit_convert=qs.CM_FS, level_cutoff=3, K=1)
    L = (qs.CM_FS * hm7.HEOM_tensor('ee')).tocsr()
    out['agg7_ee_data'], out['agg7_ee_indices'], out['agg7_ee_indptr'] = L.data, L.indices, L.indptr
    out['agg7_ee_shape'] = np.array(L.shape)
    out['agg7_ados'] = np.array([a.flatten() for a in hm7.ado_indices])
    out['agg7_rw_freq'] = hm7.rw_freq
    hm = HEOMModel(hamh, hilbert_subspace='ge', unit_convert=qs.CM_FS, level_cutoff=3, K=1)
    rho0 = np.zeros((2, 2))
    rho0[0, 0] = 1
    y0 = hm.density_matrix_to_state_vector(rho0, 'ee')
    t = np.arange(0, 200, hm.time_step)
    y = integrate(hm.equation_of_motion('ee'), y0, t, **TIGHT)
    out['dyn_t'], out['dyn_y'] = t, y
    out['dyn_time_step'] = hm.time_step
    np.savez_compressed(os.path.join(HERE, 'heom.npz'), **out)

    # ------------------------------------------------ Redfield dynamics (zvode)
    out = {}
    t, rho = qs.simulate_dynamics(m7, [1, 0, 0, 0, 0, 0, 0], 200, **TIGHT)
    out['t'], out['rho'] = t, rho
    mu = qs.UnitaryModel(ham7, hilbert_subspace='gef', unit_convert=qs.CM_FS)
    t, rho = qs.simulate_dynamics(mu, [0, 1, 0, 0, 0, 0, 0], 100, **TIGHT)
    out['unitary_t'], out['unitary_rho'] = t, rho
    np.savez_compressed(os.path.join(HERE, 'dynamics.npz'), **out)

    write_fields(qs, ham, ham7)
    write_heom_spectra(qs, HEOMModel, dim, extra)
    write_heom_fields(qs, HEOMModel, dim, extra_for_dt(qs, dim['H1'], 1.0))
    write_zofe(qs, agg)
    write_vibronic(qs)
    write_manifest(ref_path)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference',
         sys.argv[2] if len(sys.argv) > 2 else None)
