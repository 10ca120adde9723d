"""Parity of the GPU Liouvillian builders and the drop-in model / simulate
API with the oracle and with the reference's golden outputs.

Tolerances: FP64 operators 1e-12 relative (max-norm); FP64 propagation
against the oracle's RK4 on the same schedule 1e-10; against the
reference's tight-tolerance zvode outputs, the RK4 truncation error at the
step used (stated per test); MIXED precision 1e-6 relative per spectrum.
"""
import numpy as np
import pytest
import torch

import qspectra_amd as qa
from qspectra_amd import _native
from oracle import system as S
from oracle import liouville as Lv
from oracle import propagate as P
from tests.conftest import golden, rel_err

pytestmark = pytest.mark.gpu


def _inputs(name):
    g = golden(name)
    return g, dict(H1=g['in_H1'], dipoles=g['in_dipoles'], extra=float(g['extra']))


def _product_ham(inp, T=S.CM_K * 77, **kw):
    return qa.ElectronicHamiltonian(inp['H1'], bath=qa.DebyeBath(T, 35, 106),
                                    dipoles=inp['dipoles'], energy_spread_extra=inp['extra'], **kw)


def _oracle_ham(inp, T=S.CM_K * 77, **kw):
    return S.ElectronicHamiltonian(inp['H1'], bath=S.DebyeBath(T, 35, 106),
                                   dipoles=inp['dipoles'], energy_spread_extra=inp['extra'], **kw)


def _canonical_signs(U):
    idx = np.argmax(np.abs(U), axis=0)
    return np.sign(U[idx, np.arange(U.shape[1])])


def test_block_eigh_matches_lapack(gpu):
    rs = np.random.RandomState(3)
    for n, X in [(7, 1), (7, 2), (24, 1), (2, 1), (3, 2), (1, 1)]:
        H = rs.randn(5, n, n) * 80
        H = H + H.transpose(0, 2, 1) + np.diag(rs.uniform(12200, 12700, n))
        rw = rs.uniform(12300, 12500, 5)
        E, U = _native.block_eigh(torch.as_tensor(H).cuda(), X, torch.as_tensor(rw).cuda())
        E, U = E.cpu().numpy(), U.cpu().numpy()
        for b in range(5):
            Hb = S.operator_extend(H[b], 'gef'[X])
            Er, Ur = np.linalg.eigh(Hb)
            assert np.max(np.abs(E[b] - (Er - X * rw[b]))) < 1e-10
            Ur = Ur * _canonical_signs(Ur)
            assert np.max(np.abs(U[b] - Ur)) < 1e-10, (n, X)


def test_redfield_dimer_blocks_match_reference(gpu):
    g, inp = _inputs('redfield_dimer')
    ham = _product_ham(inp)
    for sec in (True, False):
        for dic in (True, False):
            m = qa.RedfieldModel(ham, unit_convert=qa.CM_FS, secular=sec, discard_imag_corr=dic)
            for s in ['gg', 'eg', 'ge', 'ee', 'fg', 'fe', 'gg,ge,eg,ee']:
                ref = g['L_%s_%d%d' % (s.replace(',', '_'), sec, dic)]
                got = m.liouvillian(s)[0].cpu().numpy()
                assert rel_err(got, ref) < 1e-12, (s, sec, dic)
                got = m.liouvillian(s, heisenberg_picture=True)[0].cpu().numpy()
                assert rel_err(got, ref.T) < 1e-12, (s, sec, dic, 'heisenberg')


def test_redfield_eigen_basis_up_to_eigenvector_signs(gpu):
    g, inp = _inputs('redfield_dimer')
    m = qa.RedfieldModel(_product_ham(inp), unit_convert=qa.CM_FS, secular=False,
                         evolve_basis='eigen')
    sg = _canonical_signs(g['U_gef'])          # reference U -> canonical signs
    e = sg[1:3]
    sign = np.outer(e, e).reshape(-1, order='F')   # s_a s_b for (a, b) in 'ee'
    ref = g['Leig_ee'] * np.outer(sign, sign)
    assert rel_err(m.liouvillian('ee')[0].cpu().numpy(), ref) < 1e-12
    assert np.max(np.abs(m.hamiltonian.E('gef') - g['E_gef'])) < 1e-9


def test_redfield_7site_blocks_and_batch(gpu):
    g, inp = _inputs('redfield_7site')
    m = qa.RedfieldModel(_product_ham(inp, disorder=100 / qa.GAUSSIAN_SD_FWHM),
                         unit_convert=qa.CM_FS, secular=False)
    for s in ['ge', 'eg', 'ee', 'fe', 'fg']:
        assert rel_err(m.liouvillian(s)[0].cpu().numpy(), g['L_' + s]) < 1e-12, s
    ms = qa.RedfieldModel(_product_ham(inp), unit_convert=qa.CM_FS, secular=True)
    assert rel_err(ms.liouvillian('ee')[0].cpu().numpy(), g['Lsec_ee']) < 1e-12
    # a batch of disorder samples in one launch vs the oracle member by member
    hams = [m.hamiltonian.sample(n) for n in range(6)]
    mb = m.with_hamiltonians(hams)
    om = Lv.LiouvilleModel(_oracle_ham(inp, disorder=100 / S.GAUSSIAN_SD_FWHM),
                           unit_convert=S.CM_FS, secular=False)
    for s in ('eg', 'ee'):
        Lb = mb.liouvillian(s).cpu().numpy()
        for n in range(6):
            assert rel_err(Lb[n], om.sample(n).evolve_matrix(s)) < 1e-12, (s, n)


def _debye_j(lam, gam):
    return lambda x: 2 * lam * gam * x / (gam ** 2 + x ** 2)


def _superohmic_j(x):
    return 35.0 * x ** 3 / 150.0 ** 2 * np.exp(-x / 150.0)


def test_arbitrary_bath_blocks(gpu):
    """ArbitraryBath (bath.py:50-56) through the tabulated builder: with the
    Debye J it reproduces the reference's discard_imag_corr blocks; with a
    super-ohmic J it matches the oracle member by member over a disorder
    batch, secular and not; the complex correlation function it lacks
    raises AttributeError as in the reference (redfield.py:48-53)."""
    g, inp = _inputs('redfield_dimer')
    T = S.CM_K * 77
    ham = qa.ElectronicHamiltonian(inp['H1'], bath=qa.ArbitraryBath(T, _debye_j(35, 106),
                                                                    2 * 35 / 106.),
                                   dipoles=inp['dipoles'], energy_spread_extra=inp['extra'])
    for sec in (True, False):
        m = qa.RedfieldModel(ham, unit_convert=qa.CM_FS, secular=sec, discard_imag_corr=True)
        for s in ['gg', 'eg', 'ge', 'ee', 'fg', 'fe']:
            ref = g['L_%s_%d1' % (s, sec)]
            assert rel_err(m.liouvillian(s)[0].cpu().numpy(), ref) < 1e-12, (s, sec)
            got = m.liouvillian(s, heisenberg_picture=True)[0].cpu().numpy()
            assert rel_err(got, ref.T) < 1e-12, (s, sec, 'heisenberg')
    with pytest.raises(AttributeError):
        qa.RedfieldModel(ham, unit_convert=qa.CM_FS).liouvillian('eg')

    _, inp = _inputs('redfield_7site')
    pham = qa.ElectronicHamiltonian(inp['H1'], bath=qa.ArbitraryBath(T, _superohmic_j, 0.0),
                                    dipoles=inp['dipoles'], energy_spread_extra=inp['extra'],
                                    disorder=100 / qa.GAUSSIAN_SD_FWHM)
    oham = S.ElectronicHamiltonian(inp['H1'], bath=S.ArbitraryBath(T, _superohmic_j, 0.0),
                                   dipoles=inp['dipoles'], energy_spread_extra=inp['extra'],
                                   disorder=100 / S.GAUSSIAN_SD_FWHM)
    for sec in (False, True):
        m = qa.RedfieldModel(pham, unit_convert=qa.CM_FS, secular=sec, discard_imag_corr=True)
        mb = m.with_hamiltonians([m.hamiltonian.sample(n) for n in range(4)])
        om = Lv.LiouvilleModel(oham, unit_convert=S.CM_FS, secular=sec, discard_imag_corr=True)
        for s in ('eg', 'ee', 'fe'):
            Lb = mb.liouvillian(s).cpu().numpy()
            for n in range(4):
                assert rel_err(Lb[n], om.sample(n).evolve_matrix(s)) < 1e-12, (s, n, sec)


def test_unitary_blocks_match_oracle(gpu):
    _, inp = _inputs('redfield_7site')
    m = qa.UnitaryModel(_product_ham(inp), unit_convert=qa.CM_FS)
    om = Lv.LiouvilleModel(_oracle_ham(inp), unit_convert=S.CM_FS, kind='unitary')
    for s in ['eg', 'ee', 'fe', 'gg,ge,eg,ee']:
        assert rel_err(m.liouvillian(s)[0].cpu().numpy(), om.evolve_matrix(s)) < 1e-12, s


def test_integrate_matches_oracle_and_reference(gpu):
    g, inp = _inputs('redfield_7site')
    m = qa.RedfieldModel(_product_ham(inp), unit_convert=qa.CM_FS, secular=False)
    om = Lv.LiouvilleModel(_oracle_ham(inp), unit_convert=S.CM_FS, secular=False)
    eom = m.equation_of_motion('ee')
    y = np.random.RandomState(0).randn(3, 49) + 0j
    assert rel_err(eom(0, y), y @ om.evolve_matrix('ee').T) < 1e-13
    V = m.dipole_operator('ee->eg', 'x', '-')
    t = np.arange(0, 100.0, 2.0)
    got = qa.integrate(eom, y, t, save_func=V.commutator, dt=2.0)
    ref = P.integrate_rk4(om.evolve_matrix('ee'), y, t, save_matrix=V.commutator.matrix, dt=2.0)
    assert got.shape == (3, 50, 7)
    assert rel_err(got, ref) < 1e-10
    gd = golden('dynamics')
    t, rho = qa.simulate_dynamics(m, [1, 0, 0, 0, 0, 0, 0], 200, dt=0.25)
    assert rel_err(rho, gd['rho']) < 1e-8          # RK4 at 0.25 fs vs tight zvode
    t, rho = qa.simulate_dynamics(m, [1, 0, 0, 0, 0, 0, 0], 200)   # default rtol step
    assert rel_err(rho, gd['rho']) < 1e-6


def test_absorption_spectra_match_reference(gpu):
    gs = golden('spectra_linear')
    g, inp = _inputs('redfield_dimer')
    md = qa.RedfieldModel(_product_ham(inp), unit_convert=qa.CM_FS)
    f, X = qa.absorption_spectra(md, 511.5, dt=1.0 / 8)
    assert np.allclose(f, gs['dimer_f'])
    assert rel_err(X, gs['dimer_X']) < 1e-8
    om = Lv.LiouvilleModel(_oracle_ham(inp), unit_convert=S.CM_FS)
    fo, Xo = P.absorption_spectra(om, 511.5, dt=1.0)
    f, X = qa.absorption_spectra(md, 511.5, dt=1.0)
    assert rel_err(X, Xo) < 1e-10
    f, X = qa.absorption_spectra(md, 511.5, dt=1.0, precision='mixed')
    assert rel_err(X, Xo) < 1e-6
    f, X = qa.absorption_spectra(md, 511.5, exact_isotropic_average=True, dt=1.0 / 8)
    assert rel_err(X, gs['dimer_X_iso']) < 1e-8
    _, inp7 = _inputs('redfield_7site')
    m7 = qa.RedfieldModel(_product_ham(inp7, disorder=100 / qa.GAUSSIAN_SD_FWHM),
                          unit_convert=qa.CM_FS, secular=False)
    f, X = qa.absorption_spectra(m7, 2 * 255.5, dt=2.0 / 8)
    assert rel_err(X, gs['agg7_X']) < 1e-7
    f, X = qa.absorption_spectra(m7, 2 * 127.5, ensemble_size=8, dt=2.0 / 8)
    assert rel_err(X, gs['agg7_ens8_X']) < 1e-7
    f, X = qa.absorption_spectra(m7, 2 * 127.5, correlation_decay_time=300.0,
                                 polarization=[0.3, 1.1], dt=2.0 / 8)
    assert rel_err(X, gs['agg7_decay_X']) < 1e-7


def test_third_order_and_2d_match_reference(gpu):
    gt = golden('spectra_third_order')
    _, inp = _inputs('redfield_7site')
    m7 = qa.RedfieldModel(_product_ham(inp), unit_convert=qa.CM_FS, secular=False)
    om = Lv.LiouvilleModel(_oracle_ham(inp), unit_convert=S.CM_FS, secular=False)
    for geom in ('-++', '+-+', '++-'):
        key = 'R3_' + geom.replace('-', 'm').replace('+', 'p')
        (t1, t2, t3), X = qa.third_order_response(m7, 2 * 15.5, population_times=gt['t2'],
                                                  geometry=geom, dt=2.0)
        _, Xo = P.third_order_response(om, 2 * 15.5, population_times=gt['t2'],
                                       geometry=geom, dt=2.0)
        assert rel_err(X, Xo) < 1e-10, geom
        (t1, t2, t3), X = qa.third_order_response(m7, 2 * 15.5, population_times=gt['t2'],
                                                  geometry=geom, dt=2.0 / 8)
        assert rel_err(X, gt[key]) < 1e-6, geom
    (f1, t2, f3), X = qa.two_dimensional_spectra(m7, 2 * 15.5, population_times=gt['t2'],
                                                 include_signal='GSB,ESE', dt=2.0 / 8)
    assert np.allclose(f1, gt['S2d_f1']) and np.allclose(f3, gt['S2d_f3'])
    assert rel_err(X, gt['S2d']) < 1e-6
    (f1, t2, f3), Xm = qa.two_dimensional_spectra(m7, 2 * 15.5, population_times=gt['t2'],
                                                  include_signal='GSB,ESE', dt=2.0 / 8,
                                                  precision='mixed')
    assert rel_err(Xm, X) < 1e-6


def test_isotropic_third_order_dimer(gpu):
    gt = golden('spectra_third_order')
    g, inp = _inputs('redfield_dimer')
    md = qa.RedfieldModel(_product_ham(inp), unit_convert=qa.CM_FS)
    (t1, t2, t3), X = qa.third_order_response(
        md, 8.5, population_times=[0., 3., 7.], exact_isotropic_average=True,
        polarization=[0, 0, qa.MAGIC_ANGLE, qa.MAGIC_ANGLE], dt=1.0 / 16)
    assert rel_err(X, gt['dimer_iso_R3']) < 1e-7


def test_fused_plan_and_graph_replay(gpu):
    """The fused plan equals the pathway-by-pathway evaluation; its HIP-graph
    replays equal the eager launch every time, with eager work in between."""
    from qspectra_amd.simulate.fused import ThirdOrderPlan
    gt = golden('spectra_third_order')
    _, inp = _inputs('redfield_7site')
    m7 = qa.RedfieldModel(_product_ham(inp), unit_convert=qa.CM_FS, secular=False)
    _, ref = qa.third_order_response(m7, 2 * 15.5, population_times=gt['t2'], dt=2.0,
                                     schedule='rk4', plan='pathways')
    plan = ThirdOrderPlan(m7, 2 * 15.5, population_times=gt['t2'], dt=2.0)
    eager = plan.run()[0].cpu().numpy()
    plan.check_status()
    assert rel_err(eager, ref) < 1e-10
    plan.capture()
    for _ in range(3):
        X = plan.run()[0].cpu().numpy()
        assert rel_err(X, eager) < 1e-13
        qa.absorption_spectra(m7, 2 * 31.5, dt=2.0)       # unrelated eager work
    pm = ThirdOrderPlan(m7, 2 * 15.5, population_times=gt['t2'], dt=2.0,
                        precision='mixed').capture()
    Xm = pm.run()[0].cpu().numpy()
    pm.check_status()
    assert rel_err(Xm, ref) < 1e-6
    for geom in ('+-+', '++-'):
        _, a = qa.third_order_response(m7, 2 * 15.5, population_times=gt['t2'], dt=2.0,
                                       geometry=geom, plan='pathways', schedule='rk4')
        _, b = qa.third_order_response(m7, 2 * 15.5, population_times=gt['t2'], dt=2.0,
                                       geometry=geom)
        assert rel_err(b, a) < 1e-10, geom


def test_blocked_schedules_and_2d_graph(gpu):
    """Time-chunked t1/t3 trajectories and the reassociated t2 stage equal
    step-by-step RK4 pathway by pathway; the 2-D spectrum of a fourier plan,
    replayed from one graph, equals the eager API every time."""
    from qspectra_amd.simulate.fused import ThirdOrderPlan
    _, inp = _inputs('redfield_7site')
    m7 = qa.RedfieldModel(_product_ham(inp), unit_convert=qa.CM_FS, secular=False)
    om = Lv.LiouvilleModel(_oracle_ham(inp), unit_convert=S.CM_FS, secular=False)
    t2 = [0.0, 30.0, 62.0]
    (t1, _, t3), ref = P.third_order_response(om, 2 * 69.5, population_times=t2, dt=2.0)
    plan = ThirdOrderPlan(m7, 2 * 69.5, population_times=t2, dt=2.0)
    kinds = set([st.kind for st, _ in plan.t1_stages.values()]
                + [st['stage'].kind for st in plan.t3_stages.values()])
    assert kinds <= {'chunks', 'dchunks', 'doubling', 'stepper'}, kinds   # autotuned
    X = plan.run()[0].cpu().numpy()
    plan.check_status()
    assert len(t1) == 70 and rel_err(X, ref) < 1e-10
    pm = ThirdOrderPlan(m7, 2 * 69.5, population_times=t2, dt=2.0, precision='mixed')
    assert rel_err(pm.run()[0].cpu().numpy(), ref) < 1e-6
    (f1, _, f3), S2 = qa.two_dimensional_spectra(m7, 2 * 69.5, population_times=t2, dt=2.0,
                                                 plan='pathways', schedule='rk4')
    g2 = ThirdOrderPlan(m7, 2 * 69.5, population_times=t2, dt=2.0, fourier=True).capture()
    ff1, _, ff3 = g2.frequencies()
    assert np.allclose(ff1, f1) and np.allclose(ff3, f3)
    for _ in range(3):
        Y = g2.run()[0].cpu().numpy()
        assert rel_err(Y, S2) < 1e-10
        qa.absorption_spectra(m7, 2 * 31.5, dt=2.0)       # eager work between replays


def test_fused_plan_ensemble_and_isotropic(gpu):
    """Plans over a batch of disorder members (per-member Liouvillians) and
    the 4th-order isotropic average equal the pathway-by-pathway path."""
    _, inp = _inputs('redfield_7site')
    ham = _product_ham(inp, disorder=80)
    m7 = qa.RedfieldModel(ham, unit_convert=qa.CM_FS, secular=False)
    kw = dict(population_times=[0.0, 40.0], dt=2.0, ensemble_size=3)
    _, a = qa.third_order_response(m7, 2 * 11.5, plan='pathways', schedule='rk4', **kw)
    _, b = qa.third_order_response(m7, 2 * 11.5, **kw)
    assert rel_err(b, a) < 1e-10
    (_, _, _), c = qa.two_dimensional_spectra(m7, 2 * 11.5, population_times=[0.0, 40.0],
                                              dt=2.0, exact_isotropic_average=True)
    (_, _, _), d = qa.two_dimensional_spectra(m7, 2 * 11.5, population_times=[0.0, 40.0],
                                              dt=2.0, exact_isotropic_average=True,
                                              plan='pathways', schedule='rk4')
    assert rel_err(c, d) < 1e-10


@pytest.mark.parametrize('kind,chunk,precision', [('doubling', None, 'fp64'),
                                                  ('chunks', 4, 'fp64'), ('chunks', 16, 'mixed'),
                                                  ('dchunks', 8, 'mixed'), ('dchunks', 16, 'fp64'),
                                                  ('stepper', None, 'fp64'),
                                                  ('stepper', None, 'mixed')])
def test_trajectory_schedules_match_oracle(gpu, kind, chunk, precision):
    """Every trajectory schedule of the 'fe' Heisenberg block (M = 147) over
    100 points, and the autotuned choice, against the oracle's RK4."""
    from qspectra_amd.schedules import Stage
    _, inp = _inputs('redfield_7site')
    m7 = qa.RedfieldModel(_product_ham(inp), unit_convert=qa.CM_FS, secular=False)
    om = Lv.LiouvilleModel(_oracle_ham(inp), unit_convert=S.CM_FS, secular=False)
    Lh = om.evolve_matrix('fe', heisenberg_picture=True)
    b = om.dipole_operator('fe->ee', 'x', '-').bra_vector
    t = np.arange(0, 200.0, 2.0)
    ref = P.integrate_rk4(Lh, b, t)
    eom = m7.equation_of_motion('fe', heisenberg_picture=True)
    y0 = _native.to_device(b[None, None, :])
    tol = 1e-10 if precision == 'fp64' else 1e-6
    st = Stage(eom, t, None, 2.0, precision, 'auto', 1, chunk=chunk, kind=kind)
    out = st.run(y0, [])[0, 0].cpu().numpy()
    assert rel_err(out, ref) < tol, (kind, chunk, precision)
    tuned = Stage(eom, t, None, 2.0, precision, 'auto', 1)
    assert rel_err(tuned.run(y0, [])[0, 0].cpu().numpy(), ref) < tol, tuned.kind


def test_impulsive_probe_matches_oracle(gpu):
    """impulsive_probe (response.py:157-231) of several probe-time states of
    the 7-site aggregate, as right-hand sides of one set of launches."""
    _, inp = _inputs('redfield_7site')
    m7 = qa.RedfieldModel(_product_ham(inp), unit_convert=qa.CM_FS, secular=False)
    om = Lv.LiouvilleModel(_oracle_ham(inp), unit_convert=S.CM_FS, secular=False)
    sub = 'gg,ge,eg,ee'
    rs = np.random.RandomState(8)
    n = len(om.liouville_subspace_index(sub))
    states = om.thermal_state(sub) + 0.05 * (rs.randn(3, n) + 1j * rs.randn(3, n))
    f, X = qa.impulsive_probe(m7, states, 2 * 63.5, dt=2.0)
    fo, Xo = P.impulsive_probe(om, states, 2 * 63.5, dt=2.0)
    assert np.allclose(f, fo) and X.shape == Xo.shape == (3, len(fo))
    assert rel_err(X, Xo) < 1e-10
    _, Xm = qa.impulsive_probe(m7, states, 2 * 63.5, dt=2.0, precision='mixed')
    assert rel_err(Xm, Xo) < 1e-6
    _, Xg = qa.impulsive_probe(m7, states[0], 2 * 63.5, dt=2.0, include_signal='GSB')
    _, Xog = P.impulsive_probe(om, states[0], 2 * 63.5, dt=2.0, include_signal='GSB')
    assert rel_err(Xg, Xog) < 1e-10


def test_reference_spectra_tests(gpu):
    """tests/test_spectra.py of the reference: the Redfield monomer's
    absorption peaks at its transition energy, rises monotonically to it and
    falls after it, also in another rotating frame; the dimer's absorption is
    the same in the site and eigen bases with and without sparse matrices."""
    from qspectra_amd.simulate.utils import is_constant
    ham = qa.ElectronicHamiltonian(np.array([[12500.]]), bath=qa.DebyeBath(qa.CM_K * 77, 35, 106),
                                   dipoles=[[1, 0, 0]])
    dyn = qa.RedfieldModel(ham, discard_imag_corr=True, unit_convert=qa.CM_FS)
    f, X = qa.absorption_spectra(dyn, 10000)
    assert abs(f[np.argmax(X)] - 12500) < 1e-7
    assert is_constant(np.diff(f))
    assert np.all(np.diff(X)[:int(len(X) / 2.0 - 1)] > 0)
    assert np.all(np.diff(X)[-int(len(X) / 2.0 - 1):] < 0)
    dyn2 = qa.RedfieldModel(ham, rw_freq=12400, discard_imag_corr=True, unit_convert=qa.CM_FS)
    f2, X2 = qa.absorption_spectra(dyn2, 10000)
    assert abs(f2[np.argmax(X2)] - 12500) < 0.5
    hd = qa.ElectronicHamiltonian(np.array([[12881., 120.], [120., 12719.]]),
                                  bath=qa.DebyeBath(qa.CM_K * 77, 35, 106),
                                  dipoles=[[1, 0, 0], [2 * np.cos(.3), 2 * np.sin(.3), 0]])
    spectra = []
    for basis in ('site', 'eigen'):
        for sp in (False, True, lambda mat: True):
            m = qa.RedfieldModel(hd, hilbert_subspace='gef', discard_imag_corr=True,
                                 unit_convert=qa.CM_FS, evolve_basis=basis, sparse_matrix=sp)
            spectra.append(qa.absorption_spectra(m, 10000, dt=1.0)[1])   # one schedule
    for X in spectra[:-1]:
        assert rel_err(X, spectra[-1]) < 1e-9


def test_edge_shapes_match_oracle(gpu):
    """Shapes at the edges of the pathway machinery against the oracle on
    the same schedule: a monomer in g/e/f (an empty two-exciton block), a
    single population time, coherence grids of one point, and an ensemble of
    one member."""
    mono_p = qa.ElectronicHamiltonian(np.array([[12500.]]), bath=qa.DebyeBath(qa.CM_K * 77, 35, 106),
                                      dipoles=[[1., 0, 0]])
    mono_o = S.ElectronicHamiltonian(np.array([[12500.]]), bath=S.DebyeBath(S.CM_K * 77, 35, 106),
                                     dipoles=[[1., 0, 0]])
    mp = qa.RedfieldModel(mono_p, hilbert_subspace='gef', unit_convert=qa.CM_FS)
    mo = Lv.LiouvilleModel(mono_o, hilbert_subspace='gef', unit_convert=S.CM_FS)
    dt = mp.time_step / 4
    for geom in ('-++', '+-+'):          # ESA needs the (empty) 'fe' block
        kw = dict(population_times=[0., 10.], geometry=geom, include_signal='GSB,ESE', dt=dt)
        (t1, t2, t3), R = qa.third_order_response(mp, 100.5, **kw)
        _, Ro = P.third_order_response(mo, 100.5, **kw)
        assert R.shape == Ro.shape
        assert np.allclose(R, Ro, rtol=0, atol=1e-10 * max(1.0, np.abs(Ro).max()))
    _, inp = _inputs('redfield_dimer')
    md = qa.RedfieldModel(_product_ham(inp), hilbert_subspace='gef', unit_convert=qa.CM_FS)
    od = Lv.LiouvilleModel(_oracle_ham(inp), hilbert_subspace='gef', unit_convert=S.CM_FS)
    for kw in (dict(coherence_time_max=30.5, population_times=[7.]),
               dict(coherence_time_max=0.5, population_times=[0., 3.])):
        (t1, t2, t3), R = qa.third_order_response(md, dt=0.25, **kw)
        _, Ro = P.third_order_response(od, dt=0.25, **kw)
        assert R.shape == Ro.shape == (len(t1), len(t2), len(t3))
        assert rel_err(R, Ro) < 1e-10
    hd = _product_ham(inp, disorder=50.0)
    ms = qa.RedfieldModel(hd, hilbert_subspace='ge', unit_convert=qa.CM_FS)
    f, X = qa.absorption_spectra(ms, 100.5, dt=0.25, ensemble_size=1)
    os_ = Lv.LiouvilleModel(_oracle_ham(inp, disorder=50.0), hilbert_subspace='ge',
                            unit_convert=S.CM_FS)
    fo, Xo = P.absorption_spectra(os_.sample(0), 100.5, dt=0.25)
    assert rel_err(X, Xo) < 1e-10
