"""Field-driven propagation (simulate_with_fields, simulate_pump,
qspectra/simulate/eom.py:79-227) on the GPU: the driven RK4 kernel against
the oracle's time-dependent RK4 on the same schedule (FP64 1e-10 relative,
MIXED 1e-6) and against the reference's tight zvode outputs
(tests/golden/fields.npz) at the default step (RK4 error, 1e-5).
"""
import numpy as np
import pytest
import torch

import qspectra_amd as qa
from qspectra_amd import _native
from qspectra_amd.dynamics.liouville_space import LinearMap
from oracle import system as S
from oracle import liouville as Lv
from oracle import propagate as P
from tests.conftest import golden, rel_err

pytestmark = pytest.mark.gpu


def _hams(name, **kw):
    g = golden(name)
    prod = qa.ElectronicHamiltonian(g['in_H1'], bath=qa.DebyeBath(S.CM_K * 77, 35, 106),
                                    dipoles=g['in_dipoles'],
                                    energy_spread_extra=float(g['extra']), **kw)
    orac = S.ElectronicHamiltonian(g['in_H1'], bath=S.DebyeBath(S.CM_K * 77, 35, 106),
                                   dipoles=g['in_dipoles'],
                                   energy_spread_extra=float(g['extra']), **kw)
    return prod, orac


def _dimer_models(kind='redfield'):
    hp, ho = _hams('redfield_dimer')
    if kind == 'unitary':
        return (qa.UnitaryModel(hp, hilbert_subspace='gef', unit_convert=qa.CM_FS),
                Lv.LiouvilleModel(ho, hilbert_subspace='gef', unit_convert=S.CM_FS,
                                  kind='unitary'))
    return (qa.RedfieldModel(hp, hilbert_subspace='gef', unit_convert=qa.CM_FS),
            Lv.LiouvilleModel(ho, hilbert_subspace='gef', unit_convert=S.CM_FS))


def _pumps():
    return (qa.GaussianPulse(12750., 20., t_peak=0., scale=0.02, freq_convert=qa.CM_FS),
            S.GaussianPulse(12750., 20., t_peak=0., scale=0.02, freq_convert=S.CM_FS))


def test_pump_matches_oracle_and_reference(gpu):
    g = golden('fields')
    mp, mo = _dimer_models()
    pp, po = _pumps()
    dt = mp.time_step / 4
    t, y = qa.simulate_pump(mp, pp, 'x', time_extra=40., dt=dt)
    to, yo = P.simulate_pump(mo, po, 'x', time_extra=40., dt=dt)
    np.testing.assert_allclose(t, to)
    assert rel_err(y, yo) < 1e-10
    _, ym = qa.simulate_pump(mp, pp, 'x', time_extra=40., dt=dt, precision='mixed')
    assert rel_err(ym, yo) < 1e-6
    t, y = qa.simulate_pump(mp, pp, 'x', time_extra=40.)          # default step
    np.testing.assert_allclose(t, g['pump_t'])
    assert rel_err(y, g['pump_y']) < 1e-5
    t, y = qa.simulate_pump(mp, pp, [1., 0.5, 0.], times=np.arange(0., 60., 3.),
                            exact_isotropic_average=True)
    np.testing.assert_allclose(t, g['iso_t'])
    assert rel_err(y, g['iso_y']) < 1e-5


def test_several_fields_and_save_funcs(gpu):
    g = golden('fields')
    mp, mo = _dimer_models()
    p1 = qa.GaussianPulse(12700., 15., t_peak=-10., scale=0.015, freq_convert=qa.CM_FS)
    p2 = qa.GaussianPulse(12900., 25., t_peak=5., scale=0.01 * np.exp(0.6j),
                          freq_convert=qa.CM_FS, t_limits_multiple=2.5)
    o1 = S.GaussianPulse(12700., 15., t_peak=-10., scale=0.015, freq_convert=S.CM_FS)
    o2 = S.GaussianPulse(12900., 25., t_peak=5., scale=0.01 * np.exp(0.6j),
                         freq_convert=S.CM_FS, t_limits_multiple=2.5)
    dt = 0.25
    t, y = qa.simulate_with_fields(mp, [p1, p2], '-+', 'xy', time_extra=20., dt=dt)
    _, yo2 = P.simulate_with_fields(mo, [o1, o2], '-+', 'xy', time_extra=20., dt=dt)
    assert rel_err(y, yo2) < 1e-10
    t, y = qa.simulate_with_fields(mp, [p1, p2], '-+', 'xy', time_extra=20.)
    np.testing.assert_allclose(t, g['two_t'])
    assert rel_err(y, g['two_y']) < 1e-5
    # three field terms (the kernel for up to four terms)
    t, y = qa.simulate_with_fields(mp, [p1, p2, p1], '-+-', ['x', 'y', 0.7],
                                   times=[0., 5., 12.5], dt=dt)
    _, yo = P.simulate_with_fields(mo, [o1, o2, o1], '-+-', ['x', 'y', 0.7],
                                   times=[0., 5., 12.5], dt=dt)
    assert rel_err(y, yo) < 1e-10
    t, y = qa.simulate_with_fields(mp, [p1, p2, p1], '-+-', ['x', 'y', 0.7],
                                   times=[0., 5., 12.5])
    assert rel_err(y, g['three_y']) < 1e-5
    # save functions: a linear map (applied on the device) and a callable
    sub = 'gg,ge,eg,ee'
    ee = mp.dipole_operator(sub, 'x', '-+')
    row = ee.expectation_value
    _, ev = qa.simulate_with_fields(mp, [p1, p2], '-+', 'xy', time_extra=20., dt=dt,
                                    save_func=row)
    assert rel_err(ev, yo2 @ np.asarray(row.matrix)) < 1e-10
    _, sq = qa.simulate_with_fields(mp, [p1, p2], '-+', 'xy', time_extra=20., dt=dt,
                                    save_func=lambda v: np.abs(v[:3]) ** 2)
    assert rel_err(sq, np.abs(yo2[:, :3]) ** 2) < 1e-9


def test_unitary_aggregate_and_ensemble(gpu):
    g = golden('fields')
    mp, mo = _dimer_models('unitary')
    pp, po = _pumps()
    _, y = qa.simulate_pump(mp, pp, 'y', time_extra=30., dt=0.25)
    _, yo = P.simulate_pump(mo, po, 'y', time_extra=30., dt=0.25)
    assert rel_err(y, yo) < 1e-10
    _, y = qa.simulate_pump(mp, pp, 'y', time_extra=30.)
    assert rel_err(y, g['unitary_y']) < 1e-5
    hp, ho = _hams('redfield_7site', disorder=100 / S.GAUSSIAN_SD_FWHM)
    m7 = qa.RedfieldModel(hp, hilbert_subspace='gef', unit_convert=qa.CM_FS, secular=False)
    o7 = Lv.LiouvilleModel(ho, unit_convert=S.CM_FS, secular=False)
    p7 = qa.GaussianPulse(12450., 20., scale=0.01, freq_convert=qa.CM_FS)
    q7 = S.GaussianPulse(12450., 20., scale=0.01, freq_convert=S.CM_FS)
    t, y = qa.simulate_pump(m7, p7, 'x', time_extra=20., dt=0.5)       # M = 64
    _, yo = P.simulate_pump(o7, q7, 'x', time_extra=20., dt=0.5)
    assert y.shape[-1] == 64 and rel_err(y, yo) < 1e-10
    _, ym = qa.simulate_pump(m7, p7, 'x', time_extra=20., dt=0.5, precision='mixed')
    assert rel_err(ym, yo) < 1e-6
    t, y = qa.simulate_pump(m7, p7, 'x', time_extra=20.)
    np.testing.assert_allclose(t, g['agg7_t'])
    assert rel_err(y, g['agg7_y']) < 1e-5
    times = np.array([0., 10., 20.])
    _, y = qa.simulate_pump(m7, p7, 'z', times=times, ensemble_size=2, dt=0.5)
    _, yo = P.ensemble_average(
        lambda mm: P.simulate_pump(mm, q7, 'z', times=times, dt=0.5), o7, 2)
    assert rel_err(y, yo) < 1e-10
    _, y = qa.simulate_pump(m7, p7, 'z', times=times, ensemble_size=2)
    assert rel_err(y, g['agg7_ens2_y']) < 1e-5


def test_pump_probe_end_to_end(gpu):
    """simulate_pump then impulsive_probe of the pumped states, against the
    reference's pump-probe signal."""
    g = golden('fields')
    mp, _ = _dimer_models()
    pp, _ = _pumps()
    _, y = qa.simulate_pump(mp, pp, 'x', time_extra=40.)
    states = y[[-1, -10, -20]]
    assert rel_err(states, g['probe_states']) < 1e-5
    f, X = qa.impulsive_probe(mp, g['probe_states'], 300.5)
    np.testing.assert_allclose(f, g['probe_f'])
    assert rel_err(X, g['probe_X']) < 1e-5
    f, X = qa.impulsive_probe(mp, g['probe_states'][0], 300.5, polarization='xy',
                              include_signal='ESE,ESA')
    assert rel_err(X, g['probe1_X']) < 1e-5


def test_driven_kernel_status_and_errors(gpu):
    """The kernel stops with status 2 when the field record is shorter than
    the schedule; bad sizes are rejected on the host."""
    M, P_ = 4, 1
    dev = 'cuda'
    A = torch.zeros((1, M, M), dtype=torch.complex128, device=dev)
    col, coef = _native.ell_pack(np.zeros((1, P_, M, M), complex))
    col = torch.as_tensor(col, device=dev)
    coef = torch.as_tensor(coef, device=dev)
    F = torch.zeros((1, 3, 3, P_), dtype=torch.complex128, device=dev)
    y0 = torch.ones((1, 1, M), dtype=torch.complex128, device=dev)
    nsub = torch.tensor([2, 2], dtype=torch.int32, device=dev)
    h = torch.tensor([0.1, 0.1], dtype=torch.float64, device=dev)
    out, status = _native.propagate_driven(A, col, coef, F, y0, nsub, h)
    assert int(status.item()) == 2
    F = torch.zeros((1, 4, 3, P_), dtype=torch.complex128, device=dev)
    out, status = _native.propagate_driven(A, col, coef, F, y0, nsub, h)
    assert int(status.item()) == 0
    assert torch.equal(out[0, 0, 1], y0[0, 0])                  # L = 0, no field
    # 65 elements: the LDS-resident kernel, the same status handling
    big = torch.zeros((1, 65, 65), dtype=torch.complex128, device=dev)
    c65, k65 = _native.ell_pack(np.zeros((1, 1, 65, 65), complex))
    y65 = torch.ones((1, 1, 65), dtype=torch.complex128, device=dev)
    out, status = _native.propagate_driven(big, torch.as_tensor(c65, device=dev),
                                           torch.as_tensor(k65, device=dev), F, y65, nsub, h)
    assert int(status.item()) == 0 and torch.equal(out[0, 0, 1], y65[0, 0])
    F3 = torch.zeros((1, 3, 3, P_), dtype=torch.complex128, device=dev)
    _, status = _native.propagate_driven(big, torch.as_tensor(c65, device=dev),
                                         torch.as_tensor(k65, device=dev), F3, y65, nsub, h)
    assert int(status.item()) == 2
    M1 = _native.MAX_DRIVEN_BLOCK + 1
    with pytest.raises(ValueError):
        _native.propagate_driven(torch.zeros((1, M1, M1), dtype=torch.complex128, device=dev),
                                 torch.zeros((1, M1, 1), dtype=torch.int32, device=dev),
                                 torch.zeros((1, M1, 1, 1), dtype=torch.complex128, device=dev),
                                 F, torch.ones((1, 1, M1), dtype=torch.complex128, device=dev),
                                 nsub, h)


def test_batched_thermal_states(gpu):
    """thermal_states of an ensemble (one eigensolver launch per manifold)
    equals thermal_state member by member, at 77 K and 300 K and in
    subspaces where the Boltzmann weights matter ('ee')."""
    from tests.conftest import golden
    g = golden('redfield_7site')
    for T in (S.CM_K * 77, S.CM_K * 3000):
        ham = qa.ElectronicHamiltonian(g['in_H1'], bath=qa.DebyeBath(T, 35, 106),
                                       dipoles=g['in_dipoles'],
                                       energy_spread_extra=float(g['extra']),
                                       disorder=100 / S.GAUSSIAN_SD_FWHM)
        m = qa.RedfieldModel(ham, hilbert_subspace='gef', unit_convert=qa.CM_FS)
        from qspectra_amd.simulate.response import _member_model
        hams = list(m.hamiltonian.sample_ensemble(5))
        mb = m.with_hamiltonians(hams)
        for sub in ('gg,ge,eg,ee', 'ee', 'gg', 'ge,eg,ee'):
            got = mb.thermal_states(sub)
            ref = np.stack([_member_model(mb, h).thermal_state(sub) for h in hams])
            assert got.shape == ref.shape and np.abs(got - ref).max() < 1e-13, (T, sub)
