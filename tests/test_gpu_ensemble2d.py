"""The batched 2-D path: one plan over a static-disorder ensemble of the
7-site aggregate (c1 grid: t1 = t3 = 128 points at 2 fs, t2 in {0, 100, 200,
500} fs, GSB + ESE + ESA; RK4 at 1 fs), contracted to the ensemble mean in
the time domain, with the Liouvillian blocks rebuilt from the members'
Hamiltonians inside every launch.

Tolerances: FP64 against the oracle's serial ensemble mean of RK4
responses on the same step, 1e-10; MIXED (including the FP32-matrix-core
split of the 147-element block) 1e-6; the ensemble-mean 2-D spectrum
against the mean of the per-member spectra of the per-member plan, 1e-12.
"""
import numpy as np
import pytest
import torch

import qspectra_amd as qa
from qspectra_amd import workloads as W
from qspectra_amd.constants import GAUSSIAN_SD_FWHM
from oracle import system as S, liouville as Lv, propagate as P
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

WORK = W.AGG7_2D
SIGMA = 100.0


def _models():
    H1, d, extra = W.aggregate7_hamiltonian()
    dis = SIGMA / GAUSSIAN_SD_FWHM
    m = qa.RedfieldModel(qa.ElectronicHamiltonian(H1, bath=qa.DebyeBath(S.CM_K * 77, 35, 106),
                                                  dipoles=d, disorder=dis,
                                                  energy_spread_extra=extra),
                         unit_convert=qa.CM_FS, secular=False)
    om = Lv.LiouvilleModel(S.ElectronicHamiltonian(H1, bath=S.DebyeBath(S.CM_K * 77, 35, 106),
                                                   dipoles=d, disorder=dis,
                                                   energy_spread_extra=extra),
                           unit_convert=S.CM_FS, secular=False)
    return m, om


def _oracle_mean(om, members):
    total = 0
    for n in members:
        total = total + P.third_order_response(
            om.sample(n), WORK['coherence_time_max'], population_times=WORK['population_times'],
            geometry=WORK['geometry'], include_signal=WORK['include_signal'], dt=WORK['dt'])[1]
    return total / len(members)


def _plan(mb, **kw):
    from qspectra_amd.simulate.fused import ThirdOrderPlan
    return ThirdOrderPlan(mb, WORK['coherence_time_max'],
                          population_times=WORK['population_times'], geometry=WORK['geometry'],
                          include_signal=WORK['include_signal'], dt=WORK['dt'], **kw)


def test_ensemble_mean_plan_rebuilt_from_hamiltonians(gpu):
    m, om = _models()
    B = 6
    mb = m.with_hamiltonians(list(m.hamiltonian.sample_ensemble(B)))
    refA = _oracle_mean(om, range(B))
    refB = _oracle_mean(om, range(B, 2 * B))
    other = torch.as_tensor(np.stack([m.hamiltonian.sample(n).H1_nonrotating
                                      for n in range(B, 2 * B)])).cuda()
    for prec, chunk, tol in (('fp64', None, 1e-10), ('fp64', 128, 1e-10),
                             ('mixed', None, 1e-6), ('mixed', 128, 1e-6)):
        plan = _plan(mb, precision=prec, chunk=chunk, ensemble_mean=True, rebuild=True)
        plan.capture()
        X = plan.run().cpu().numpy()
        plan.check_status()
        assert X.shape == (1, 128, 4, 128)
        assert rel_err(X[0], refA) < tol, (prec, chunk, rel_err(X[0], refA))
        # new Hamiltonians written in place: the replay rebuilds every block
        H1 = plan.hamiltonian_inputs[0]
        saved = H1.clone()
        H1.copy_(other)
        XB = plan.run().cpu().numpy()
        assert rel_err(XB[0], refB) < tol, (prec, chunk)
        H1.copy_(saved)
        assert rel_err(plan.run().cpu().numpy(), X) < 1e-13


def test_ensemble_mean_spectrum_matches_per_member_plan(gpu):
    m, _ = _models()
    B = 40                                     # >= 32: one chunk per trajectory
    mb = m.with_hamiltonians(list(m.hamiltonian.sample_ensemble(B)))
    per = _plan(mb, precision='fp64', fourier=True)
    Xper = per.run().cpu().numpy().mean(axis=0)
    for prec, tol in (('fp64', 1e-12), ('mixed', 1e-6)):
        ens = _plan(mb, precision=prec, fourier=True, ensemble_mean=True, rebuild=True).capture()
        assert ens._t3_list[0]['stage'].nc == 1
        for _ in range(2):
            X = ens.run().cpu().numpy()
            ens.check_status()
            assert rel_err(X[0], Xper) < tol, (prec, rel_err(X[0], Xper))
        assert np.allclose(ens.frequencies()[0], per.frequencies()[0])
