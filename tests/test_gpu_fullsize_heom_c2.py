"""Full-size parity of the two batched workloads beside the 2-D spectrum:

* the depth-4 hierarchy of the 7-site aggregate (300 K, K = 1, level_cutoff
  5: 3060 ADOs x 49 = 149,940 elements per member), two disorder members in
  one batch, against the oracle's sparse generator (heom.py:298-443
  restated) built for each member: right-hand side 1e-13, 20 RK4 steps of
  0.5 fs 1e-11 (FP64), MIXED 1e-5;
* the absorption ensemble of aggregates of every size 2..24 sites plus
  disorder members of the 7-site aggregate, 4096 points at 1 fs, one batch,
  member by member against the oracle's RK4 absorption spectra: FP64 1e-10,
  MIXED 1e-6.
"""
import numpy as np
import pytest
import torch

import qspectra_amd as qa
from qspectra_amd import workloads as W
from qspectra_amd.simulate.utils import propagate_heom
from oracle import system as S, liouville as Lv, propagate as P
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def test_depth4_hierarchy_matches_oracle_generator(gpu):
    m = W.heom_aggregate7_model()
    members = [m.hamiltonian.sample(n) for n in range(2)]
    eom = m.with_hamiltonians(members).equation_of_motion('ee')
    assert eom.n_ado == 3060 and eom.size == 3060 * 49
    rw = float(m.hamiltonian.rw_freq)
    rs = np.random.RandomState(4)
    y = rs.randn(2, eom.size) + 1j * rs.randn(2, eom.size)
    got = eom(0, y)
    y0 = np.zeros((2, eom.size), complex)
    y0[:, :49] = rs.randn(2, 49) + 1j * rs.randn(2, 49)
    t = np.arange(0, 10.0 + 1e-9, 1.0)                      # 20 steps of 0.5 fs
    y0d = torch.as_tensor(y0).cuda()
    r64 = propagate_heom(eom, y0d, t, dt=0.5).cpu().numpy()
    r32 = propagate_heom(eom, y0d, t, dt=0.5, precision='mixed').cpu().numpy()
    for b, h in enumerate(members):
        oh = S.ElectronicHamiltonian(h.H1_nonrotating, bath=S.DebyeBath(S.CM_K * 300, 35, 106),
                                     dipoles=h.dipoles).in_rotating_frame(rw)
        L = Lv.heom_evolve_matrix(oh, 'ge', 'ee', S.CM_FS, K=1, level_cutoff=5)
        assert L.shape == (eom.size, eom.size)
        assert rel_err(np.asarray(got[b]), L @ y[b]) < 1e-13, b
        ref = P.integrate_rk4(L, y0[b], t, dt=0.5)[:, :49]
        assert rel_err(r64[b], ref) < 1e-11, b
        assert rel_err(r32[b], ref) < 1e-5, b


class _FixedStep(Lv.LiouvilleModel):
    time_step = 1.0


def test_absorption_ensemble_every_size_2_to_24(gpu):
    from qspectra_amd.batched import absorption_batch
    hams = list(W.ensemble_absorption_hamiltonians(n_disorder=4, n_aggregates=0))
    rng = np.random.RandomState(2)
    for n in range(2, 25):
        J = rng.normal(0, 80, (n, n))
        H = np.triu(J, 1)
        H = H + H.T + np.diag(rng.uniform(12200, 12700, n))
        dip = rng.normal(size=(n, 3))
        dip /= np.linalg.norm(dip, axis=1, keepdims=True)
        hams.append(qa.ElectronicHamiltonian(H, bath=W.drude_bath(), dipoles=dip))
    rw, tmax = 12450.0, 4096.0
    f, X = absorption_batch(hams, tmax, rw_freq=rw, dt=1.0)
    _, Xm = absorption_batch(hams, tmax, rw_freq=rw, dt=1.0, precision='mixed')
    assert X.shape == (len(hams), len(f))
    for i, h in enumerate(hams):
        om = _FixedStep(S.ElectronicHamiltonian(h.H1_nonrotating,
                                                bath=S.DebyeBath(S.CM_K * 77, 35, 106),
                                                dipoles=h.dipoles),
                        rw_freq=rw, hilbert_subspace='ge', unit_convert=S.CM_FS)
        fo, Xo = P.absorption_spectra(om, tmax, dt=1.0)
        assert np.allclose(f, fo)
        assert rel_err(X[i], Xo) < 1e-10, (i, h.n_sites)
        assert rel_err(Xm[i], Xo) < 1e-6, (i, h.n_sites)
