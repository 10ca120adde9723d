"""Batched absorption spectra of aggregates of uneven size against the oracle,
member by member (response.py:103-154 on each Hamiltonian)."""
import numpy as np
import pytest

import qspectra_amd as qa
from oracle import system as S
from oracle import liouville as Lv
from oracle import propagate as P
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

BATH = (S.CM_K * 77, 35.0, 106.0)


class _FixedStep(Lv.LiouvilleModel):
    time_step = 1.0


def _aggregates(sizes, seed=2):
    rng = np.random.RandomState(seed)
    out = []
    for n in sizes:
        J = rng.normal(0.0, 80.0, (n, n))
        H = (J + J.T) / 2
        H[np.diag_indices(n)] = 12500.0 + rng.normal(0.0, 100.0, n)
        out.append((H, rng.normal(size=(n, 3))))
    return out


def test_absorption_batch_uneven_sizes(gpu):
    from qspectra_amd.batched import absorption_batch
    systems = _aggregates([2, 3, 2, 5, 3, 4, 2])
    hams = [qa.ElectronicHamiltonian(H, bath=qa.DebyeBath(*BATH), dipoles=d)
            for H, d in systems]
    f, X = absorption_batch(hams, 512, rw_freq=12500.0, dt=1.0)
    fm, Xm = absorption_batch(hams, 512, rw_freq=12500.0, dt=1.0, precision='mixed')
    assert X.shape == (len(hams), len(f)) and np.allclose(f, fm)
    for i, (H, d) in enumerate(systems):
        om = _FixedStep(S.ElectronicHamiltonian(H, bath=S.DebyeBath(*BATH), dipoles=d),
                        rw_freq=12500.0, hilbert_subspace='ge', unit_convert=S.CM_FS)
        fo, Xo = P.absorption_spectra(om, 512, dt=1.0)
        assert np.allclose(f, fo)
        assert rel_err(X[i], Xo) < 1e-10, i
        assert rel_err(Xm[i], Xo) < 1e-6, i


def test_absorption_batch_disorder_members(gpu):
    """Static-disorder members evaluated in one batch equal the ensemble
    average of the drop-in API member by member."""
    from qspectra_amd.batched import absorption_batch
    H, d = _aggregates([4], seed=5)[0]
    ham = qa.ElectronicHamiltonian(H, bath=qa.DebyeBath(*BATH), dipoles=d, disorder=100,
                                   random_seed=7)
    members = list(ham.sample_ensemble(6))
    f, X = absorption_batch(members, 256, rw_freq=12500.0, dt=1.0)
    for i, h in enumerate(members[:3]):
        om = _FixedStep(S.ElectronicHamiltonian(h.H_1exc, bath=S.DebyeBath(*BATH),
                                                dipoles=h.dipoles),
                        rw_freq=12500.0, hilbert_subspace='ge',
                        unit_convert=S.CM_FS)
        _, Xo = P.absorption_spectra(om, 256, dt=1.0)
        assert rel_err(X[i], Xo) < 1e-10, i
