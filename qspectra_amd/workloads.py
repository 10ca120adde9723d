"""Seeded synthetic systems used by the benchmark and the tests.

All energies in cm^-1, times in fs.  The Drude bath is lambda = 35 cm^-1,
cutoff 106 cm^-1 at 77 K unless stated.
"""
import numpy as np

from .bath import DebyeBath
from .constants import CM_FS, CM_K, GAUSSIAN_SD_FWHM

# unit transition dipoles of the FMO complex (pigments 1-7)
FMO_DIPOLES = np.array([[3.019, 3.442, 0.797, 3.213, 2.969, 0.547, 1.983],
                        [2.284, -2.023, -3.871, 2.145, -2.642, 3.562, 2.837],
                        [1.506, 0.431, 0.853, 1.112, -0.661, -1.851, 2.015]]).T
FMO_DIPOLES = FMO_DIPOLES / np.linalg.norm(FMO_DIPOLES, axis=1, keepdims=True)


def drude_bath(kelvin=77.0):
    return DebyeBath(CM_K * kelvin, 35.0, 106.0)


def spread_extra_for_step(H1, dt):
    """energy_spread_extra making the default model time step ``dt`` fs:
    1 / (2 (max|E_gef| + extra) CM_FS) = dt in the rotating frame at the
    mean one-exciton energy (hamiltonian.py freq_step rule)."""
    H1 = np.asarray(H1, float)
    n = len(H1)
    e = np.linalg.eigvalsh(H1)
    rw = e.mean()
    Ef = np.linalg.eigvalsh(_two_exciton(H1)) if n > 1 else np.zeros(0)
    E = np.concatenate([[0.0], e - rw, Ef - 2 * rw])
    return 1.0 / (2 * dt * CM_FS) - max(E.max(), -E.min())


def _two_exciton(H1):
    from .operator_tools import operator_1_to_2
    return operator_1_to_2(H1)


def dimer_hamiltonian():
    """Dimer at 12881 / 12719 cm^-1, coupling 120 cm^-1, unit dipoles 30
    degrees apart; default time step 1 fs."""
    H1 = np.array([[12881., 120.], [120., 12719.]])
    th = np.deg2rad(30.0)
    d = np.array([[1., 0, 0], [np.cos(th), np.sin(th), 0]])
    return H1, d, spread_extra_for_step(H1, 1.0)


def aggregate7_hamiltonian(seed=1):
    """7-site FMO-like aggregate: site energies U(12200, 12700), symmetric
    couplings N(0, 60), both from RandomState(seed); FMO unit dipoles;
    default time step 2 fs."""
    rs = np.random.RandomState(seed)
    E = rs.uniform(12200, 12700, 7)
    J = rs.normal(0, 60, (7, 7))
    H1 = np.triu(J, 1)
    H1 = H1 + H1.T + np.diag(E)
    return H1, FMO_DIPOLES.copy(), spread_extra_for_step(H1, 2.0)


def aggregate7_model(secular=False, disorder_sigma=None, kelvin=77.0):
    from .dynamics.redfield import RedfieldModel
    from .hamiltonian import ElectronicHamiltonian
    H1, d, extra = aggregate7_hamiltonian()
    disorder = None if disorder_sigma is None else disorder_sigma / GAUSSIAN_SD_FWHM
    ham = ElectronicHamiltonian(H1, bath=drude_bath(kelvin), dipoles=d, disorder=disorder,
                                energy_spread_extra=extra)
    return RedfieldModel(ham, hilbert_subspace='gef', unit_convert=CM_FS, secular=secular)


# the 2-D workload: t1 = t3 = 128 points at 2 fs, four population times
AGG7_2D = dict(coherence_time_max=2 * 127.5, population_times=[0., 100., 200., 500.],
               geometry='-++', include_signal='GSB,ESE,ESA', dt=2.0)


def rk4_steps_2d(n1=128, t2=(0., 100., 200., 500.), n3=128, dt=2.0, pathways=3):
    """Propagated state-steps of one 2-D spectrum: per pathway 1 RHS over
    the t1 grid, n1 RHS over the t2 grid (from t0 = 0), 1 RHS over t3."""
    s2 = int(round(max(t2) / dt))
    return pathways * ((n1 - 1) + n1 * s2 + (n3 - 1))
