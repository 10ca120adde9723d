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


# the 2-D workload: t1 = t3 = 128 points at 2 fs (the model's time step),
# four population times.  The RK4 step is 1 fs: two RK4 steps per 2 fs grid
# interval put the spectrum within 3e-7 of the reference's tight-tolerance
# zvode result (RK4 at 2 fs: 4.4e-6; the reference at its default zvode
# tolerance: 1.7e-6), inside the 1e-6 the metric is quoted at.
AGG7_2D = dict(coherence_time_max=2 * 127.5, population_times=[0., 100., 200., 500.],
               geometry='-++', include_signal='GSB,ESE,ESA', dt=1.0)
AGG7_2D_GRID_STEP = 2.0


def rk4_steps_2d(n1=128, t2=(0., 100., 200., 500.), n3=128, dt=2.0, pathways=3):
    """Propagated state-steps of one 2-D spectrum at step dt: per pathway 1
    RHS over the t1 grid, n1 RHS over the t2 grid (from t0 = 0), 1 RHS over
    t3 (with the 2 fs grid spacing and dt = 2: 96,762)."""
    q = int(round(AGG7_2D_GRID_STEP / dt))
    s2 = int(round(max(t2) / dt))
    return pathways * (q * (n1 - 1) + n1 * s2 + q * (n3 - 1))


def ensemble_absorption_hamiltonians(n_disorder=4096, n_aggregates=1024, sigma=100.0):
    """Absorption ensemble: n_disorder static-disorder members (diagonal
    Gaussian, standard deviation sigma) of the 7-site aggregate, drawn as
    ElectronicHamiltonian.sample_ensemble draws them (member n from
    RandomState([0, n])), plus n_aggregates aggregates of N ~ U{2..24}
    sites with E ~ U(12200, 12700), symmetric J ~ N(0, 80) and random unit
    dipoles, all from RandomState(2).  Returns ElectronicHamiltonians."""
    from .hamiltonian import ElectronicHamiltonian
    H1, d, _ = aggregate7_hamiltonian()
    parent = ElectronicHamiltonian(H1, bath=drude_bath(), dipoles=d,
                                   disorder=sigma / GAUSSIAN_SD_FWHM)
    hams = list(parent.sample_ensemble(n_disorder)) if n_disorder else []
    rs = np.random.RandomState(2)
    for _ in range(n_aggregates):
        n = rs.randint(2, 25)
        E = rs.uniform(12200, 12700, n)
        J = rs.normal(0, 80, (n, n))
        H = np.triu(J, 1)
        H = H + H.T + np.diag(E)
        dip = rs.normal(size=(n, 3))
        dip /= np.linalg.norm(dip, axis=1, keepdims=True)
        hams.append(ElectronicHamiltonian(H, bath=drude_bath(), dipoles=dip))
    return hams


def gradient_target_parameters(seed=3):
    """Perturbed parameters of the 7-site aggregate for the gradient
    workload: site energies +-50, couplings +-10, lambda +-5 cm^-1 (uniform,
    RandomState(seed)).  Returns (E, J upper triangle, lam) for the
    unperturbed and the perturbed system."""
    H1, _, _ = aggregate7_hamiltonian()
    iu = np.triu_indices(7, 1)
    E0, J0, lam0 = np.diag(H1).copy(), H1[iu].copy(), 35.0
    rs = np.random.RandomState(seed)
    return ((E0, J0, lam0),
            (E0 + rs.uniform(-50, 50, 7), J0 + rs.uniform(-10, 10, len(J0)),
             lam0 + rs.uniform(-5, 5)))


def heom_aggregate7_model(kelvin=300.0, level_cutoff=5):
    """HEOM model of the 7-site aggregate: Drude bath at 300 K, K = 1
    Matsubara term, hierarchy depth 4 (level_cutoff 5 -> 3060 ADOs), 'ge'
    Hilbert space, disorder standard deviation 100 cm^-1."""
    from .dynamics.heom import HEOMModel
    from .hamiltonian import ElectronicHamiltonian
    H1, d, _ = aggregate7_hamiltonian()
    ham = ElectronicHamiltonian(H1, bath=drude_bath(kelvin), dipoles=d,
                                disorder=100.0 / GAUSSIAN_SD_FWHM)
    return HEOMModel(ham, hilbert_subspace='ge', unit_convert=CM_FS, level_cutoff=level_cutoff,
                     K=1)


def example_dimer_models():
    """The electronic dimer of the reference's example comparing HEOM,
    Redfield and ZOFE dynamics: site energies 12881 / 12719 cm^-1, coupling
    120 cm^-1, dipoles (1, 0, 0) and (2 cos 0.3, 2 sin 0.3, 0), Drude bath
    (35 cm^-1, 106 cm^-1, 77 K); Redfield (imaginary part of C discarded)
    and HEOM (level_cutoff 3, K = 1, no low-temperature correction) models
    on the g/e/f space."""
    from .dynamics.heom import HEOMModel
    from .dynamics.redfield import RedfieldModel
    from .hamiltonian import ElectronicHamiltonian
    ham = ElectronicHamiltonian(np.array([[12881., 120.], [120., 12719.]]), bath=drude_bath(77.0),
                                dipoles=[[1., 0., 0.], [2. * np.cos(.3), 2. * np.sin(.3), 0.]])
    red = RedfieldModel(ham, hilbert_subspace='gef', discard_imag_corr=True, unit_convert=CM_FS)
    heom = HEOMModel(ham, hilbert_subspace='gef', unit_convert=CM_FS, level_cutoff=3,
                     low_temp_corr=False)
    return red, heom


# the example's 2-D spectra: 1 ps coherence times, 50 population times to 1 ps
EXAMPLE_DIMER_2D = dict(coherence_time_max=1000, population_times=np.linspace(0, 1000, 50),
                        geometry='-++', polarization='xxxx', include_signal='GSB,ESE,ESA')


# the FMO complex of the reference's ZOFE example: one-exciton Hamiltonian
# (cm^-1) and the 16-pseudomode fit of the 77 K Drude spectral density
FMO_H1 = np.array([[12400, -87.7, 5.5, -5.9, 6.7, -13.7, -9.9],
                   [-87.7, 12520, 30.8, 8.2, 0.7, 11.8, 4.3],
                   [5.5, 30.8, 12200, -53.5, -2.2, -9.6, 6.],
                   [-5.9, 8.2, -53.5, 12310, -70.7, -17., -63.3],
                   [6.7, 0.7, -2.2, -70.7, 12470, 81.1, -1.3],
                   [-13.7, 11.8, -9.6, -17., 81.1, 12620, 39.7],
                   [-9.9, 4.3, 6., -63.3, -1.3, 39.7, 12430]])
PM_OMEGA = [-500., -200., -90., 1., 21., 60., 80., 130., 200., 300., 400., 500., 600., 800.,
            1100., 1500.]
PM_HUANG = [-2.5133e-03, -7.5398e-03, -2.5133e-02, 5.0265e+01, 2.2619e+00, 4.5239e-02,
            2.7646e-01, 9.2991e-03, 2.2619e-02, 1.5080e-02, 3.0159e-03, 3.5186e-03, 2.8274e-04,
            1.7593e-03, 4.3982e-04, 4.3982e-04]
PM_GAMMA = [500., 100., 50., 50., 50., 50., 80., 40., 80., 150., 200., 200., 80., 250., 200.,
            300.]


def pseudomode_bath(n_sites):
    """The 16 pseudomodes, identical for every site (arrays [16, n_sites])."""
    from .bath import PseudomodeBath
    on = np.ones(n_sites, complex)
    Om, ga, hu = [np.array([v[pm] * on for pm in range(len(v))])
                  for v in (PM_OMEGA, PM_GAMMA, PM_HUANG)]
    return PseudomodeBath(len(PM_OMEGA), Om, ga, hu)


def fmo_zofe_model(hilbert_subspace='e'):
    from .dynamics.zofe import ZOFEModel
    from .hamiltonian import ElectronicHamiltonian
    ham = ElectronicHamiltonian(FMO_H1, bath=pseudomode_bath(7), dipoles=FMO_DIPOLES)
    return ZOFEModel(ham, hilbert_subspace=hilbert_subspace, unit_convert=CM_FS)


def jonas_models():
    """The vibronic examples of the reference: the Jonas dimer (two sites at
    11500 / 11650 cm^-1, coupling 66, one two-level mode of 200 cm^-1 per
    site with coupling -32, Drude bath 77 K, lambda 52.5, gamma 106;
    Redfield, real C(w), g/e/f), its monomer (Redfield, complex C(w), g/e)
    and the unitary five-level vibronic monomer (11500, mode 200,
    coupling -100, g/e)."""
    from . import ElectronicHamiltonian, RedfieldModel, UnitaryModel, VibronicHamiltonian
    bath = DebyeBath(CM_K * 77, 1.5 * 35, 106)
    H_el = np.array([[11500., 66.], [66., 11650.]])
    dipoles = [[1., 0., 0.], [0., 1., 0.]]
    ham = VibronicHamiltonian(ElectronicHamiltonian(H_el, dipoles=dipoles, bath=bath),
                              [2, 2], [200., 200.], -32. * np.eye(2))
    dimer = RedfieldModel(ham, hilbert_subspace='gef', unit_convert=CM_FS,
                          discard_imag_corr=True)
    mono = VibronicHamiltonian(ElectronicHamiltonian(H_el[:1, :1], dipoles=dipoles[:1],
                                                     bath=bath),
                               [2], [200.], [[-32.]])
    monomer = RedfieldModel(mono, hilbert_subspace='ge', unit_convert=CM_FS)
    uham = VibronicHamiltonian(ElectronicHamiltonian([[11500.]], dipoles=[[1., 0., 0.]]),
                               [5], [200.], [[-100.]])
    unitary = UnitaryModel(uham, hilbert_subspace='ge', unit_convert=CM_FS)
    return dimer, monomer, unitary


JONAS_GSB = dict(coherence_time_max=1000, population_times=np.linspace(0, 1000, num=100),
                 geometry='-++', polarization='xxxx', include_signal='GSB',
                 exact_isotropic_average=True)
