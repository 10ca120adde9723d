"""Oracle: system description (constants, operators, polarization, bath, H).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Literal numpy restatement
of the reference's L1 layer; citations are to the qspectra sources.
"""
from itertools import product

import numpy as np
import scipy.linalg

# qspectra/constants.py:8-18
CM_FS_LINEAR = 2.99792458e-5
CM_FS = np.pi * 2 * CM_FS_LINEAR
CM_K = 0.69503476
GAUSSIAN_SD_FWHM = 1.0 / (2 * np.sqrt(2 * np.log(2)))


# ---------------------------------------------------------------- operators
def all_states(N, subspace='gef'):
    """qspectra/operator_tools.py:122-152 (hard-core boson state list)."""
    states = []
    if 'g' in subspace:
        states.append([])
    if 'e' in subspace:
        states.extend([i] for i in range(N))
    if 'f' in subspace:
        states.extend([i, j] for i in range(N) for j in range(i + 1, N))
    return states


def operator_1_to_2(op1):
    """qspectra/operator_tools.py:155-191, double loop over f states."""
    states = all_states(len(op1), 'f')
    out = np.zeros((len(states), len(states)), dtype=np.asarray(op1).dtype)
    for m, (i, j) in enumerate(states):
        for n, (k, l) in enumerate(states):
            out[m, n] = (op1[j, l] * (i == k) + op1[j, k] * (i == l)
                         + op1[i, l] * (j == k) + op1[i, k] * (j == l))
    return out


def operator_extend(op1, subspace='gef'):
    """qspectra/operator_tools.py:194-235 (block-diagonal g/e/f extension)."""
    op1 = np.asarray(op1)
    blocks = []
    if 'g' in subspace:
        blocks.append(np.array([[0]]))
    if 'e' in subspace:
        blocks.append(op1)
    if 'f' in subspace:
        blocks.append(operator_1_to_2(op1))
    return scipy.linalg.block_diag(*blocks).astype(op1.dtype)


def transition_operator(n, n_sites, subspace='gef', include_transitions='-+'):
    """qspectra/operator_tools.py:238-270."""
    states = all_states(n_sites, subspace)
    out = np.zeros((len(states), len(states)))
    for i, si in enumerate(states):
        for j, sj in enumerate(states):
            if (('+' in include_transitions and si == sorted(sj + [n]))
                    or ('-' in include_transitions and sorted(si + [n]) == sj)):
                out[i, j] = 1
    return out


def n_excitations(n_sites=1, n_vibrational_states=1):
    """qspectra/operator_tools.py:279-285."""
    return np.array([1, n_sites, n_sites * (n_sites - 1) // 2]) * n_vibrational_states


def hilbert_subspace_index(subspace, all_subspaces, n_sites, n_vibrational_states=1):
    """qspectra/operator_tools.py:311-335."""
    n_exc = n_excitations(n_sites, n_vibrational_states)
    breaks = [0] + list(np.cumsum(n_exc[['gef'.index(s) for s in all_subspaces]]))
    if subspace not in all_subspaces:
        raise ValueError('%s not in %s' % (subspace, all_subspaces))
    k = all_subspaces.index(subspace)
    return slice(int(breaks[k]), int(breaks[k + 1]))


def extract_subspace(s):
    """qspectra/operator_tools.py:292-298."""
    return sorted(set(s) - {',', '-', '>'}, key=lambda x: 'gef'.index(x))


def full_liouville_subspace(s):
    """qspectra/operator_tools.py:301-308."""
    h = extract_subspace(s)
    return ','.join(a + b for a in h for b in h)


def basis_transform_operator(X, U):
    """qspectra/operator_tools.py:43-87 (kron(U,U) in Liouville space)."""
    X = np.asarray(X)
    U = np.asarray(U)
    if X.shape[-1] == len(U) ** 2:
        U = np.kron(U, U)
    return U.T.conj().dot(X).dot(U)


def basis_transform_vector(rho, U):
    """qspectra/operator_tools.py:90-119."""
    rho = np.asarray(rho)
    U = np.asarray(U)
    if rho.shape[-1] == len(U) ** 2:
        U = np.kron(U, U)
    return np.tensordot(rho, U.T.conj(), axes=(-1, -1))


# ------------------------------------------------------------- polarization
_COORD = {'x': np.array([1, 0, 0]), 'y': np.array([0, 1, 0]),
          'z': np.array([0, 0, 1])}
FOURTH_ORDER_INVARIANTS = [((0, 1), (2, 3)), ((0, 2), (1, 3)), ((0, 3), (1, 2))]
MAGIC_ANGLE = np.arccos(1 / np.sqrt(3))


def polarization_vector(p):
    """qspectra/polarization.py:14-36."""
    if isinstance(p, str):
        return _COORD[p]
    if np.isscalar(p):
        return np.array([np.cos(p), np.sin(p), 0])
    p = np.asarray(p, float).reshape(-1)
    if len(p) != 3:
        raise ValueError('polarization vectors must have length 3')
    return p


def check_polarizations(p, length):
    """qspectra/polarization.py:39-43."""
    pol = np.array([polarization_vector(x) for x in p])
    if len(pol) != length:
        raise ValueError('%s polarizations required' % length)
    return pol


def invariant_weights_4th_order(polarizations):
    """qspectra/polarization.py:55-70 (Hochstrasser weights)."""
    p = check_polarizations(polarizations, 4)
    cos = p.dot(p.T)
    products = [cos[x] * cos[y] for x, y in FOURTH_ORDER_INVARIANTS]
    return (5 * np.eye(3) - np.ones((3, 3))).dot(products) / 30


def invariant_polarizations(invariant):
    """qspectra/polarization.py:73-84."""
    return [''.join(p) for p in product('xyz', repeat=4)
            if all(p[a] == p[b] for a, b in invariant)]


def random_rotation_matrix(random_state):
    """qspectra/polarization.py:87-102 (Arvo)."""
    x1, x2, x3 = random_state.rand(3)
    th, ph = 2 * np.pi * x1, 2 * np.pi * x2
    R = [[np.cos(th), np.sin(th), 0], [-np.sin(th), np.cos(th), 0], [0, 0, 1]]
    v = [np.cos(ph) * np.sqrt(x3), np.sin(ph) * np.sqrt(x3), np.sqrt(1 - x3)]
    Hh = np.identity(3) - 2 * np.outer(v, v)
    return -Hh.dot(R)


# --------------------------------------------------------------------- bath
class DebyeBath(object):
    """qspectra/bath.py:70-102 (+ Bath.corr_func_real :17-31)."""

    def __init__(self, temperature, reorg_energy, cutoff_freq):
        self.temperature = float(temperature)
        self.reorg_energy = float(reorg_energy)
        self.cutoff_freq = float(cutoff_freq)

    def spectral_density_func(self, x):
        return (2 * self.reorg_energy * self.cutoff_freq * x
                / (self.cutoff_freq ** 2 + x ** 2))

    def corr_func_real(self, x):
        """bath.py:17-31."""
        T = self.temperature
        if x == 0:
            return T * 2 * self.reorg_energy / self.cutoff_freq
        Janti = (self.spectral_density_func(x) if x >= 0
                 else -self.spectral_density_func(-x))
        return (1 / np.expm1(x / T) + 1) * Janti

    def corr_func_complex(self, x, matsubara_cutoff=1000):
        """bath.py:84-102 (Matsubara sum truncated at 1000 terms, k=0 term 0)."""
        T, lam, g = self.temperature, self.reorg_energy, self.cutoff_freq
        nu = 2 * np.pi * np.arange(matsubara_cutoff) * T
        if x == 0:
            return lam * (2 * T / g - 1j)
        return lam * g * ((1 / np.tan(g / (2 * T)) - 1j) / (g - 1j * x)
                          + 4 * T * np.sum(nu / ((nu ** 2 - g ** 2) * (nu - 1j * x))))


class ArbitraryBath(object):
    """qspectra/bath.py:50-56 with Bath.corr_func_real :17-31: a bath given
    by a spectral density function J(x) and lim J(x)/x at 0.  Only the real
    correlation function exists (no corr_func_complex, as in the reference)."""

    def __init__(self, temperature, spectral_density_func, spectral_density_limit_at_zero):
        self.temperature = temperature
        self.spectral_density_func = spectral_density_func
        self.spectral_density_limit_at_zero = spectral_density_limit_at_zero

    def corr_func_real(self, x):
        """bath.py:17-31: (n(x) + 1) J_anti(x), T J0 at x = 0."""
        T = self.temperature
        if x == 0:
            return T * self.spectral_density_limit_at_zero
        J = self.spectral_density_func
        Janti = J(x) if x >= 0 else -J(-x)
        return (1 / np.expm1(x / T) + 1) * Janti


# -------------------------------------------------------------- Hamiltonian
def thermal_state(H, temperature):
    """qspectra/hamiltonian.py:49-78 (+ ground_state :25-46)."""
    H = np.asarray(H)
    if temperature > 0:
        rho = scipy.linalg.expm(-H / float(temperature))
        rho = rho / np.trace(rho)
    else:
        E, U = scipy.linalg.eigh(H)
        rho = np.mean([np.outer(U[:, i], U[:, i]) for i in range(len(U))
                       if E[i] == E[0]], axis=0)
    return rho.astype(complex)


class ElectronicHamiltonian(object):
    """qspectra/hamiltonian.py:464-631 restricted to what the path uses.

    ``rw_freq`` / ``sampled_from`` mimic the bookkeeping of
    ``Hamiltonian.in_rotating_frame`` (:177-209) and ``sample`` (:233-274):
    a sampled Hamiltonian keeps its parent's rw_freq and time_step.
    """

    def __init__(self, H_1exc, bath=None, dipoles=None, disorder=None,
                 random_seed=0, energy_spread_extra=None):
        self.H_1exc = np.asarray(H_1exc)
        self.bath = bath
        self.dipoles = None if dipoles is None else np.asarray(dipoles, float)
        self.disorder = disorder
        self.random_seed = random_seed
        self.energy_spread_extra = energy_spread_extra
        self.rw_freq = 0.0
        self.parent = self          # the non-sampled Hamiltonian

    @property
    def n_sites(self):
        return len(self.H_1exc)

    def n_states(self, subspace):
        return int(sum(n_excitations(self.n_sites)[['gef'.index(s) for s in subspace]]))

    def H(self, subspace, rotating=True):
        """hamiltonian.py:539-550: operator_extend of H_1exc - rw*I."""
        H1 = self.H_1exc - (self.rw_freq * np.eye(self.n_sites) if rotating else 0)
        return operator_extend(H1, subspace)

    def in_rotating_frame(self, rw_freq=None):
        """hamiltonian.py:177-209; default rw = mean E_e of the original."""
        if rw_freq is None:
            rw_freq = np.mean(scipy.linalg.eigh(self.parent.H_1exc)[0])
        new = ElectronicHamiltonian(self.H_1exc, self.bath, self.dipoles,
                                    self.disorder, self.random_seed,
                                    self.energy_spread_extra)
        new.rw_freq = rw_freq
        if self.parent is not self:
            new.parent = self.parent.in_rotating_frame(rw_freq)
        return new

    def eig(self, subspace):
        """hamiltonian.py:310-328: eigh of the non-rotating H, then shift."""
        E, U = scipy.linalg.eigh(self.H(subspace, rotating=False))
        if 'e' in subspace:
            E[hilbert_subspace_index('e', subspace, self.n_sites)] -= self.rw_freq
        if 'f' in subspace:
            E[hilbert_subspace_index('f', subspace, self.n_sites)] -= 2 * self.rw_freq
        return E, U

    def E(self, subspace):
        return self.eig(subspace)[0]

    def U(self, subspace):
        return self.eig(subspace)[1]

    @property
    def freq_step(self):
        """hamiltonian.py:382-399 (fixed under sampling)."""
        E = self.parent.E('gef')
        if self.energy_spread_extra is None:
            extra = 0.01 * np.mean(scipy.linalg.eigh(self.parent.H_1exc)[0])
        else:
            extra = self.energy_spread_extra
        return 2 * (max(E.max(), -E.min()) + extra)

    @property
    def time_step(self):
        """hamiltonian.py:401-412."""
        return 1.0 / self.freq_step

    def sample(self, n):
        """hamiltonian.py:233-274 + ElectronicHamiltonian._sample :552-578
        (diagonal Gaussian disorder with sigma = FWHM * GAUSSIAN_SD_FWHM,
        RandomState(list(atleast_1d(seed)) + [n]) of the non-sampled parent)."""
        base = self.parent
        seed = list(np.atleast_1d(base.random_seed)) + [n]
        rs = np.random.RandomState(seed)
        if base.disorder is None:
            dH = 0
        elif callable(base.disorder):
            dH = base.disorder(rs)
        else:
            dH = np.diag(base.disorder * GAUSSIAN_SD_FWHM * rs.randn(base.n_sites))
        new = ElectronicHamiltonian(base.H_1exc + dH, base.bath, base.dipoles,
                                    base.disorder, seed, base.energy_spread_extra)
        new.rw_freq = self.rw_freq
        new.parent = base
        return new

    def dipole_operator(self, subspace='gef', polarization='x', transitions='-+'):
        """hamiltonian.py:580-591."""
        ops = [transition_operator(n, self.n_sites, subspace, transitions)
               for n in range(self.n_sites)]
        return np.einsum('nij,nk,k->ij', ops, self.dipoles,
                         polarization_vector(polarization))

    def number_operator(self, site, subspace='gef'):
        """hamiltonian.py:593-598."""
        d = np.zeros(self.n_sites)
        d[site] = 1
        return operator_extend(np.diag(d), subspace)

    def system_bath_couplings(self, subspace='gef'):
        """hamiltonian.py:600-608."""
        return np.array([self.number_operator(n, subspace) for n in range(self.n_sites)])

    def thermal_state(self, subspace):
        """hamiltonian.py:163-175 on the non-rotating H."""
        T = getattr(self.bath, 'temperature', 0) if self.bath is not None else 0
        return thermal_state(self.H(subspace, rotating=False), T)


class VibronicHamiltonian(object):
    """qspectra/hamiltonian.py:633-807: an electronic Hamiltonian (this
    module's) extended by harmonic modes, on the product basis |e>|v>
    (electronic state outer, mode occupations inner, the last mode fastest,
    the order of operator_tools.py:5-31 tensor / extend_vib_operator).

    The matrix is written element by element over that basis:
    <e v|H|e' v'> = H_el[e, e'] d(v, v') + d(e, e') d(v, v') sum_m w_m v_m
                    + d(e, e') sum_{n in e, m} c_nm <v|b_m + b_m^+|v'>,
    with <v|b_m + b_m^+|v'> = sqrt(max(v_m, v'_m)) when v and v' differ
    only in mode m, by one quantum.  The rotating frame, ensemble samples
    and bath are those of the electronic part (:741-751)."""

    def __init__(self, electronic, n_vibrational_levels, vib_energies, elec_vib_couplings,
                 energy_spread_extra=None):
        self.electronic = electronic
        self.bath = electronic.bath
        self.levels = [int(x) for x in np.atleast_1d(n_vibrational_levels)]
        self.vib_energies = np.asarray(vib_energies, float)
        self.couplings = np.asarray(elec_vib_couplings, float)
        self.energy_spread_extra = energy_spread_extra
        self.rw_freq = electronic.rw_freq
        self.parent = self

    @property
    def n_sites(self):
        return self.electronic.n_sites

    @property
    def n_vibrational_states(self):
        return int(np.prod(self.levels))

    def vib_states(self):
        return list(product(*[range(k) for k in self.levels]))

    def n_states(self, subspace):
        return self.electronic.n_states(subspace) * self.n_vibrational_states

    def H(self, subspace, rotating=True):
        Hel = self.electronic.H(subspace, rotating)
        el = all_states(self.n_sites, subspace)
        vib = self.vib_states()
        basis = [(i, v) for i in range(len(el)) for v in vib]
        # complex when the electronic part is (the reference's tensor products keep it)
        H = np.zeros((len(basis), len(basis)), dtype=np.result_type(Hel, float))
        for r, (i, v) in enumerate(basis):
            for c, (j, u) in enumerate(basis):
                h = Hel[i, j] if v == u else 0.0
                if i == j:
                    if v == u:
                        h += float(np.dot(self.vib_energies, v))
                    diff = [m for m in range(len(v)) if v[m] != u[m]]
                    if len(diff) == 1 and abs(v[diff[0]] - u[diff[0]]) == 1:
                        m = diff[0]
                        q = np.sqrt(max(v[m], u[m]))
                        h += sum(self.couplings[n, m] for n in el[i]) * q
                H[r, c] = h
        return H

    def _extend(self, op):
        """operator on the electronic states -> same operator x identity on
        the modes (hamiltonian.py:763-768), element by element."""
        op = np.asarray(op)
        nv = self.n_vibrational_states
        out = np.zeros((op.shape[0] * nv, op.shape[1] * nv), dtype=op.dtype)
        for k in range(nv):
            out[k::nv, k::nv] = op
        return out

    def in_rotating_frame(self, rw_freq=None):
        """hamiltonian.py:177-209 with :741-745; default rw = mean one-exciton
        eigenvalue of the full vibronic Hamiltonian of the original."""
        if rw_freq is None:
            rw_freq = float(np.mean(scipy.linalg.eigh(self.parent.H('e', rotating=False))[0]))
        new = VibronicHamiltonian(self.electronic.in_rotating_frame(rw_freq), self.levels,
                                  self.vib_energies, self.couplings, self.energy_spread_extra)
        if self.parent is not self:
            new.parent = self.parent.in_rotating_frame(rw_freq)
        return new

    def eig(self, subspace):
        """hamiltonian.py:310-328 on the vibronic matrix."""
        E, U = scipy.linalg.eigh(self.H(subspace, rotating=False))
        nv = self.n_vibrational_states
        if 'e' in subspace:
            E[hilbert_subspace_index('e', subspace, self.n_sites, nv)] -= self.rw_freq
        if 'f' in subspace:
            E[hilbert_subspace_index('f', subspace, self.n_sites, nv)] -= 2 * self.rw_freq
        return E, U

    def E(self, subspace):
        return self.eig(subspace)[0]

    def U(self, subspace):
        return self.eig(subspace)[1]

    @property
    def transition_energy(self):
        """hamiltonian.py:373-380 on the original (non-rotating) vibronic H."""
        return float(np.mean(scipy.linalg.eigh(self.parent.H('e', rotating=False))[0]))

    @property
    def freq_step(self):
        """hamiltonian.py:382-399 (fixed under sampling)."""
        E = self.parent.E('gef')
        extra = (0.01 * self.transition_energy if self.energy_spread_extra is None
                 else self.energy_spread_extra)
        return 2 * (max(E.max(), -E.min()) + extra)

    @property
    def time_step(self):
        return 1.0 / self.freq_step

    def sample(self, n):
        """hamiltonian.py:233-274 with :747-751: the electronic part sampled."""
        base = self.parent
        new = VibronicHamiltonian(self.electronic.sample(n), self.levels, self.vib_energies,
                                  self.couplings, self.energy_spread_extra)
        new.parent = base
        return new

    def dipole_operator(self, subspace='gef', polarization='x', transitions='-+'):
        return self._extend(self.electronic.dipole_operator(subspace, polarization, transitions))

    def number_operator(self, site, subspace='gef'):
        return self._extend(self.electronic.number_operator(site, subspace))

    def system_bath_couplings(self, subspace='gef'):
        return np.array([self._extend(k)
                         for k in self.electronic.system_bath_couplings(subspace)])

    def thermal_state(self, subspace):
        """hamiltonian.py:163-175 on the non-rotating vibronic H."""
        T = getattr(self.bath, 'temperature', 0) if self.bath is not None else 0
        return thermal_state(self.H(subspace, rotating=False), T)


# --------------------------------------------------------------------- pulses
class CustomPulse(object):
    """pulse.py:51-71: field given by call(t, rw_freq)."""

    def __init__(self, t_init, t_final, call):
        self.t_init, self.t_final, self.call = t_init, t_final, call

    def __call__(self, t, rw_freq):
        return self.call(t, rw_freq)


class GaussianPulse(object):
    """pulse.py:74-114: scale * exp(i c (w_c - rw) (t - t_p) - (t - t_p)^2 / (2 s^2)),
    s = GAUSSIAN_SD_FWHM * fwhm, limits t_p -/+ t_limits_multiple * s."""

    def __init__(self, carrier_freq, fwhm, t_peak=0, scale=1, freq_convert=1,
                 t_limits_multiple=3):
        sigma = GAUSSIAN_SD_FWHM * fwhm
        self.two_sigma_squared = 2 * sigma ** 2
        self.t_init = t_peak - t_limits_multiple * sigma
        self.t_final = t_peak + t_limits_multiple * sigma
        self.t_peak, self.carrier_freq = t_peak, carrier_freq
        self.scale, self.freq_convert = scale, freq_convert

    def __call__(self, t, rw_freq):
        return self.scale * np.exp(1j * self.freq_convert * (self.carrier_freq - rw_freq)
                                   * (t - self.t_peak)
                                   - (t - self.t_peak) ** 2 / self.two_sigma_squared)
