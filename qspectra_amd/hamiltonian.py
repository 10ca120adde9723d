"""System Hamiltonians with the API of qspectra/hamiltonian.py
(Hamiltonian, ElectronicHamiltonian).

Eigensystems come from the batched FP64 Jacobi kernel (qsx_block_eigh), one
excitation manifold at a time; the thermal state is built from them.  The
rotating-frame / sampling bookkeeping follows hamiltonian.py:96-274: a
sampled Hamiltonian keeps the rotating-wave frequency and time step of its
parent, and ``sample`` and ``in_rotating_frame`` commute.
"""
from abc import ABCMeta, abstractmethod
from numbers import Number
import warnings

import numpy as np
import torch

from . import _native
from .constants import GAUSSIAN_SD_FWHM
from .operator_tools import (operator_extend, transition_operators, unit_vec,
                             hilbert_subspace_index, basis_transform_vector,
                             basis_transform_operator, SubspaceError, tensor, vib_annihilate,
                             vib_create, extend_vib_operator)
from .polarization import polarization_vector, random_rotation_matrix
from .utils import imemoize, check_random_state, inspect_repr


class HamiltonianError(Exception):
    """Missing Hamiltonian data (dipoles, bath)."""


def check_hermitian(matrix):
    matrix = np.asarray(matrix)
    if not np.allclose(matrix.conj().T, matrix):
        raise ValueError('matrix input must to be Hermitian')
    return matrix


def _state_from_eig(E, U, temperature):
    """Thermal (T > 0) or ground (T <= 0) density matrix from an
    eigensystem: U diag(w) U^T / sum(w), with w = exp(-(E - E_min)/T), or the
    mean projector on the lowest level (hamiltonian.py:25-78)."""
    if temperature > 0:
        w = np.exp(-(E - E.min()) / float(temperature))
    else:
        w = (E == E.min()).astype(float)
    rho = (U * (w / w.sum())) @ U.conj().T
    return rho.astype(complex)


def electronic_thermal_states(hamiltonians, subspace):
    """thermal_state(subspace) of several ElectronicHamiltonians at once,
    [B, N, N]: one batched eigensolver launch per manifold of the
    non-rotating one-exciton Hamiltonians, then the Boltzmann (T > 0) or
    ground-state (T <= 0) weights of _state_from_eig per member."""
    hams = list(hamiltonians)
    n = hams[0].n_sites
    dev = _native.device()
    H1 = torch.as_tensor(np.stack([h.H1_nonrotating for h in hams])).to(dev)
    rw = torch.zeros(len(hams), dtype=torch.float64, device=dev)
    Es, Us = [], []
    for s in 'gef':
        X = 'gef'.index(s)
        if s not in subspace or _native.block_size(X, n) == 0:
            continue
        E, U = _native.block_eigh(H1, X, rw)
        Es.append(E.cpu().numpy())
        Us.append(U.cpu().numpy())
    E = np.concatenate(Es, axis=1)                                  # [B, N]
    B, N = E.shape
    U = np.zeros((B, N, N))
    k = 0
    for u in Us:
        U[:, k:k + u.shape[1], k:k + u.shape[1]] = u
        k += u.shape[1]
    T = np.array([getattr(getattr(h._not_rotating, 'bath', None), 'temperature', 0) or 0
                  for h in hams], dtype=float)
    dE = E - E.min(axis=1, keepdims=True)
    hot = T > 0
    w = np.where(hot[:, None], np.exp(-dE / np.where(hot, T, 1.0)[:, None]), dE == 0)
    w = w / w.sum(axis=1, keepdims=True)
    return np.einsum('bij,bj,bkj->bik', U, w, U).astype(complex)


def add_braket(labels):
    out = []
    for label in labels:
        if isinstance(label, str) or not np.iterable(label):
            out.append('|{}>'.format(label))
        else:
            out.append(''.join('|{}>'.format(i) for i in label))
    return out


class Hamiltonian(metaclass=ABCMeta):
    """Base class; subclasses provide ``H(subspace)``."""

    _implements_eq = False

    def __init__(self, energy_spread_extra=None, site_labels=None):
        self.energy_spread_extra = energy_spread_extra
        self._not_sampled = self
        self._not_rotating = self
        self.rw_freq = 0
        self.site_labels = site_labels

    @property
    def _original(self):
        return self._not_rotating._not_sampled

    def __repr__(self):
        return inspect_repr(self)

    def __eq__(self, other):
        return self._eq(other, max_depth=1)

    def __ne__(self, other):
        return not self == other

    __hash__ = object.__hash__

    def _eq(self, other, max_depth):
        if not self._implements_eq:
            return self is other
        return (not max_depth
                or (self.rw_freq == other.rw_freq and self.site_labels == other.site_labels
                    and self._not_sampled._eq(other._not_sampled, max_depth - 1)
                    and self._not_rotating._eq(other._not_rotating, max_depth - 1)))

    @abstractmethod
    def H(self, subspace):
        """System Hamiltonian matrix on the given Hilbert subspace."""

    def n_states(self, subspace):
        return len(self.H(subspace))

    @imemoize
    def in_rotating_frame(self, rw_freq=None):
        """Copy shifted to the rotating frame at rw_freq (default: the mean
        one-exciton energy of the original Hamiltonian); idempotent."""
        if rw_freq is None:
            rw_freq = self._original.transition_energy
        ham = self._not_rotating._in_rotating_frame(rw_freq)
        ham._not_rotating = self._not_rotating
        if self._not_sampled is not self:
            ham._not_sampled = self._not_sampled.in_rotating_frame(rw_freq)
        ham.rw_freq = rw_freq
        return ham

    def _in_rotating_frame(self, rw_freq):
        raise NotImplementedError('%s does not implement rotating frame transformations'
                                  % type(self).__name__)

    def sample_ensemble(self, ensemble_size=1, random_orientations=False):
        for n in range(ensemble_size):
            yield self.sample(n, random_orientations)

    def sample(self, n=None, random_orientations=False):
        """The n-th static-disorder sample (deterministic in n); it keeps this
        Hamiltonian's rotating-wave frequency and time step."""
        if n is None:
            n = np.random.randint(2 ** 30)
        ham = self._not_sampled._sample(n, random_orientations)
        if self._not_rotating is not self:
            ham._not_rotating = self._not_rotating.sample(n, random_orientations)
        ham._not_sampled = self._not_sampled
        ham.rw_freq = self.rw_freq
        return ham

    def _sample(self, n, random_orientations):
        raise NotImplementedError('%s does not implement ensemble sampling' % type(self).__name__)

    def dipole_operator(self, subspace='gef', polarization='x', transitions='-+'):
        raise NotImplementedError('%s does not implement dipole operators' % type(self).__name__)

    def system_bath_couplings(self, subspace='gef'):
        raise NotImplementedError('%s does not implement system-bath couplings'
                                  % type(self).__name__)

    @imemoize
    def eig(self, subspace):
        """(E, U) of the non-rotating Hamiltonian on ``subspace``, with the e
        and f energies then shifted by -rw and -2 rw; the order is that of a
        dense eigensolver on the lab-frame block-diagonal matrix (ascending
        lab-frame energy), as hamiltonian.py:310-328 returns it."""
        E, U, excitations = self._eig_blocks(subspace)
        order = np.argsort(E + excitations * self.rw_freq, kind='stable')
        return E[order], U[:, order]

    def _eig_blocks(self, subspace):
        raise NotImplementedError

    def E(self, subspace):
        return self.eig(subspace)[0]

    def U(self, subspace):
        return self.eig(subspace)[1]

    def transform_vector_to_eigenbasis(self, rho, subspace):
        return basis_transform_vector(rho, self.U(subspace))

    def transform_vector_from_eigenbasis(self, rho, subspace):
        return basis_transform_vector(rho, self.U(subspace).T.conj())

    def transform_operator_to_eigenbasis(self, op, subspace):
        return basis_transform_operator(op, self.U(subspace))

    def transform_operator_from_eigenbasis(self, op, subspace):
        return basis_transform_operator(op, self.U(subspace).T.conj())

    @property
    def transition_energy(self):
        """Mean one-exciton energy."""
        return float(np.mean(self.E('e')))

    @property
    def freq_step(self):
        """Nyquist sampling frequency covering every energy of the
        non-sampled Hamiltonian plus ``energy_spread_extra`` (default 1% of
        the original transition energy); fixed under sampling."""
        E = self._not_sampled.E('gef')
        extra = (0.01 * self._original.transition_energy if self.energy_spread_extra is None
                 else self.energy_spread_extra)
        return 2 * (max(E.max(), -E.min()) + extra)

    @property
    def time_step(self):
        return 1.0 / self.freq_step

    @imemoize
    def ground_state(self, subspace):
        E, U = self._not_rotating.eig(subspace)
        return _state_from_eig(E, U, 0)

    @imemoize
    def thermal_state(self, subspace):
        """Boltzmann state of the non-rotating Hamiltonian at the bath
        temperature (ground state without a bath temperature)."""
        try:
            T = self._not_rotating.bath.temperature
        except AttributeError:
            T = 0
        E, U = self._not_rotating.eig(subspace)
        return _state_from_eig(E, U, T)

    def basis_labels(self, subspace, braket=False):
        return add_braket(self.site_labels) if braket else self.site_labels

    def hilbert_subspace_index(self, subspace, all_subspaces):
        return hilbert_subspace_index(subspace, all_subspaces, self.n_sites,
                                      self.n_vibrational_states)


def diagonal_gaussian_disorder(fwhm, n_sites):
    def disorder(random_state):
        return np.diag((fwhm * GAUSSIAN_SD_FWHM) * random_state.randn(n_sites))
    return disorder


class ElectronicHamiltonian(Hamiltonian):
    """Aggregate of two-level sites with one-exciton Hamiltonian H_1exc, an
    identical independent bath per site, transition dipoles [n, 3] and
    diagonal static disorder given as a FWHM (or a function of a
    RandomState returning an n x n disorder matrix)."""

    _implements_eq = True

    def __init__(self, H_1exc, bath=None, dipoles=None, disorder=None, random_seed=0,
                 energy_spread_extra=None, site_labels=None):
        self.H_1exc = check_hermitian(H_1exc)
        self.bath = bath
        self.dipoles = np.asarray(dipoles) if dipoles is not None else None
        self.disorder = disorder
        self.random_seed = random_seed
        self.n_vibrational_states = 1
        super(ElectronicHamiltonian, self).__init__(energy_spread_extra, site_labels)

    def _eq(self, other, max_depth):
        return (isinstance(other, ElectronicHamiltonian)
                and np.array_equal(self.H_1exc, other.H_1exc) and self.bath == other.bath
                and np.all(self.dipoles == other.dipoles) and self.disorder == other.disorder
                and np.all(np.asarray(self.random_seed) == np.asarray(other.random_seed))
                and self.energy_spread_extra == other.energy_spread_extra
                and super(ElectronicHamiltonian, self)._eq(other, max_depth))

    __hash__ = object.__hash__

    @property
    def n_sites(self):
        return len(self.H_1exc)

    def n_states(self, subspace):
        """len(H(subspace)): each manifold named in ``subspace`` counts once
        (operator_extend builds one block per manifold)."""
        sizes = (1, self.n_sites, self.n_sites * (self.n_sites - 1) // 2)
        return sum(size for s, size in zip('gef', sizes) if s in subspace)

    @imemoize
    def H(self, subspace):
        return operator_extend(self.H_1exc, subspace)

    def _in_rotating_frame(self, rw_freq):
        return type(self)(self.H_1exc - rw_freq * np.identity(self.n_sites), self.bath,
                          self.dipoles, self.disorder, self.random_seed,
                          self.energy_spread_extra, self.site_labels)

    def _sample(self, n, random_orientations):
        if self.disorder is None:
            if not random_orientations:
                warnings.warn('called sample with `disorder=None` and `random_orientations='
                              'False`: sampled Hamiltonian is identical to original',
                              RuntimeWarning, stacklevel=3)
            disorder = lambda rs: 0
        elif isinstance(self.disorder, Number):
            disorder = diagonal_gaussian_disorder(self.disorder, self.n_sites)
        else:
            disorder = self.disorder
        seed = list(np.atleast_1d(self.random_seed)) + [n]
        rs = check_random_state(seed)
        H_1exc = self.H_1exc + disorder(rs)
        dipoles = self.dipoles
        if random_orientations:
            dipoles = self.dipoles @ random_rotation_matrix(rs).T
        return type(self)(H_1exc, self.bath, dipoles, self.disorder, seed,
                          self.energy_spread_extra, self.site_labels)

    # ---------------------------------------------------------- device data
    @property
    def H1_nonrotating(self):
        """One-exciton Hamiltonian in the lab frame (with this sample's
        disorder), the input of the GPU eigen- and Liouvillian builders."""
        return np.asarray(self._not_rotating.H_1exc, dtype=float)

    def _eig_blocks(self, subspace):
        if np.iscomplexobj(self.H_1exc) and np.any(np.imag(self.H_1exc) != 0):
            raise NotImplementedError('complex Hermitian H_1exc: the GPU eigensolver handles '
                                      'real symmetric Hamiltonians')
        dev = _native.device()
        H1 = torch.as_tensor(self.H1_nonrotating[None]).to(dev)
        rw = torch.tensor([float(self.rw_freq)], dtype=torch.float64, device=dev)
        Es, Us, xs = [], [], []
        for s in 'gef':
            if s not in subspace:
                continue
            X = 'gef'.index(s)
            if _native.block_size(X, self.n_sites) == 0:
                continue
            E, U = _native.block_eigh(H1, X, rw)
            Es.append(E[0].cpu().numpy())
            Us.append(U[0].cpu().numpy())
            xs.append(np.full(len(Es[-1]), X))
        E = np.concatenate(Es)
        U = np.zeros((len(E), len(E)))
        k = 0
        for u in Us:
            U[k:k + len(u), k:k + len(u)] = u
            k += len(u)
        return E, U, np.concatenate(xs)

    # --------------------------------------------------------- operators
    def dipole_operator(self, subspace='gef', polarization='x', transitions='-+'):
        if self.dipoles is None:
            raise HamiltonianError('transition dipole moments undefined')
        mu = self.dipoles @ polarization_vector(polarization)
        return np.tensordot(mu, transition_operators(self.n_sites, subspace, transitions),
                            axes=1)

    def number_operator(self, site, subspace='gef'):
        return operator_extend(np.diag(unit_vec(site, self.n_sites, dtype=float)), subspace)

    def system_bath_couplings(self, subspace='gef'):
        if self.bath is None:
            raise HamiltonianError('bath undefined')
        return np.array([self.number_operator(n, subspace) for n in range(self.n_sites)])

    def basis_labels(self, subspace='gef', braket=False):
        labels = self._fock_labels(subspace, self.site_labels)
        return add_braket(labels) if braket else labels

    def _fock_labels(self, subspace, labels=None):
        digits = np.array([10 ** (self.n_sites - i - 1) for i in range(self.n_sites)], dtype='O')
        full = np.diag(operator_extend(np.diag(digits), subspace))
        fock = [str(i).zfill(self.n_sites) for i in full]
        if labels is None:
            return fock
        custom = [','.join(lab for i, lab in enumerate(labels) if st[i] == '1') for st in fock]
        if 'g' in subspace:
            custom[0] = 'g'
        return custom


class VibronicHamiltonian(Hamiltonian):
    """An electronic Hamiltonian extended by explicit vibrations
    (qspectra/hamiltonian.py:633-807): H = H_el x 1 + 1 x H_vib
    + sum_{n,m} c_nm |n><n| x (b_m + b_m^+), on the product of the
    electronic states and the Fock states of the modes.  Its Liouvillian
    blocks are built on the GPU from the dense Hilbert blocks
    (qsx_redfield_block_dense / qsx_unitary_block_dense), its eigensystems by
    qsx_dense_eigh."""

    _implements_eq = True

    def __init__(self, electronic, n_vibrational_levels, vib_energies, elec_vib_couplings,
                 energy_spread_extra=None, site_labels=None):
        self.electronic = electronic
        self.bath = electronic.bath
        self.n_vibrational_levels = np.asarray(n_vibrational_levels)
        self.vib_energies = np.asarray(vib_energies)
        self.elec_vib_couplings = np.asarray(elec_vib_couplings)
        super(VibronicHamiltonian, self).__init__(energy_spread_extra, site_labels)

    def _eq(self, other, max_depth):
        return (isinstance(other, VibronicHamiltonian) and self.electronic == other.electronic
                and np.all(self.n_vibrational_levels == other.n_vibrational_levels)
                and np.all(self.vib_energies == other.vib_energies)
                and np.all(self.elec_vib_couplings == other.elec_vib_couplings)
                and super(VibronicHamiltonian, self)._eq(other, max_depth))

    __hash__ = object.__hash__

    @property
    def n_sites(self):
        return self.electronic.n_sites

    @property
    def dipoles(self):
        return self.electronic.dipoles

    @property
    def n_vibrational_states(self):
        return int(np.prod(self.n_vibrational_levels))

    @property
    def H_vibrational(self):
        H = np.zeros((self.n_vibrational_states,) * 2)
        for m, (levels, energy) in enumerate(zip(self.n_vibrational_levels, self.vib_energies)):
            H = H + energy * extend_vib_operator(self.n_vibrational_levels, m,
                                                 np.diag(np.arange(levels)))
        return H

    def H_electronic_vibrational(self, subspace='gef'):
        N = self.electronic.n_states(subspace) * self.n_vibrational_states
        H = np.zeros((N, N))
        for i in range(self.n_sites):
            el = self.electronic.number_operator(i, subspace)
            for m, levels in enumerate(self.n_vibrational_levels):
                q = vib_annihilate(levels) + vib_create(levels)
                H = H + self.elec_vib_couplings[i, m] * tensor(
                    el, extend_vib_operator(self.n_vibrational_levels, m, q))
        return H

    def n_states(self, subspace):
        return self.electronic.n_states(subspace) * self.n_vibrational_states

    @imemoize
    def H(self, subspace='gef'):
        return (self.el_to_sys_operator(self.electronic.H(subspace))
                + self.vib_to_sys_operator(self.H_vibrational, subspace)
                + self.H_electronic_vibrational(subspace))

    def _in_rotating_frame(self, rw_freq):
        return type(self)(self.electronic.in_rotating_frame(rw_freq), self.n_vibrational_levels,
                          self.vib_energies, self.elec_vib_couplings, self.energy_spread_extra,
                          self.site_labels)

    def _sample(self, n, random_orientations):
        return type(self)(self.electronic.sample(n, random_orientations),
                          self.n_vibrational_levels, self.vib_energies, self.elec_vib_couplings,
                          self.energy_spread_extra, self.site_labels)

    def el_to_sys_operator(self, el_operator):
        return tensor(el_operator, np.eye(self.n_vibrational_states))

    def vib_to_sys_operator(self, vib_operator, subspace='gef'):
        return tensor(np.eye(self.electronic.n_states(subspace)), vib_operator)

    def dipole_operator(self, *args, **kwargs):
        return self.el_to_sys_operator(self.electronic.dipole_operator(*args, **kwargs))

    def system_bath_couplings(self, *args, **kwargs):
        return self.el_to_sys_operator(self.electronic.system_bath_couplings(*args, **kwargs))

    def number_operator(self, site, subspace='gef'):
        return self.el_to_sys_operator(self.electronic.number_operator(site, subspace))

    def vib_basis_labels(self):
        labels = np.zeros(self.n_vibrational_states)
        n = len(self.n_vibrational_levels)
        for m, levels in enumerate(self.n_vibrational_levels):
            labels = labels + 10 ** (n - m - 1) * np.diag(
                extend_vib_operator(self.n_vibrational_levels, m, np.diag(np.arange(levels))))
        return [('{:0' + str(n) + '}').format(int(i)) for i in labels]

    def basis_labels(self, subspace='gef', braket=False):
        labels = [(e, v) for e in self.electronic.basis_labels(subspace)
                  for v in self.vib_basis_labels()]
        return add_braket(labels) if braket else labels

    # ---------------------------------------------------------- device data
    def dense_block(self, X, hilbert_subspace):
        """This Hamiltonian (in its frame) restricted to Hilbert block X of
        the given subspace, as a dense float64 matrix."""
        idx = self.hilbert_subspace_index(X, hilbert_subspace)
        H = np.asarray(self.H(hilbert_subspace))
        if np.iscomplexobj(H) and np.any(np.imag(H) != 0):
            raise NotImplementedError('complex Hermitian vibronic Hamiltonians: the GPU '
                                      'eigensolver is real symmetric')
        return np.ascontiguousarray(np.real(H)[idx, idx])

    def _eig_blocks(self, subspace):
        dev = _native.device()
        Es, Us, xs = [], [], []
        for s in 'gef':
            if s not in subspace or self.electronic.n_states(s) == 0:
                continue
            Hb = torch.as_tensor(self.dense_block(s, subspace)[None]).to(dev)
            E, U = _native.dense_eigh(Hb)
            Es.append(E[0].cpu().numpy())
            Us.append(U[0].cpu().numpy())
            xs.append(np.full(len(Es[-1]), 'gef'.index(s)))
        E = np.concatenate(Es)
        U = np.zeros((len(E), len(E)))
        k = 0
        for u in Us:
            U[k:k + len(u), k:k + len(u)] = u
            k += len(u)
        return E, U, np.concatenate(xs)
