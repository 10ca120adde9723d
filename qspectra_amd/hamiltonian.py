"""System Hamiltonians with the API of qspectra/hamiltonian.py
(Hamiltonian, ElectronicHamiltonian).

Eigensystems come from the batched FP64 Jacobi kernel (qsx_block_eigh), one
excitation manifold at a time; the thermal state is built from them.

Variants.  Every Hamiltonian is a variant of the one a user constructed
(its *base*: lab frame, no disorder drawn), identified by a key (draw,
rw_freq): ``draw`` is None or (n, random_orientations) for the n-th disorder
sample, ``rw_freq`` the rotating-wave frequency (0 = lab frame).  One
factory on the base builds and caches every variant, so rotating and
sampling act on the key and commute (they even return the same object),
drawing a sample of a sample replaces the draw (disorder is always added to
the base), and a sample keeps its parent's rw_freq and hence its time and
frequency grids -- the semantics of hamiltonian.py:96-274.
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
    """The matrix as an array; ValueError unless it is square and equals its
    conjugate transpose to numpy's default tolerances (hamiltonian.py:18-22)."""
    m = np.asarray(matrix)
    square = m.ndim == 2 and m.shape[0] == m.shape[1]
    if not (square and np.allclose(m, np.conj(m).T)):
        raise ValueError('expected a Hermitian matrix, got one of shape %s that is not'
                         % (m.shape,))
    return m


def _is_complex(H):
    """True for a matrix with a non-zero imaginary part."""
    return np.iscomplexobj(H) and bool(np.any(np.imag(H) != 0))


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
    T = np.array([getattr(getattr(h.lab_frame, 'bath', None), 'temperature', 0) or 0
                  for h in hams], dtype=float)
    dE = E - E.min(axis=1, keepdims=True)
    hot = T > 0
    w = np.where(hot[:, None], np.exp(-dE / np.where(hot, T, 1.0)[:, None]), dE == 0)
    w = w / w.sum(axis=1, keepdims=True)
    return np.einsum('bij,bj,bkj->bik', U, w, U.conj()).astype(complex)


def add_braket(labels):
    out = []
    for label in labels:
        if isinstance(label, str) or not np.iterable(label):
            out.append('|{}>'.format(label))
        else:
            out.append(''.join('|{}>'.format(i) for i in label))
    return out


class Hamiltonian(metaclass=ABCMeta):
    """Base class; subclasses provide ``H(subspace)`` and, to support the
    variants described in the module docstring, ``_in_rotating_frame`` and
    ``_sample`` (each building a new instance from the base's data)."""

    _implements_eq = False

    def __init__(self, energy_spread_extra=None, site_labels=None):
        self.energy_spread_extra = energy_spread_extra
        self.site_labels = site_labels
        self.rw_freq = 0
        self._variant_of = None      # the base this is a variant of (None: itself)
        self._draw = None            # (n, random_orientations) of a disorder sample

    # ------------------------------------------------------------ variants
    @property
    def base(self):
        """The Hamiltonian as constructed: lab frame, no disorder drawn."""
        return self if self._variant_of is None else self._variant_of

    def variant(self, draw, rw_freq):
        """The base's variant with disorder ``draw`` (None, or (n,
        random_orientations)) in the frame rotating at ``rw_freq``; built
        once per key and cached on the base."""
        base = self.base
        key = (draw, float(rw_freq))
        if key == (None, 0.0):
            return base
        table = base.__dict__.setdefault('_variants', {})
        ham = table.get(key)
        if ham is None:
            ham = base._sample(*draw) if draw is not None else base
            if key[1] != 0.0:
                ham = ham._in_rotating_frame(key[1])
            ham._variant_of, ham._draw, ham.rw_freq = base, draw, rw_freq
            table[key] = ham
        return ham

    @property
    def lab_frame(self):
        """This variant without the rotating frame (same disorder draw)."""
        return self.variant(self._draw, 0.0)

    @property
    def undisordered(self):
        """This variant without the disorder draw (same frame)."""
        return self.variant(None, self.rw_freq)

    def in_rotating_frame(self, rw_freq=None):
        """This Hamiltonian in the frame rotating at rw_freq (default: the
        base's mean one-exciton energy); applying it again re-targets the
        frame instead of shifting twice."""
        if rw_freq is None:
            rw_freq = self.base.transition_energy
        return self.variant(self._draw, rw_freq)

    def sample_ensemble(self, ensemble_size=1, random_orientations=False):
        for n in range(ensemble_size):
            yield self.sample(n, random_orientations)

    def sample(self, n=None, random_orientations=False):
        """The n-th static-disorder sample (deterministic in n; random n by
        default), in this Hamiltonian's frame: it keeps rw_freq, so its time
        and frequency grids are this Hamiltonian's."""
        if n is None:
            n = np.random.randint(2 ** 30)
        return self.variant((int(n), bool(random_orientations)), self.rw_freq)

    # ------------------------------------------------------------ identity
    def __repr__(self):
        return inspect_repr(self)

    def __eq__(self, other):
        return self._eq(other, max_depth=1)

    def __ne__(self, other):
        return not self == other

    __hash__ = object.__hash__

    def _eq(self, other, max_depth):
        """Data equality (subclasses compare their fields and call this);
        variants must also agree in frame, draw and base."""
        if not self._implements_eq:
            return self is other
        if not max_depth:
            return True
        if (self.rw_freq, self.site_labels, self._draw) != \
                (other.rw_freq, other.site_labels, other._draw):
            return False
        if (self._variant_of is None) != (other._variant_of is None):
            return False
        return self._variant_of is None or self.base._eq(other.base, max_depth - 1)

    @abstractmethod
    def H(self, subspace):
        """System Hamiltonian matrix on the given Hilbert subspace."""

    def n_states(self, subspace):
        return len(self.H(subspace))

    def _in_rotating_frame(self, rw_freq):
        raise NotImplementedError('%s does not implement rotating frame transformations'
                                  % type(self).__name__)

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
        E = self.undisordered.E('gef')
        extra = (0.01 * self.base.transition_energy if self.energy_spread_extra is None
                 else self.energy_spread_extra)
        return 2 * (max(E.max(), -E.min()) + extra)

    @property
    def time_step(self):
        return 1.0 / self.freq_step

    @imemoize
    def ground_state(self, subspace):
        E, U = self.lab_frame.eig(subspace)
        return _state_from_eig(E, U, 0)

    @imemoize
    def thermal_state(self, subspace):
        """Boltzmann state of the lab-frame Hamiltonian at the bath
        temperature (ground state without a bath temperature)."""
        lab = self.lab_frame
        T = getattr(getattr(lab, 'bath', None), 'temperature', 0) or 0
        E, U = lab.eig(subspace)
        return _state_from_eig(E, U, T)

    def basis_labels(self, subspace, braket=False):
        return add_braket(self.site_labels) if braket else self.site_labels

    def hilbert_subspace_index(self, subspace, all_subspaces):
        return hilbert_subspace_index(subspace, all_subspaces, self.n_sites,
                                      self.n_vibrational_states)


class GaussianSiteDisorder(object):
    """Independent Gaussian shifts of the n site energies, of full width at
    half maximum ``fwhm`` (hamiltonian.py:458-461): called with a
    RandomState, draws n standard normals in one call and returns the
    diagonal disorder matrix."""

    def __init__(self, fwhm, n_sites):
        self.sigma = float(fwhm) * GAUSSIAN_SD_FWHM
        self.n_sites = int(n_sites)

    def __call__(self, random_state):
        return np.diag(random_state.randn(self.n_sites) * self.sigma)


diagonal_gaussian_disorder = GaussianSiteDisorder


class ElectronicHamiltonian(Hamiltonian):
    """Aggregate of two-level sites with one-exciton Hamiltonian H_1exc, an
    identical independent bath per site, transition dipoles [n, 3] and
    diagonal static disorder given as a FWHM (or a function of a
    RandomState returning an n x n disorder matrix)."""

    _implements_eq = True

    n_vibrational_states = 1

    def __init__(self, H_1exc, bath=None, dipoles=None, disorder=None, random_seed=0,
                 energy_spread_extra=None, site_labels=None):
        super(ElectronicHamiltonian, self).__init__(energy_spread_extra, site_labels)
        self.H_1exc = check_hermitian(H_1exc)
        self.dipoles = None if dipoles is None else np.asarray(dipoles)
        self.bath, self.disorder, self.random_seed = bath, disorder, random_seed

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

    def _disorder_draw(self):
        """The disorder as a function of a RandomState (a FWHM number becomes
        Gaussian site disorder; None draws nothing)."""
        if self.disorder is None:
            return None
        if isinstance(self.disorder, Number):
            return GaussianSiteDisorder(self.disorder, self.n_sites)
        return self.disorder

    def _sample(self, n, random_orientations):
        """Member n: a RandomState seeded with the seed list extended by n
        (hamiltonian.py:566-567) draws the disorder, then the orientation."""
        draw = self._disorder_draw()
        if draw is None and not random_orientations:
            warnings.warn('sampling a Hamiltonian without disorder or random orientations '
                          'returns a copy of it', RuntimeWarning, stacklevel=4)
        seed = np.atleast_1d(self.random_seed).tolist() + [n]
        rng = check_random_state(seed)
        H = self.H_1exc if draw is None else self.H_1exc + draw(rng)
        dipoles = (self.dipoles @ random_rotation_matrix(rng).T if random_orientations
                   else self.dipoles)
        return type(self)(H, bath=self.bath, dipoles=dipoles, disorder=self.disorder,
                          random_seed=seed, energy_spread_extra=self.energy_spread_extra,
                          site_labels=self.site_labels)

    # ---------------------------------------------------------- device data
    @property
    def H1_nonrotating(self):
        """One-exciton Hamiltonian in the lab frame (with this sample's
        disorder), the input of the GPU eigen- and Liouvillian builders:
        float64, or complex128 for a complex-Hermitian H_1exc."""
        H = self.lab_frame.H_1exc
        return np.asarray(H, dtype=complex if _is_complex(H) else float)

    def _eig_blocks(self, subspace):
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
        U = np.zeros((len(E), len(E)), dtype=np.result_type(*Us))
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
        the given subspace, as a dense float64 matrix (complex128 when the
        Hamiltonian is complex Hermitian)."""
        idx = self.hilbert_subspace_index(X, hilbert_subspace)
        H = np.asarray(self.H(hilbert_subspace))
        H = H if _is_complex(H) else np.real(H)
        return np.ascontiguousarray(H[idx, idx])

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
        U = np.zeros((len(E), len(E)), dtype=np.result_type(*Us))
        k = 0
        for u in Us:
            U[k:k + len(u), k:k + len(u)] = u
            k += len(u)
        return E, U, np.concatenate(xs)
