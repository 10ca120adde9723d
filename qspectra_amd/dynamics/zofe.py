"""ZOFE (zeroth-order functional expansion) master equation with the API of
qspectra/dynamics/zofe.py (ZOFESpaceOperator, ZOFEModel).

The state is the reference's: vec(rho) followed by vec(O), O the auxiliary
operators of shape (n_pseudomodes, n_sites, N, N), both column-major
(zofe.py:84-118).  The equation (zofe.py:121-225) is nonlinear in the state
(products of O and rho); it runs in the ZOFE kernel (qsx_zofe_propagate /
qsx_zofe_apply), one work-group per right-hand side, the whole state in LDS.
There is no Heisenberg picture (as in the reference): spectra use the
Schroedinger picture with expectation values as save functions.
"""
import numpy as np
import torch

from .. import _native
from ..bath import PseudomodeBath
from ..utils import copy_with_new_cache
from .base import DynamicalModel, SystemOperator
from .liouville_space import LinearMap


class ZOFESpaceOperator(SystemOperator):
    """A Hilbert-space operator acting on ZOFE states: products act on rho
    and on every auxiliary operator O[p, s] (zofe.py:8-39).  ``matrix`` is
    the operator; ``expectation_value`` is a LinearMap on the state (tr of
    the operator times rho), so integrate fuses it into the kernel."""

    def __init__(self, operator, liouv_subspace_map, dynamical_model):
        self.operator = np.asarray(operator)
        self.dynamical_model = dynamical_model

    @property
    def matrix(self):
        return self.operator

    def left_multiply(self, state):
        rho, oop = self.dynamical_model.state_vec_to_operators(state)
        return self.dynamical_model.operators_to_state_vec(
            self.operator.dot(rho), np.einsum('cd,abde->abce', self.operator, oop))

    def right_multiply(self, state):
        rho, oop = self.dynamical_model.state_vec_to_operators(state)
        return self.dynamical_model.operators_to_state_vec(rho.dot(self.operator),
                                                           oop.dot(self.operator))

    @property
    def expectation_value(self):
        """tr(operator rho) as a row on the state: operator[i, j] pairs with
        rho[j, i], at index j + N i of vec(rho); zero on O."""
        row = np.zeros(self.dynamical_model.state_size, dtype=complex)
        row[:self.operator.size] = self.operator.reshape(-1)
        return LinearMap(row)


class DeviceZOFE(object):
    """eom(t, y) of a (batched) ZOFE model with its operators on the GPU:
    H [B, N, N] (rotating frame), L [1, S, N, N] (negated couplings),
    Gamma, w [1, P, S]."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    @property
    def batch(self):
        return self.H.shape[0]

    def __call__(self, t, y):
        y = np.asarray(y, dtype=complex)
        B = self.batch
        yt = _native.to_device(y.reshape(B, -1, self.size))
        out = _native.zofe_apply(self, yt)
        return out.cpu().numpy().reshape(y.shape)


class ZOFEModel(DynamicalModel):
    """ZOFE master equation for identical independent pseudomode baths
    (zofe.py:42-225); the Hamiltonian's bath must be a PseudomodeBath.
    ham_hermit / rho_hermit select the reference's equivalent forms of the
    rho equation (zofe.py:170-186)."""
    system_operator = ZOFESpaceOperator

    def __init__(self, hamiltonian, rw_freq=None, hilbert_subspace='gef', unit_convert=1,
                 ham_hermit=False, rho_hermit=False):
        super(ZOFEModel, self).__init__(hamiltonian, rw_freq, hilbert_subspace, unit_convert)
        if not isinstance(self.hamiltonian.bath, PseudomodeBath):
            raise NotImplementedError('ZOFE only implemented for baths of type PseudomodeBath')
        numb_pm = self.hamiltonian.bath.numb_pm
        n_sit = self.hamiltonian.n_sites
        n_stat = self.hamiltonian.n_states(self.hilbert_subspace)
        self.oop_shape = (numb_pm, n_sit, n_stat, n_stat)
        self.ham_hermit = ham_hermit
        self.rho_hermit = rho_hermit

    @property
    def state_size(self):
        P, S, N, _ = self.oop_shape
        return N * N * (1 + P * S)

    # ------------------------------------------------------------ batching
    def _hamiltonians(self):
        return getattr(self, '_batch', None) or [self.hamiltonian]

    def with_hamiltonians(self, hamiltonians):
        new = copy_with_new_cache(self)
        new.hamiltonian = hamiltonians[0]
        new._batch = list(hamiltonians)
        return new

    # ---------------------------------------------------------- state maps
    def density_matrix_to_state_vector(self, rho0, liouville_subspace=None):
        state0 = np.asarray(rho0).reshape(-1, order='F')
        return np.append(state0, np.zeros(self.oop_shape, dtype=complex).reshape(-1, order='F'))

    def state_vector_to_density_matrix(self, rhos):
        return np.array([self.state_vec_to_operators(k)[0] for k in rhos])

    def thermal_state(self, _):
        rho0 = self.hamiltonian.thermal_state(self.hilbert_subspace)
        return self.density_matrix_to_state_vector(rho0, None)

    def map_between_subspaces(self, state, from_subspace, to_subspace):
        return state

    def state_vec_to_operators(self, rho_oop_vec):
        n_stat = self.oop_shape[-1]
        n2 = n_stat ** 2
        rho_oop_vec = np.asarray(rho_oop_vec)
        rho = rho_oop_vec[:n2].reshape((n_stat, n_stat), order='F')
        oop = rho_oop_vec[n2:].reshape(self.oop_shape, order='F')
        return rho, oop

    def operators_to_state_vec(self, rho, oop):
        return np.append(np.asarray(rho).reshape(-1, order='F'),
                         np.asarray(oop).reshape(-1, order='F'))

    def dipole_operator(self, liouv_subspace_map, polarization, transitions='-+'):
        op = self.hamiltonian.dipole_operator(self.hilbert_subspace, polarization, transitions)
        return self.system_operator(op, liouv_subspace_map, self)

    # ------------------------------------------------- equation of motion
    def equation_of_motion(self, liouville_subspace, heisenberg_picture=False):
        if heisenberg_picture:
            raise NotImplementedError('ZOFE not implemented in the Heisenberg picture')
        cache = self.__dict__.setdefault('_eoms', {})
        if 'eom' in cache:
            return cache['eom']
        dev = _native.device()
        hams = self._hamiltonians()
        bath = self.hamiltonian.bath
        H = np.stack([np.asarray(h.H(self.hilbert_subspace), dtype=complex) for h in hams])
        L = -np.asarray(self.hamiltonian.system_bath_couplings(self.hilbert_subspace),
                        dtype=complex)
        Omega = np.asarray(bath.Omega)
        Gamma = Omega ** 2 * np.asarray(bath.huang)
        w = 1j * Omega + np.asarray(bath.gamma)
        P, S, N, _ = self.oop_shape
        Gamma = np.broadcast_to(Gamma, (P, S))
        w = np.broadcast_to(w, (P, S))
        flags = (1 if self.ham_hermit else 0) | (2 if self.rho_hermit else 0)
        eom = DeviceZOFE(H=_native.to_device(H, dev=dev), L=_native.to_device(L[None], dev=dev),
                         Gamma=_native.to_device(np.ascontiguousarray(Gamma)[None], dev=dev),
                         w=_native.to_device(np.ascontiguousarray(w)[None], dev=dev),
                         flags=flags, unit_convert=self.unit_convert, N=N, S=S, P=P,
                         size=self.state_size, model=self)
        cache['eom'] = eom
        return eom
