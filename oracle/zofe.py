"""Oracle: the ZOFE (zeroth-order functional expansion) master equation.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  numpy restatement of
qspectra/dynamics/zofe.py (ZOFESpaceOperator :8-39, ZOFEModel :42-225) and
the PseudomodeBath parameters (qspectra/bath.py:105-145).

State vector: vec_F(rho) (N^2, column-major) followed by vec_F(O) of the
auxiliary operator O with shape (n_pseudomodes, n_sites, N, N), also
column-major (zofe.py:96-118).
"""
import numpy as np

from .liouville import LiouvilleModel


class PseudomodeBath(object):
    """bath.py:105-145: Lorentzians of the correlation spectrum; Omega,
    gamma, huang of shape (numb_pm, n_sites)."""

    def __init__(self, numb_pm, Omega, gamma, huang):
        self.numb_pm, self.Omega, self.gamma, self.huang = numb_pm, Omega, gamma, huang


class ZOFEOperator(object):
    """zofe.py:8-39: products act on rho and on every O[p, s]."""

    def __init__(self, operator, model):
        self.operator = np.asarray(operator)
        self.model = model

    def left_multiply(self, state):
        rho, oop = self.model.split(state)
        return self.model.join(self.operator.dot(rho),
                               np.einsum('cd,abde->abce', self.operator, oop))

    def right_multiply(self, state):
        rho, oop = self.model.split(state)
        return self.model.join(rho.dot(self.operator), oop.dot(self.operator))

    def commutator(self, state):
        return self.left_multiply(state) - self.right_multiply(state)

    def expectation_value(self, state):
        rho, _ = self.model.split(state)
        return np.sum(self.operator * rho.T)            # tr(M rho)


class ZOFEOracleModel(object):
    """zofe.py:42-225 on the electronic Hamiltonian of oracle/system.py; the
    Hilbert-space parts (rotating frame, dipoles, couplings, thermal state)
    come from the oracle's Liouville model of the same Hamiltonian."""

    def __init__(self, hamiltonian, rw_freq=None, hilbert_subspace='gef', unit_convert=1,
                 ham_hermit=False, rho_hermit=False):
        self.lspace = LiouvilleModel(hamiltonian, rw_freq, hilbert_subspace, unit_convert,
                                     kind='unitary')
        self.hamiltonian = self.lspace.hamiltonian
        self.rw_freq = self.lspace.rw_freq
        self.hilbert_subspace = hilbert_subspace
        self.unit_convert = unit_convert
        self.ham_hermit, self.rho_hermit = ham_hermit, rho_hermit
        bath = self.hamiltonian.bath
        self.N = len(self.hamiltonian.H(hilbert_subspace))
        self.oop_shape = (bath.numb_pm, self.hamiltonian.n_sites, self.N, self.N)

    @property
    def time_step(self):
        return self.lspace.time_step

    # state maps (zofe.py:84-118)
    def split(self, v):
        n2 = self.N ** 2
        return (v[:n2].reshape(self.N, self.N, order='F'),
                v[n2:].reshape(self.oop_shape, order='F'))

    def join(self, rho, oop):
        return np.append(rho.reshape(-1, order='F'), oop.reshape(-1, order='F'))

    def density_matrix_to_state_vector(self, rho0, liouville_subspace=None):
        return self.join(np.asarray(rho0, dtype=complex),
                         np.zeros(self.oop_shape, dtype=complex))

    def state_vector_to_density_matrix(self, states):
        return np.array([self.split(v)[0] for v in states])

    def thermal_state(self, _):
        return self.density_matrix_to_state_vector(
            self.hamiltonian.thermal_state(self.hilbert_subspace))

    def map_between_subspaces(self, state, from_subspace, to_subspace):
        return state

    def dipole_operator(self, liouv_subspace_map, polarization, transitions='-+'):
        op = self.hamiltonian.dipole_operator(self.hilbert_subspace, polarization, transitions)
        return ZOFEOperator(op, self)

    # the master equation (zofe.py:121-202, equation_of_motion :204-225)
    def rhs(self, y):
        ham = self.hamiltonian.H(self.hilbert_subspace)
        L = -np.asarray(self.hamiltonian.system_bath_couplings(self.hilbert_subspace))
        bath = self.hamiltonian.bath
        Gamma = bath.Omega ** 2 * bath.huang
        w = 1j * bath.Omega + bath.gamma
        rho, oop = self.split(y)
        Osum = oop.sum(axis=0)                                          # [site, N, N]
        Ld = L.conj().transpose(0, 2, 1)                                # L_s^dagger
        a = np.einsum('sij,sjk->ik', Ld, Osum)                          # sum_s L^+ O_s
        b = -1j * ham - a
        c = np.einsum('sij,jk,slk->il', L, rho, Osum.conj())            # sum_s L rho O^+
        d = b.dot(rho) + c
        if not self.rho_hermit:
            big = np.einsum('sij,jk,skl->il', Osum, rho, Ld)            # sum_s O rho L^+
            f = (rho.dot(b.conj().T) if self.ham_hermit
                 else rho.dot(1j * ham - a.conj().T)) + big
        else:
            f = (d.conj().T if self.ham_hermit
                 else rho.dot(1j * ham - a.conj().T) + c.conj().T)
        rhodot = d + f
        oopdot = (Gamma[:, :, None, None] * L[None]
                  - w[:, :, None, None] * oop
                  + np.einsum('ij,psjk->psik', b, oop)
                  - np.einsum('psij,jk->psik', oop, b))
        return self.unit_convert * self.join(rhodot, oopdot)

    def equation_of_motion(self, liouville_subspace, heisenberg_picture=False):
        if heisenberg_picture:
            raise NotImplementedError('ZOFE not implemented in the Heisenberg picture')
        return lambda t, y: self.rhs(y)
