"""The DynamicalModel / SystemOperator API (qspectra/dynamics/base.py)."""
from abc import ABCMeta, abstractmethod

from ..utils import copy_with_new_cache, inspect_repr


class DynamicalModel(metaclass=ABCMeta):
    """Specifies how a system evolves freely and under applied fields.

    hamiltonian is moved to the rotating frame at rw_freq (default: its
    mean one-exciton energy); unit_convert converts energies to inverse time.
    """

    def __init__(self, hamiltonian, rw_freq=None, hilbert_subspace='gef', unit_convert=1):
        self.hamiltonian = hamiltonian.in_rotating_frame(rw_freq)
        self.rw_freq = self.hamiltonian.rw_freq
        self.hilbert_subspace = hilbert_subspace
        self.unit_convert = unit_convert

    def __repr__(self):
        return inspect_repr(self)

    def density_matrix_to_state_vector(self, rho, liouville_subspace=None):
        return rho

    def state_vector_to_density_matrix(self, rho):
        return rho

    @abstractmethod
    def thermal_state(self, liouville_subspace):
        """Thermal state vector on the Liouville subspace."""

    @abstractmethod
    def equation_of_motion(self, liouville_subspace, heisenberg_picture=False):
        """eom(t, y) -> dy/dt on the subspace (NotImplementedError when the
        Heisenberg picture is not available)."""

    @abstractmethod
    def map_between_subspaces(self, state, from_subspace, to_subspace):
        """State restricted/zero-extended from one Liouville subspace to another."""

    def dipole_operator(self, liouv_subspace_map, polarization, transitions='-+'):
        op = self.hamiltonian.dipole_operator(self.hilbert_subspace, polarization, transitions)
        return self.system_operator(op, liouv_subspace_map, self)

    def dipole_destroy(self, liouville_subspace_map, polarization):
        return self.dipole_operator(liouville_subspace_map, polarization, '-')

    def dipole_create(self, liouville_subspace_map, polarization):
        return self.dipole_operator(liouville_subspace_map, polarization, '+')

    def sample_ensemble(self, *args, **kwargs):
        """Yields copies of this model over sampled Hamiltonians."""
        for ham in self.hamiltonian.sample_ensemble(*args, **kwargs):
            new = copy_with_new_cache(self)
            new.hamiltonian = ham
            yield new

    @property
    def time_step(self):
        """Default output time step (rotating frame), in time units; a
        ``time_step_override`` attribute (set by batched drivers so that
        systems of different size share one grid) takes precedence."""
        override = self.__dict__.get('time_step_override')
        if override is not None:
            return override
        return self.hamiltonian.time_step / self.unit_convert

    def hilbert_subspace_index(self, subspace):
        return self.hamiltonian.hilbert_subspace_index(subspace, self.hilbert_subspace)


class SystemOperator(metaclass=ABCMeta):
    """An operator whose commutator / products / expectation value act on
    state vectors of a DynamicalModel."""

    def commutator(self, state):
        return self.left_multiply(state) - self.right_multiply(state)

    @abstractmethod
    def left_multiply(self, state):
        pass

    @abstractmethod
    def right_multiply(self, state):
        pass

    @abstractmethod
    def expectation_value(self, state):
        pass
