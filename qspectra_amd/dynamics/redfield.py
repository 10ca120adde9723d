"""Redfield theory (qspectra/dynamics/redfield.py API) with the Liouvillian
blocks built on the GPU in FP64 (qsx_redfield_block): eigensystem, Debye
C(w), Redfield tensor (May & Kuhn Eqs. 3.319, 3.322) and the basis change
to the site basis, for a whole batch of Hamiltonians per launch."""
import numpy as np

from .. import _native
from ..bath import debye_parameters, is_debye_form
from .liouville_space import LiouvilleSpaceModel, all_liouville_subspaces

# The reference moves the Redfield Liouvillian to the site basis with
# kron(U, U) (operator_tools.py:43-56, redfield.py:95-104), which is the
# basis change of rho -> U^H rho U only for a real U; for a complex-Hermitian
# H its result depends on the phases eigh happens to give the eigenvectors,
# so there is no reference result to reproduce.
COMPLEX_REDFIELD = ('Redfield dynamics of a complex-Hermitian Hamiltonian: the reference '
                    'transforms the Liouvillian with kron(U, U), which for complex '
                    'eigenvectors depends on their arbitrary phases; use UnitaryModel, '
                    'HEOMModel or ZOFEModel, or a real Hamiltonian')


class RedfieldModel(LiouvilleSpaceModel):
    """Redfield dynamics; each site couples to an identical independent bath.

    secular: keep only population transfer and coherence decay terms.
    discard_imag_corr: use the real part of the bath correlation function.
    """

    def __init__(self, hamiltonian, rw_freq=None, hilbert_subspace='gef', unit_convert=1,
                 secular=True, discard_imag_corr=False, evolve_basis='site',
                 sparse_matrix=False):
        super(RedfieldModel, self).__init__(hamiltonian, rw_freq, hilbert_subspace,
                                            unit_convert, evolve_basis, sparse_matrix)
        self.secular = secular
        self.discard_imag_corr = discard_imag_corr

    def _flags(self, heisenberg):
        return ((_native.QSX_SECULAR if self.secular else 0)
                | (_native.QSX_REAL_CORR if self.discard_imag_corr else 0)
                | (_native.QSX_EIGEN_BASIS if self.evolve_basis == 'eigen' else 0)
                | (_native.QSX_HEISENBERG if heisenberg else 0))

    def _build_block(self, X, Y, heisenberg, H1, rw):
        if H1.is_complex():
            raise NotImplementedError(COMPLEX_REDFIELD)
        bath = self.hamiltonian.bath
        if not is_debye_form(bath):
            return self._build_tabulated_block(X, Y, heisenberg, H1, rw)
        T, lam, gam = debye_parameters(bath)
        return _native.redfield_block(H1, rw, T, lam, gam, X, Y, self._flags(heisenberg),
                                      self.unit_convert)

    def _correlation_function(self):
        """The bath function redfield_tensor evaluates (redfield.py:48-53):
        the real one for discard_imag_corr, else the complex one (a bath
        without it raises AttributeError, as in the reference)."""
        bath = self.hamiltonian.bath
        return bath.corr_func_real if self.discard_imag_corr else bath.corr_func_complex

    def _build_tabulated_block(self, X, Y, heisenberg, H1, rw):
        """A bath given by Python functions (ArbitraryBath, user baths): the
        eigen-energies come from the GPU eigensolver, the host evaluates
        C(E_d - E_c) for each member's pairs, the rest of the build runs on
        the GPU (qsx_redfield_block_tabulated)."""
        corr = self._correlation_function()
        tabs = []
        for blk in (X, Y):
            E, _ = _native.block_eigh(H1, blk, rw)
            E = E.cpu().numpy()
            tab = np.empty(E.shape + E.shape[-1:], dtype=complex)
            for b in range(E.shape[0]):
                for d in range(E.shape[1]):
                    for c in range(E.shape[1]):
                        tab[b, d, c] = corr(E[b, d] - E[b, c])
            tabs.append(_native.to_device(tab))
        return _native.redfield_block_tabulated(H1, rw, tabs[0], tabs[1], X, Y,
                                                self._flags(heisenberg), self.unit_convert)

    def _block_rebuilder(self, X, Y, heisenberg, inputs):
        if not is_debye_form(self.hamiltonian.bath):
            raise NotImplementedError('in-graph rebuilds need the GPU form of the bath '
                                      '(DebyeBath or UncoupledBath)')
        T, lam, gam = debye_parameters(self.hamiltonian.bath)
        H1, rw = inputs
        if H1.is_complex():
            raise NotImplementedError(COMPLEX_REDFIELD)
        ws = _native.redfield_workspace(H1.shape[-1], X, Y, H1.shape[0], H1.device)
        flags = self._flags(heisenberg)

        def rebuild(L, stream=None):
            _native.redfield_block(H1, rw, T, lam, gam, X, Y, flags, self.unit_convert,
                                   stream=stream, L=L, ws=ws)
        return rebuild

    def _build_dense_block(self, X, Y, heisenberg, dense):
        T, lam, gam = debye_parameters(self.hamiltonian.bath)
        (HX, oX), (HY, oY) = dense[X], dense[Y]
        if HX.is_complex() or HY.is_complex():
            raise NotImplementedError(COMPLEX_REDFIELD)
        return _native.redfield_block_dense(HX, HY, oX, oY, T, lam, gam,
                                            self._flags(heisenberg), self.unit_convert)


def redfield_evolve(hamiltonian, subspace='ge', evolve_basis='site', **kwargs):
    """Full Redfield Liouvillian on the Hilbert subspace (numpy), without
    unit conversion (redfield.py:95-104)."""
    model = RedfieldModel.__new__(RedfieldModel)
    LiouvilleSpaceModel.__init__(model, hamiltonian, hamiltonian.rw_freq, subspace, 1,
                                 evolve_basis)
    model.hamiltonian = hamiltonian
    model.secular = kwargs.get('secular', True)
    model.discard_imag_corr = kwargs.get('discard_imag_corr', False)
    return model.evolution_super_operator
