"""Unitary dynamics L = -i [H, .] (qspectra/dynamics/unitary.py), blocks
built on the GPU (qsx_unitary_block).  As in the reference, L is always
taken in the site basis; dipoles and thermal states follow evolve_basis."""
from .. import _native
from .liouville_space import LiouvilleSpaceModel


class UnitaryModel(LiouvilleSpaceModel):
    def _build_block(self, X, Y, heisenberg, H1, rw):
        return _native.unitary_block(H1, rw, self.unit_convert, X, Y, heisenberg)

    def _build_dense_block(self, X, Y, heisenberg, dense):
        return _native.unitary_block_dense(dense[X][0], dense[Y][0], self.unit_convert,
                                           heisenberg)
