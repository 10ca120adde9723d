"""Liouville-space bookkeeping and the device-backed equation of motion
(qspectra/dynamics/liouville_space.py API).

Conventions kept from the reference:
  * density matrices are column-stacked (vec index r + N*c);
  * a Liouville subspace 'ab,cd' selects the sorted column-major positions of
    the blocks rho_ab, rho_cd (liouville_space.py:9-29);
  * the Heisenberg-picture generator is the plain transpose L^T
    (liouville_space.py:325-330);
  * bra_vector is the row-stacked operator restricted to the 'from' indices,
    i.e. bra . vec(rho) = tr(O rho) (liouville_space.py:176-181).

The Liouvillian is block-diagonal over Liouville blocks XY (rho_ab with a in
Hilbert manifold X, b in Y); models build only the blocks a subspace needs,
on the GPU, for a whole batch of Hamiltonians at once.
"""
import itertools

import numpy as np
import torch

from .. import _native
from ..operator_tools import SubspaceError, n_excitations, full_liouville_subspace
from ..utils import imemoize
from .base import DynamicalModel, SystemOperator


# ------------------------------------------------------------ index maps
def _block_starts(full_subspace, n_sites, n_vib=1):
    starts, total = {}, 0
    for exc, size in zip('gef', n_excitations(n_sites, n_vib)):
        if exc in full_subspace:
            starts[exc] = (total, int(size))
            total += int(size)
    return starts, total


def liouville_subspace_index(liouville_subspace, full_subspace, n_sites,
                             n_vibrational_states=1):
    """Sorted vec indices (r + N c) of the blocks in ``liouville_subspace``."""
    starts, N = _block_starts(full_subspace, n_sites, n_vibrational_states)
    keep = np.zeros((N, N), dtype=bool)
    for row, col in liouville_subspace.split(','):
        if row not in starts or col not in starts:
            raise SubspaceError("{}{} not in subspace '{}'".format(row, col, full_subspace))
        (r0, nr), (c0, nc) = starts[row], starts[col]
        keep[r0:r0 + nr, c0:c0 + nc] = True
    return np.flatnonzero(keep.reshape(-1, order='F'))


def all_liouville_subspaces(hilbert_subspace):
    return ','.join(''.join(s) for s in itertools.product(hilbert_subspace, repeat=2))


def matrix_to_ket_vec(matrix):
    return np.asarray(matrix).reshape(-1, order='F')


def ket_vec_to_matrix(ket_vec):
    ket_vec = np.asarray(ket_vec)
    N = int(round(np.sqrt(ket_vec.size)))
    return ket_vec.reshape((N, N), order='F')


def matrix_to_bra_vec(matrix):
    return np.asarray(matrix).reshape(-1, order='C')


def tensor_to_super(tensor_operator):
    """super[i + N a, j + N b] = T[i, a, j, b]."""
    T = np.asarray(tensor_operator)
    N = T.shape[0]
    return T.transpose(1, 0, 3, 2).reshape(N * N, N * N)


def super_left_matrix(operator):
    return np.kron(np.identity(len(operator)), operator)


def super_right_matrix(operator):
    return np.kron(np.asarray(operator).T, np.identity(len(operator)))


def super_commutator_matrix(operator):
    return super_left_matrix(operator) - super_right_matrix(operator)


def block_layout(liouville_subspace, hilbert_subspace, n_sites, n_vib=1):
    """For each Liouville block XY in the subspace: (X, Y, positions in the
    sorted subspace vector, local indices a + nX*b).  Blocks are the
    Hilbert-manifold pairs the subspace's indices fall into."""
    starts, N = _block_starts(hilbert_subspace, n_sites, n_vib)
    idx = liouville_subspace_index(liouville_subspace, hilbert_subspace, n_sites, n_vib)
    rows, cols = idx % N, idx // N
    layout = []
    for X, (x0, nx) in starts.items():
        for Y, (y0, ny) in starts.items():
            sel = np.flatnonzero((rows >= x0) & (rows < x0 + nx) & (cols >= y0) & (cols < y0 + ny))
            if sel.size:
                local = (rows[sel] - x0) + nx * (cols[sel] - y0)
                layout.append((X, Y, sel, local))
    return layout, idx


# ------------------------------------------------------------ linear maps
class LinearMap(object):
    """A fixed linear map v -> A v on Liouville vectors.  ``matrix`` is 2-D
    [K, M] (vector output) or 1-D [M] (scalar output).  The propagation
    engine fuses a LinearMap save function into the RK4 kernel."""

    def __init__(self, matrix):
        self.matrix = np.asarray(matrix)

    def __call__(self, state):
        return self.matrix.dot(np.asarray(state))

    dot = __call__


class LiouvilleSpaceOperator(SystemOperator):
    """A Hilbert-space operator acting between Liouville subspaces
    ('from->to', or one subspace for both)."""

    def __init__(self, operator, liouv_subspace_map, dynamical_model):
        self.operator = np.asarray(operator)
        subs = (liouv_subspace_map.split('->') if '->' in liouv_subspace_map
                else [liouv_subspace_map] * 2)
        self.from_indices, self.to_indices = [dynamical_model.liouville_subspace_index(s)
                                              for s in subs]
        self.super_op_mesh = np.ix_(self.to_indices, self.from_indices)

    @property
    def bra_vector(self):
        return matrix_to_bra_vec(np.asarray(self.operator, dtype=complex))[self.from_indices]

    def _restricted(self, left):
        """Rows of kron(I, O) (left) or kron(O^T, I) (right) restricted to
        the subspace indices, without forming the N^2 x N^2 product."""
        O = np.asarray(self.operator)
        N = len(O)
        ti, fi = self.to_indices, self.from_indices
        r_to, c_to = ti % N, ti // N
        r_fr, c_fr = fi % N, fi // N
        if left:   # (O rho)_{rc} = sum_k O_rk rho_kc
            return O[r_to[:, None], r_fr[None, :]] * (c_to[:, None] == c_fr[None, :])
        return O[c_fr[None, :], c_to[:, None]] * (r_to[:, None] == r_fr[None, :])

    @imemoize
    def left_matrix(self):
        return self._restricted(True)

    @imemoize
    def right_matrix(self):
        return self._restricted(False)

    @property
    def left_multiply(self):
        return LinearMap(self.left_matrix())

    @property
    def right_multiply(self):
        return LinearMap(self.right_matrix())

    @property
    def commutator(self):
        return LinearMap(self.left_matrix() - self.right_matrix())

    @property
    def expectation_value(self):
        """tr(O rho) on the 'to' subspace, as a row vector over 'from'."""
        tr = np.identity(len(self.operator)).reshape(-1)[self.to_indices]
        return LinearMap(tr.dot(self.left_matrix()))


# --------------------------------------------------------- device EOM
class DeviceEOM(object):
    """eom(t, y) = L y with L ([B, M, M] complex128) resident on the GPU.

    Calling it applies L with qsx_apply_dense_c128 (numpy in, numpy out; a
    leading batch axis of size B is accepted when B > 1).  The propagation
    engine reads ``L`` directly."""

    def __init__(self, L, liouville_subspace, heisenberg_picture):
        self.L = L
        self.liouville_subspace = liouville_subspace
        self.heisenberg_picture = heisenberg_picture

    @property
    def batch(self):
        return self.L.shape[0]

    @property
    def size(self):
        return self.L.shape[-1]

    def __call__(self, t, y):
        y = np.asarray(y, dtype=complex)
        B, M = self.batch, self.size
        if B == 1:
            yt = y.reshape(1, -1, M)
        else:
            yt = y.reshape(B, -1, M)
        out = _native.apply_dense(self.L, _native.to_device(yt, dev=self.L.device))
        return out.cpu().numpy().reshape(y.shape)


# --------------------------------------------------------------- model
class LiouvilleSpaceModel(DynamicalModel):
    """DynamicalModel for Liouville-space dynamics; subclasses implement
    ``_build_block(X, Y, heisenberg, H1, rw)`` returning device blocks.

    evolve_basis: 'site' or 'eigen'.  sparse_matrix is accepted for API
    compatibility; the GPU kernels always use the dense blocks.
    """
    system_operator = LiouvilleSpaceOperator

    def __init__(self, hamiltonian, rw_freq=None, hilbert_subspace='gef', unit_convert=1,
                 evolve_basis='site', sparse_matrix=False):
        super(LiouvilleSpaceModel, self).__init__(hamiltonian, rw_freq, hilbert_subspace,
                                                  unit_convert)
        self.evolve_basis = evolve_basis
        self.sparse_matrix = sparse_matrix

    @property
    def evolve_basis(self):
        return self._evolve_basis

    @evolve_basis.setter
    def evolve_basis(self, val):
        if val not in ('site', 'eigen'):
            raise ValueError('invalid basis')
        self._evolve_basis = val

    def liouville_subspace_index(self, subspace):
        return liouville_subspace_index(subspace, self.hilbert_subspace,
                                        self.hamiltonian.n_sites,
                                        self.hamiltonian.n_vibrational_states)

    def density_matrix_to_state_vector(self, rho0, liouville_subspace):
        state = matrix_to_ket_vec(rho0)
        return self.map_between_subspaces(state, full_liouville_subspace(liouville_subspace),
                                          liouville_subspace)

    def state_vector_to_density_matrix(self, rho):
        rho = np.asarray(rho)
        N = int(round(np.sqrt(rho.shape[-1])))
        return rho.reshape(-1, N, N, order='F')

    def dipole_operator(self, liouv_subspace_map, polarization, transitions='-+'):
        op = self.hamiltonian.dipole_operator(self.hilbert_subspace, polarization, transitions)
        if self.evolve_basis == 'eigen':
            op = self.hamiltonian.transform_operator_to_eigenbasis(op, self.hilbert_subspace)
        return self.system_operator(op, liouv_subspace_map, self)

    def thermal_state(self, liouville_subspace):
        if liouville_subspace == 'gg' and self.hamiltonian.n_vibrational_states == 1:
            return np.ones(1, dtype=complex)   # the electronic ground manifold is one state
        rho0 = self.hamiltonian.thermal_state(liouville_subspace)
        rho = self.map_between_subspaces(matrix_to_ket_vec(rho0),
                                         full_liouville_subspace(liouville_subspace),
                                         liouville_subspace)
        if self.evolve_basis == 'eigen':
            rho = self.hamiltonian.transform_vector_to_eigenbasis(rho, liouville_subspace)
        return rho

    def thermal_states(self, liouville_subspace):
        """thermal_state of every batch member, [B, M]: site-basis models of
        real electronic Hamiltonians take one batched eigensolver launch per
        manifold; other models are evaluated member by member."""
        from ..hamiltonian import ElectronicHamiltonian, electronic_thermal_states
        from ..utils import copy_with_new_cache
        hams = self._hamiltonians()
        if len(hams) == 1:
            return np.asarray(self.thermal_state(liouville_subspace))[None]
        if liouville_subspace == 'gg' and self.hamiltonian.n_vibrational_states == 1:
            return np.ones((len(hams), 1), dtype=complex)
        batched = (self.evolve_basis == 'site'
                   and all(type(h) is ElectronicHamiltonian for h in hams)
                   and not any(np.iscomplexobj(h.H_1exc) and np.any(np.imag(h.H_1exc) != 0)
                               for h in hams))
        if not batched:
            out = []
            for h in hams:
                member = copy_with_new_cache(self)
                member.__dict__.pop('_batch', None)
                member.hamiltonian = h
                out.append(member.thermal_state(liouville_subspace))
            return np.stack(out)
        rho = electronic_thermal_states(hams, liouville_subspace)       # [B, N, N]
        vec = rho.transpose(0, 2, 1).reshape(len(hams), -1)             # column-stacked
        fi, ti = [self.liouville_subspace_index(s)
                  for s in (full_liouville_subspace(liouville_subspace), liouville_subspace)]
        N = self.hamiltonian.n_states(self.hilbert_subspace)
        full = np.zeros((len(hams), N * N), dtype=complex)
        full[:, fi] = vec
        return full[:, ti]

    def map_between_subspaces(self, state, from_subspace, to_subspace):
        fi, ti = [self.liouville_subspace_index(s) for s in (from_subspace, to_subspace)]
        N = self.hamiltonian.n_states(self.hilbert_subspace)
        full = np.zeros(N * N, dtype=complex)
        full[fi] = state
        return full[ti]

    # ---------------------------------------------------------- Liouvillian
    def _hamiltonians(self):
        return getattr(self, '_batch', None) or [self.hamiltonian]

    def with_hamiltonians(self, hamiltonians):
        """Copy of this model evaluating a batch of Hamiltonians (ensemble
        members) together; Liouvillians get a leading batch axis."""
        from ..utils import copy_with_new_cache
        new = copy_with_new_cache(self)
        new.hamiltonian = hamiltonians[0]
        new._batch = list(hamiltonians)
        return new

    def _device_inputs(self):
        hams = self._hamiltonians()
        dev = _native.device()
        H1 = torch.as_tensor(np.stack([h.H1_nonrotating for h in hams])).to(dev)
        rw = torch.tensor([float(h.rw_freq) for h in hams], dtype=torch.float64, device=dev)
        return H1, rw

    def _build_block(self, X, Y, heisenberg, H1, rw):
        raise NotImplementedError('subclass must build Liouvillian blocks')

    def _build_dense_block(self, X, Y, heisenberg, dense):
        raise NotImplementedError('%s has no dense-block builder' % type(self).__name__)

    def _dense_inputs(self):
        """For Hamiltonians without a one-exciton form (vibronic): per
        Hilbert block X, the rotating-frame block of every member
        (float64 [B, nX, nX]) and the diagonals of the system-bath
        couplings over the block ([n, nX]; None without a bath)."""
        hams = self._hamiltonians()
        dev = _native.device()
        out = {}
        for X in 'gef':
            if X not in self.hilbert_subspace or self.hamiltonian.n_states(X) == 0:
                continue
            HX = torch.as_tensor(np.stack([h.dense_block(X, self.hilbert_subspace)
                                           for h in hams])).to(dev)
            occ = None
            if self.hamiltonian.bath is not None:
                idx = self.hamiltonian.hilbert_subspace_index(X, self.hilbert_subspace)
                K = np.asarray(self.hamiltonian.system_bath_couplings(self.hilbert_subspace))
                sub = K[:, idx, idx]
                if np.any(sub - np.einsum('kii->ki', sub)[:, :, None] * np.eye(sub.shape[-1])
                          != 0):
                    raise NotImplementedError('non-diagonal system-bath couplings')
                occ = torch.as_tensor(np.ascontiguousarray(
                    np.real(np.einsum('kii->ki', sub)))).to(dev)
            out[X] = (HX, occ)
        return out

    @imemoize
    def liouvillian(self, liouville_subspace, heisenberg_picture=False):
        """Device tensor [B, M, M] of unit_convert * L on the subspace (its
        transpose for the Heisenberg picture), B = batch size."""
        n = self.hamiltonian.n_sites
        layout, idx = block_layout(liouville_subspace, self.hilbert_subspace, n,
                                   self.hamiltonian.n_vibrational_states)
        if hasattr(self.hamiltonian, 'dense_block'):
            dense = self._dense_inputs()
            build = lambda X, Y: self._build_dense_block(X, Y, heisenberg_picture, dense)
            B = len(self._hamiltonians())
            dev = _native.device()
        else:
            H1, rw = self._device_inputs()
            build = lambda X, Y: self._build_block('gef'.index(X), 'gef'.index(Y),
                                                   heisenberg_picture, H1, rw)
            B, dev = H1.shape[0], H1.device
        if len(layout) == 1 and np.array_equal(layout[0][3], np.arange(idx.size)):
            X, Y = layout[0][:2]
            return build(X, Y)
        L = torch.zeros((B, idx.size, idx.size), dtype=torch.complex128, device=dev)
        for X, Y, sel, local in layout:
            blk = build(X, Y)
            s = torch.as_tensor(sel, device=dev)
            l = torch.as_tensor(local, device=dev)
            L[:, s[:, None], s[None, :]] = blk[:, l[:, None], l[None, :]]
        return L

    def block_builder(self, liouville_subspace, heisenberg_picture=False):
        """(inputs, rebuild) for a single-block Liouville subspace of an
        electronic Hamiltonian: inputs = (H1 [B, n, n], rw [B]) device
        tensors, rebuild(L, stream=None) re-runs the builder from them into L
        [B, M, M] (the block liouvillian() returns).  A caller that writes
        new Hamiltonians into H1 in place and rebuilds gets their blocks;
        the pair of tensors is shared by every block of the model."""
        n = self.hamiltonian.n_sites
        layout, idx = block_layout(liouville_subspace, self.hilbert_subspace, n,
                                   self.hamiltonian.n_vibrational_states)
        if hasattr(self.hamiltonian, 'dense_block') or len(layout) != 1 or \
                not np.array_equal(layout[0][3], np.arange(idx.size)):
            raise NotImplementedError('rebuilds need a single Liouville block of an '
                                      'electronic Hamiltonian')
        inputs = self.__dict__.get('_builder_inputs')
        if inputs is None:
            inputs = self.__dict__['_builder_inputs'] = self._device_inputs()
        X, Y = ('gef'.index(layout[0][0]), 'gef'.index(layout[0][1]))
        return inputs, self._block_rebuilder(X, Y, heisenberg_picture, inputs)

    def _block_rebuilder(self, X, Y, heisenberg, inputs):
        raise NotImplementedError('%s has no block rebuilder' % type(self).__name__)

    @imemoize
    def equation_of_motion(self, liouville_subspace, heisenberg_picture=False):
        return DeviceEOM(self.liouvillian(liouville_subspace, heisenberg_picture),
                         liouville_subspace, heisenberg_picture)

    @property
    def evolution_super_operator(self):
        """The full N^2 x N^2 Liouvillian (numpy), assembled from blocks."""
        subspace = all_liouville_subspaces(self.hilbert_subspace)
        L = self.liouvillian(subspace)[0].cpu().numpy()
        idx = self.liouville_subspace_index(subspace)
        N = self.hamiltonian.n_states(self.hilbert_subspace)
        full = np.zeros((N * N, N * N), dtype=complex)
        full[np.ix_(idx, idx)] = L
        return full
