"""Operators on the g/e/f manifolds of an aggregate of n two-level sites
(hard-core bosons), with the conventions of qspectra/operator_tools.py.

States are listed ground state first, then the n single excitations, then
the n(n-1)/2 double excitations (i, j), i < j, in lexicographic order.
Everything here is host-side bookkeeping on small matrices.
"""
from functools import reduce

import functools

import numpy as np


class SubspaceError(Exception):
    """An invalid Hilbert or Liouville subspace was requested."""


def tensor(*args):
    return reduce(np.kron, args)


def vib_annihilate(N):
    """Annihilation operator of an N-level vibrational mode."""
    return np.diag(np.sqrt(np.arange(1, N)), k=1)


def vib_create(N):
    """Creation operator of an N-level vibrational mode."""
    return np.diag(np.sqrt(np.arange(1, N)), k=-1)


def extend_vib_operator(n_vibrational_levels, m, vib_operator):
    """Embeds the operator of mode m into the product space of all modes."""
    levels = np.asarray(n_vibrational_levels)
    return tensor(np.eye(int(np.prod(levels[:m]))), vib_operator,
                  np.eye(int(np.prod(levels[m + 1:]))))


def unit_vec(n, N, dtype=complex):
    v = np.zeros(N, dtype=dtype)
    v[n] = 1
    return v


def _liouville_or_hilbert(X, U):
    if U.ndim != 2 or U.shape[0] != U.shape[1]:
        raise ValueError('basis transformation must be a square matrix')
    if X.shape[-1] == len(U):
        return U
    if X.shape[-1] == len(U) ** 2:
        return np.kron(U, U)
    raise ValueError('basis transformation incompatible with operator dimensions')


def basis_transform_operator(X, U):
    """U^dagger X U, with U -> kron(U, U) for Liouville-space operators."""
    X = np.asarray(X)
    U = np.asarray(U)
    if X.ndim != 2:
        raise ValueError('operator must have ndim=2')
    U = _liouville_or_hilbert(X, U)
    return U.conj().T @ X @ U


def basis_transform_vector(rho, U):
    """U^dagger rho for state vectors (broadcast over leading axes)."""
    rho = np.asarray(rho)
    U = _liouville_or_hilbert(rho, np.asarray(U))
    return np.tensordot(rho, U.conj().T, axes=(-1, -1))


def all_states(N, subspace='gef'):
    """States of the requested manifolds as lists of excited sites."""
    states = []
    if 'g' in subspace:
        states.append([])
    if 'e' in subspace:
        states += [[i] for i in range(N)]
    if 'f' in subspace:
        states += [[i, j] for i in range(N) for j in range(i + 1, N)]
    return states


def _pairs(N):
    i, j = np.triu_indices(N, k=1)
    return i, j


def operator_1_to_2(operator1):
    """Two-excitation matrix of sum_nm A_nm a^+_n a_m (hard-core bosons):
    element ((i,j),(k,l)) = A_jl d_ik + A_jk d_il + A_il d_jk + A_ik d_jl."""
    A = np.asarray(operator1)
    i, j = _pairs(len(A))
    I, J = i[:, None], j[:, None]
    K, L = i[None, :], j[None, :]
    return (A[J, L] * (I == K) + A[J, K] * (I == L)
            + A[I, L] * (J == K) + A[I, K] * (J == L))


def operator_extend(operator1, subspace='gef'):
    """Block-diagonal extension of a one-excitation operator to the
    requested manifolds (the ground-state block is 0)."""
    A = np.asarray(operator1)
    blocks = []
    if 'g' in subspace:
        blocks.append(np.zeros((1, 1), A.dtype))
    if 'e' in subspace:
        blocks.append(A)
    if 'f' in subspace:
        blocks.append(operator_1_to_2(A))
    size = sum(len(b) for b in blocks)
    out = np.zeros((size, size), dtype=A.dtype)
    start = 0
    for b in blocks:
        out[start:start + len(b), start:start + len(b)] = b
        start += len(b)
    return out


def _state_masks(n_sites, subspace):
    return np.array([sum(1 << s for s in st) for st in all_states(n_sites, subspace)],
                    dtype=np.int64)


@functools.lru_cache(maxsize=256)
def transition_operators(n_sites, subspace='gef', include_transitions='-+'):
    """transition_operator of every site, [n_sites, N, N] (read-only, cached
    per argument set: the dipole operators of an ensemble share them)."""
    masks = _state_masks(n_sites, subspace)
    bits = np.int64(1) << np.arange(n_sites, dtype=np.int64)
    Mi, Mj = masks[None, :, None], masks[None, None, :]
    b = bits[:, None, None]
    out = np.zeros((n_sites, len(masks), len(masks)))
    if '+' in include_transitions:
        out[(Mi == (Mj | b)) & ((Mj & b) == 0)] = 1
    if '-' in include_transitions:
        out[(Mj == (Mi | b)) & ((Mi & b) == 0)] = 1
    out.setflags(write=False)
    return out


def transition_operator(n, n_sites, subspace='gef', include_transitions='-+'):
    """Matrix of the excitation creation ('+') and/or annihilation ('-') at
    site n; element (i, j) is 1 when state i = state j with site n added
    ('+') or state j = state i with site n added ('-')."""
    return transition_operators(n_sites, subspace, include_transitions)[n].copy()


def n_excitations(n_sites=1, n_vibrational_states=1):
    """Numbers of 0-, 1- and 2-excitation states."""
    return np.array([1, n_sites, n_sites * (n_sites - 1) // 2]) * n_vibrational_states


def excitation_to_number(excitation):
    return 'gef'.index(excitation)


def extract_subspace(subspaces_string):
    """Smallest Hilbert subspace containing a Liouville subspace or map."""
    return sorted(set(subspaces_string) - {',', '-', '>'}, key=excitation_to_number)


def full_liouville_subspace(subspaces_string):
    h = extract_subspace(subspaces_string)
    return ','.join(a + b for a in h for b in h)


def hilbert_subspace_index(subspace, all_subspaces, n_sites, n_vibrational_states=1):
    """Slice selecting manifold ``subspace`` within ``all_subspaces``."""
    if subspace not in all_subspaces:
        raise SubspaceError("{} not in set of all subspaces '{}'".format(subspace, all_subspaces))
    sizes = n_excitations(n_sites, n_vibrational_states)
    start = 0
    for s in all_subspaces:
        size = int(sizes['gef'.index(s)])
        if s == subspace:
            return slice(start, start + size)
        start += size
