"""Oracle: Liouville-space bookkeeping and the Redfield / unitary / HEOM
Liouvillians.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Literal restatement of
qspectra/dynamics/{liouville_space,redfield,unitary,heom}.py: the full
N^2 x N^2 superoperator is built and sliced exactly as the reference does, so
the product's block-wise GPU builder is checked against the reference's own
construction order.
"""
import itertools

import numpy as np
from scipy.sparse import csr_matrix, lil_matrix

from .system import (n_excitations, full_liouville_subspace,
                     basis_transform_operator, basis_transform_vector)


# ------------------------------------------------------ vectorisation rules
def matrix_to_ket_vec(m):
    """liouville_space.py:46-50 (column stacking)."""
    return np.asarray(m).reshape(-1, order='F')


def ket_vec_to_matrix(v):
    """liouville_space.py:53-58."""
    N = int(np.sqrt(np.prod(np.shape(v))))
    return np.asarray(v).reshape((N, N), order='F')


def matrix_to_bra_vec(m):
    """liouville_space.py:61-65 (row stacking)."""
    return np.asarray(m).reshape(-1, order='C')


def liouville_subspace_index(liouville_subspace, full_subspace, n_sites,
                             n_vibrational_states=1):
    """liouville_space.py:9-29: sorted column-major positions of the blocks."""
    total, cuts = 0, {}
    for exc, n in zip('gef', n_excitations(n_sites, n_vibrational_states)):
        if exc in full_subspace:
            cuts[exc] = np.arange(n) + total
            total += n
    keep = np.zeros((total, total))
    for row, col in liouville_subspace.split(','):
        if row not in cuts or col not in cuts:
            raise KeyError('%s%s not in subspace %r' % (row, col, full_subspace))
        keep[np.ix_(cuts[row], cuts[col])] = 1
    return matrix_to_ket_vec(keep).nonzero()[0]


def all_liouville_subspaces(hilbert_subspace):
    """liouville_space.py:32-43."""
    return ','.join(''.join(s) for s in itertools.product(hilbert_subspace, repeat=2))


def tensor_to_super(t):
    """liouville_space.py:68-90: super[i + N a, j + N b] = T[i, a, j, b]."""
    N = t.shape[0]
    s = np.empty((N ** 2, N ** 2), dtype=t.dtype)
    for i in range(N):
        for j in range(N):
            s[i::N, j::N] = t[i, :, j, :]
    return s


def super_left_matrix(op):
    """liouville_space.py:102-115."""
    return np.kron(np.identity(len(op)), op)


def super_right_matrix(op):
    """liouville_space.py:118-131."""
    return np.kron(np.asarray(op).T, np.identity(len(op)))


def super_commutator_matrix(op):
    """liouville_space.py:93-99."""
    return super_left_matrix(op) - super_right_matrix(op)


class LiouvilleSpaceOperator(object):
    """liouville_space.py:151-209 (sliced superoperators; bra = tr(O rho))."""

    def __init__(self, operator, liouv_subspace_map, model):
        self.operator = np.asarray(operator)
        subs = (liouv_subspace_map.split('->') if '->' in liouv_subspace_map
                else [liouv_subspace_map, liouv_subspace_map])
        self.from_indices, self.to_indices = [model.liouville_subspace_index(s) for s in subs]
        self.mesh = np.ix_(self.to_indices, self.from_indices)

    @property
    def bra_vector(self):
        return matrix_to_bra_vec(np.asarray(self.operator, complex))[self.from_indices]

    def left_matrix(self):
        return super_left_matrix(self.operator)[self.mesh]

    def right_matrix(self):
        return super_right_matrix(self.operator)[self.mesh]

    def commutator_matrix(self):
        return super_commutator_matrix(self.operator)[self.mesh]

    def left_multiply(self, state):
        return self.left_matrix().dot(state)

    def right_multiply(self, state):
        return self.right_matrix().dot(state)

    def commutator(self, state):
        return self.commutator_matrix().dot(state)

    def expectation_row(self):
        tr = np.identity(len(self.operator)).reshape(-1)[self.to_indices]
        return tr.dot(self.left_matrix())

    def expectation_value(self, state):
        return self.expectation_row().dot(state)


# ------------------------------------------------------------------ Redfield
def secular_terms(n):
    """redfield.py:77-83."""
    I = np.identity(n, dtype=bool)
    return np.einsum('ab,cd->abcd', I, I) | np.einsum('ac,bd->abcd', I, I)


def redfield_tensor(ham, subspace='ge', secular=True, discard_imag_corr=False):
    """redfield.py:9-74 (May & Kuhn Eqs. 3.319, 3.322)."""
    n = ham.n_states(subspace)
    E, U = ham.eig(subspace)
    K = [basis_transform_operator(c, U) for c in ham.system_bath_couplings(subspace)]
    corr = ham.bath.corr_func_real if discard_imag_corr else ham.bath.corr_func_complex
    C = np.array([[corr(Ei - Ej) for Ej in E] for Ei in E])
    Gamma = np.einsum('iab,icd,dc->abcd', K, K, C)
    I = np.identity(n)
    Gs = np.einsum('abbc->ac', Gamma)
    R = (np.einsum('ac,bd->abcd', I, Gs).conj()
         + np.einsum('bd,ac->abcd', I, Gs)
         - np.einsum('cabd->abcd', Gamma).conj()
         - np.einsum('dbac->abcd', Gamma))
    if secular:
        R = R * secular_terms(n)
    return R


def redfield_evolve(ham, subspace='ge', evolve_basis='site', **kw):
    """redfield.py:95-104."""
    H = np.diag(ham.E(subspace))
    R = tensor_to_super(redfield_tensor(ham, subspace, **kw))
    L = -1j * super_commutator_matrix(H) - R
    if evolve_basis == 'site':
        return basis_transform_operator(L, ham.U(subspace).T.conj())
    if evolve_basis == 'eigen':
        return L
    raise ValueError('invalid basis')


class LiouvilleModel(object):
    """liouville_space.py:212-349 + redfield.py:107-152 + unitary.py:5-9.

    kind: 'redfield' or 'unitary'.
    """

    def __init__(self, hamiltonian, rw_freq=None, hilbert_subspace='gef',
                 unit_convert=1, kind='redfield', secular=True,
                 discard_imag_corr=False, evolve_basis='site'):
        self.hamiltonian = hamiltonian.in_rotating_frame(rw_freq)
        self.rw_freq = self.hamiltonian.rw_freq
        self.hilbert_subspace = hilbert_subspace
        self.unit_convert = unit_convert
        self.kind = kind
        self.secular = secular
        self.discard_imag_corr = discard_imag_corr
        self.evolve_basis = evolve_basis
        self._L = None

    def sample(self, n):
        """base.py:120-128: copy with a resampled Hamiltonian (fresh cache)."""
        new = LiouvilleModel.__new__(LiouvilleModel)
        new.__dict__.update(self.__dict__)
        new.hamiltonian = self.hamiltonian.sample(n)
        new._L = None
        return new

    @property
    def time_step(self):
        """base.py:130-136."""
        return self.hamiltonian.time_step / self.unit_convert

    def liouville_subspace_index(self, s):
        return liouville_subspace_index(s, self.hilbert_subspace, self.hamiltonian.n_sites,
                                        getattr(self.hamiltonian, 'n_vibrational_states', 1))

    @property
    def evolution_super_operator(self):
        if self._L is None:
            if self.kind == 'redfield':
                self._L = self.unit_convert * redfield_evolve(
                    self.hamiltonian, self.hilbert_subspace, self.evolve_basis,
                    secular=self.secular, discard_imag_corr=self.discard_imag_corr)
            else:
                H = self.unit_convert * self.hamiltonian.H(self.hilbert_subspace)
                self._L = -1j * super_commutator_matrix(H)
        return self._L

    def evolve_matrix(self, subspace, heisenberg_picture=False):
        """liouville_space.py:316-341 (slice, transpose for Heisenberg)."""
        idx = self.liouville_subspace_index(subspace)
        L = self.evolution_super_operator[np.ix_(idx, idx)]
        return L.T if heisenberg_picture else L

    def dipole_operator(self, liouv_subspace_map, polarization, transitions='-+'):
        """liouville_space.py:281-293."""
        op = self.hamiltonian.dipole_operator(self.hilbert_subspace, polarization, transitions)
        if self.evolve_basis == 'eigen':
            U = self.hamiltonian.U(self.hilbert_subspace)
            op = basis_transform_operator(op, U)
        return LiouvilleSpaceOperator(op, liouv_subspace_map, self)

    def map_between_subspaces(self, state, from_subspace, to_subspace):
        """liouville_space.py:343-349."""
        fi, ti = [self.liouville_subspace_index(s) for s in (from_subspace, to_subspace)]
        N = self.hamiltonian.n_states(self.hilbert_subspace)
        new = np.zeros(N * N, dtype=complex)
        new[fi] = state
        return new[ti]

    def thermal_state(self, liouville_subspace):
        """liouville_space.py:300-309."""
        rho0 = self.hamiltonian.thermal_state(liouville_subspace)
        rho = self.map_between_subspaces(matrix_to_ket_vec(rho0),
                                         full_liouville_subspace(liouville_subspace),
                                         liouville_subspace)
        if self.evolve_basis == 'eigen':
            rho = basis_transform_vector(rho, self.hamiltonian.U(liouville_subspace))
        return rho

    def density_matrix_to_state_vector(self, rho0, liouville_subspace):
        """liouville_space.py:251-260."""
        return self.map_between_subspaces(matrix_to_ket_vec(rho0),
                                          full_liouville_subspace(liouville_subspace),
                                          liouville_subspace)


# ---------------------------------------------------------------------- HEOM
def matsubara_frequencies(K, gamma, T):
    """heom.py:61-67."""
    v = 2 * np.pi * T * np.arange(K + 1)
    v[0] = gamma
    return v


def corr_func_coeffs(K, gamma, T, reorg_en, nu):
    """heom.py:69-89 (aki_temp_corr=False branch)."""
    c = [reorg_en * gamma * (1 / np.tan(gamma / (2 * T)) - 1j)]
    for k in range(1, K + 1):
        c.append(4 * reorg_en * gamma * T * nu[k] / (nu[k] ** 2 - gamma ** 2))
    return np.array(c, dtype=complex)


def multichoose(n, c):
    """heom.py:155-174 (stars and bars, recursive order)."""
    if not c:
        return [[0] * n]
    if not n:
        return []
    if n == 1:
        return [[c]]
    return ([[0] + v for v in multichoose(n - 1, c)]
            + [[v[0] + 1] + v[1:] for v in multichoose(n, c - 1)])


def ado_mappings(N, K, level_cutoff):
    """heom.py:92-152: all N x (K+1) nonneg. integer matrices, sum < level_cutoff."""
    perms = []
    for c in range(level_cutoff):
        perms.extend(multichoose(N * (K + 1), c))
    inv = {tuple(v): i for i, v in enumerate(perms)}
    return [np.array(v).reshape(N, K + 1) for v in perms], inv


def heom_tensor(ham, hilbert_subspace, liouville_subspace, K, level_cutoff,
                low_temp_corr=True, modified_HEOM=False):
    """heom.py:298-443 (lil assembly, without the aki_temp_corr branch).

    ``ham`` is the rotating-frame Hamiltonian of the model.
    """
    ados, inv = ado_mappings(ham.n_sites, K, level_cutoff)
    idx = liouville_subspace_index(liouville_subspace, hilbert_subspace, ham.n_sites)
    mesh = np.ix_(idx, idx)
    M = idx.size
    bath = ham.bath
    gamma, T, lam = bath.cutoff_freq, bath.temperature, bath.reorg_energy
    nu = matsubara_frequencies(K, gamma, T)
    c = corr_func_coeffs(K, gamma, T, lam, nu)
    nu_inf = matsubara_frequencies(K + 5000, gamma, T)
    c_inf = corr_func_coeffs(K + 5000, gamma, T, lam, nu_inf)
    Delta = np.sum((c_inf / nu_inf)[K + 1:])
    H = ham.H(hilbert_subspace)
    Lh = super_commutator_matrix(H)[mesh]
    PL = [super_left_matrix(p)[mesh] for p in ham.system_bath_couplings(hilbert_subspace)]
    PR = [super_right_matrix(p)[mesh] for p in ham.system_bath_couplings(hilbert_subspace)]
    L = lil_matrix((len(ados) * M, len(ados) * M), dtype=np.complex128)
    I = np.eye(M)
    sl = [slice(n * M, (n + 1) * M) for n in range(len(ados))]
    for n, m in enumerate(ados):
        blk = -1j * Lh - I * np.sum(m.dot(nu))
        if low_temp_corr:
            blk = blk - Delta * sum(pl + pr - 2 * pl.dot(pr) for pl, pr in zip(PL, PR))
        L[sl[n], sl[n]] = blk
        for (j, k), njk in np.ndenumerate(m):
            m[j, k] += 1
            p = inv.get(tuple(m.flatten()))
            m[j, k] -= 2
            q = inv.get(tuple(m.flatten()))
            m[j, k] += 1
            if p is not None:
                coef = np.sqrt((njk + 1) * np.abs(c[k])) if modified_HEOM else 1
                L[sl[n], sl[p]] = -1j * coef * (PL[j] - PR[j])
            if q is not None:
                coef = np.sqrt(njk / np.abs(c[k])) if modified_HEOM else njk
                L[sl[n], sl[q]] = -1j * coef * (c[k] * PL[j] - np.conj(c[k]) * PR[j])
    return L, ados


def heom_evolve_matrix(ham, hilbert_subspace, liouville_subspace, unit_convert,
                       K, level_cutoff, low_temp_corr=True, modified_HEOM=False,
                       heisenberg_picture=False):
    """heom.py:228-244: unit_convert * tensor, transpose for Heisenberg, CSR."""
    L, _ = heom_tensor(ham, hilbert_subspace, liouville_subspace, K, level_cutoff,
                       low_temp_corr, modified_HEOM)
    L = unit_convert * L
    if heisenberg_picture:
        L = L.T
    return csr_matrix(L)


class HEOMOperator(object):
    """heom.py:14-58: a Liouville-space operator applied to every ADO block
    of an ADO-major hierarchy vector; bra on ADO 0."""

    def __init__(self, lspace_op, ado_count):
        self.lspace_op = lspace_op
        self.ado_count = ado_count

    def commutator_matrix(self):
        return np.kron(np.eye(self.ado_count), self.lspace_op.commutator_matrix())

    def commutator(self, state):
        return self.commutator_matrix().dot(state)

    @property
    def bra_vector(self):
        b = self.lspace_op.bra_vector
        out = np.zeros(b.size * self.ado_count, dtype=complex)
        out[:b.size] = b
        return out


class HEOMOracleModel(object):
    """heom.py:177-296 as a model for the response drivers: the generator
    of heom_evolve_matrix, dipole operators mapped over ADOs, the thermal
    state on ADO 0 (the Liouville-space part is the unitary model's)."""

    def __init__(self, hamiltonian, rw_freq=None, hilbert_subspace='gef', unit_convert=1,
                 level_cutoff=3, K=1, low_temp_corr=True, modified_HEOM=False):
        self.lspace = LiouvilleModel(hamiltonian, rw_freq, hilbert_subspace, unit_convert,
                                     kind='unitary')
        self.hamiltonian = self.lspace.hamiltonian
        self.rw_freq = self.lspace.rw_freq
        self.hilbert_subspace = hilbert_subspace
        self.unit_convert = unit_convert
        self.level_cutoff, self.K = level_cutoff, K
        self.low_temp_corr, self.modified_HEOM = low_temp_corr, modified_HEOM
        self.ado_count = len(ado_mappings(self.hamiltonian.n_sites, K, level_cutoff)[0])

    @property
    def time_step(self):
        return self.lspace.time_step

    def evolve_matrix(self, subspace, heisenberg_picture=False):
        return heom_evolve_matrix(self.hamiltonian, self.hilbert_subspace, subspace,
                                  self.unit_convert, self.K, self.level_cutoff,
                                  self.low_temp_corr, self.modified_HEOM,
                                  heisenberg_picture).toarray()

    def dipole_operator(self, liouv_subspace_map, polarization, transitions='-+'):
        return HEOMOperator(self.lspace.dipole_operator(liouv_subspace_map, polarization,
                                                        transitions), self.ado_count)

    def thermal_state(self, liouville_subspace):
        rho = self.lspace.thermal_state(liouville_subspace)
        out = np.zeros(rho.size * self.ado_count, dtype=complex)
        out[:rho.size] = rho
        return out

    def map_between_subspaces(self, state, from_subspace, to_subspace):
        fs, ts = [self.lspace.liouville_subspace_index(s).size
                  for s in (from_subspace, to_subspace)]
        out = np.zeros(ts * self.ado_count, dtype=complex)
        for n in range(self.ado_count):
            out[n * ts:(n + 1) * ts] = self.lspace.map_between_subspaces(
                state[n * fs:(n + 1) * fs], from_subspace, to_subspace)
        return out
