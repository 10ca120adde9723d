"""Hierarchical equations of motion for a Drude bath
(qspectra/dynamics/heom.py API: HEOMModel, HEOMSpaceOperator,
matsubara_frequencies, corr_func_coeffs, ADO_mappings, multichoose).

The hierarchy is not assembled as a sparse matrix for propagation.  Its
structure is reduced to tables the GPU kernel (qsx_heom_propagate) reads:
  * per ADO n: the damping shift -sum_jk m_jk nu_k,
  * per Liouville element p = (r, c): the low-temperature correction
    -Delta sum_j (n_j(r) + n_j(c) - 2 n_j(r) n_j(c)),
  * per ADO: incoming edges (source ADO, weight vector, coefficient) for the
    +1 neighbours, -i [P_j, .], and the -1 neighbours,
    -i m_jk (c_k P_j^L - c_k^* P_j^R) (heom.py:410-442), both diagonal in p in
    the site basis; the Heisenberg picture uses the transposed edges,
  * the same edges binned by (ADO, bin (j, k), kind) -- the form the kernel
    reads -- and, per element p, the sites occupied in its row or column
    state: the weights of every other site vanish, so the kernel walks only
    those bins,
  * per sample: the system Hamiltonian block for -i [H, .].
``HEOM_tensor`` assembles the packed edges into the reference's sparse
matrix, which pins the tables on the host; the binned form is checked
against the packed one.
"""
import numpy as np
import torch
from scipy.sparse import csr_matrix, coo_matrix

from .. import _native
from ..operator_tools import full_liouville_subspace
from .base import DynamicalModel, SystemOperator
from .liouville_space import (LiouvilleSpaceOperator, super_commutator_matrix, _block_starts)
from .unitary import UnitaryModel


# Liouville blocks of at most this many elements take the dense -i[H, .]
# block in the stage kernel; single blocks XY above it take the commutator
# form with H_X and H_Y (the square blocks XX always do)
DENSE_BLOCK_MAX = 64


def matsubara_frequencies(K, gamma, T):
    """nu_0 = gamma (Drude pole), nu_k = 2 pi k T (Matsubara), k = 1..K."""
    v = 2 * np.pi * T * np.arange(K + 1)
    v[0] = gamma
    return v


def corr_func_coeffs(K, gamma, T, reorg_en, matsu_freqs, aki_temp_corr=False):
    """c_0 = lambda gamma (cot(gamma/2T) - i), c_k = 4 lambda gamma T nu_k /
    (nu_k^2 - gamma^2): C(t) = sum_k c_k exp(-nu_k t) (J. Chem. Phys. 130,
    234111)."""
    if aki_temp_corr:
        c0 = reorg_en * gamma * (2 * T / gamma - 1j)
    else:
        c0 = reorg_en * gamma * (1 / np.tan(gamma / (2 * T)) - 1j)
    nu = np.asarray(matsu_freqs, dtype=float)
    ck = 4 * reorg_en * gamma * T * nu[1:K + 1] / (nu[1:K + 1] ** 2 - gamma ** 2)
    return [c0] + list(ck)


def multichoose(n, c):
    """All ways to put c identical balls in n bins, in the recursive order
    of the reference (first bin empty first, then with one more ball)."""
    if c < 0 or n < 0:
        raise ValueError('multichoose needs nonnegative arguments')
    if c == 0:
        return [[0] * n]
    if n == 0:
        return []
    if n == 1:
        return [[c]]
    return ([[0] + v for v in multichoose(n - 1, c)]
            + [[v[0] + 1] + v[1:] for v in multichoose(n, c - 1)])


def ADO_mappings(N, K, level_cutoff):
    """ADOs: all N x (K+1) nonnegative integer matrices with total below
    ``level_cutoff``, level by level in multichoose order.  Returns the list
    of matrices and a matrix -> index function (None when absent)."""
    perms = []
    for c in range(level_cutoff):
        perms.extend(multichoose(N * (K + 1), c))
    lookup = {tuple(v): i for i, v in enumerate(perms)}

    def mat_to_ind(mat):
        return lookup.get(tuple(np.asarray(mat).flatten()))
    return [np.array(v).reshape(N, K + 1) for v in perms], mat_to_ind


def _occupations(hilbert_subspace, n_sites):
    """occ[s, j] = number of excitations of site j in Hilbert state s."""
    rows = []
    if 'g' in hilbert_subspace:
        rows.append(np.zeros(n_sites))
    if 'e' in hilbert_subspace:
        rows.extend(np.eye(n_sites))
    if 'f' in hilbert_subspace:
        for i in range(n_sites):
            for j in range(i + 1, n_sites):
                v = np.zeros(n_sites)
                v[i] = v[j] = 1
                rows.append(v)
    return np.array(rows)


class HEOMSpaceOperator(SystemOperator):
    """A Liouville-space operator applied to every ADO of a hierarchy; the
    expectation value reads ADO 0 (heom.py:22-58)."""

    def __init__(self, operator, liouv_subspace_map, dynamical_model):
        self.lspace_op = LiouvilleSpaceOperator(operator, liouv_subspace_map,
                                                dynamical_model.lspace_model)
        self.ado_count = dynamical_model.ado_count
        self.to_size = self.lspace_op.to_indices.size
        self.from_size = self.lspace_op.from_indices.size

    @property
    def bra_vector(self):
        bra = np.zeros(self.from_size * self.ado_count, dtype=complex)
        bra[:self.from_size] = self.lspace_op.bra_vector
        return bra

    def _map(self, matrix, state):
        s = np.asarray(state).reshape(self.ado_count, self.from_size)
        return (s @ np.asarray(matrix).T).reshape(-1)

    def left_multiply(self, state):
        return self._map(self.lspace_op.left_matrix(), state)

    def right_multiply(self, state):
        return self._map(self.lspace_op.right_matrix(), state)

    def commutator(self, state):
        return self._map(self.lspace_op.left_matrix() - self.lspace_op.right_matrix(), state)

    def expectation_value(self, state):
        return self.lspace_op.expectation_value(np.asarray(state)[:self.from_size])


class DeviceHEOM(object):
    """eom(t, y) of a hierarchy with its tables resident on the GPU."""

    def __init__(self, tables):
        self.__dict__.update(tables)

    @property
    def batch(self):
        return self.Hop.shape[0]

    @property
    def size(self):
        return self.n_ado * self.M


    def __call__(self, t, y):
        y = np.asarray(y, dtype=complex)
        yt = _native.to_device(y.reshape(self.batch, -1))
        out = _native.heom_apply(self, yt)
        return out.cpu().numpy().reshape(y.shape)

    def repeated(self, R):
        """The same hierarchy with every sample repeated R times (sample-
        major), so that R right-hand sides per sample propagate as batch
        entries of one launch (the tables are shared; only the per-sample
        operator block is copied)."""
        new = DeviceHEOM(dict(self.__dict__))
        new.Hop = self.Hop.repeat_interleave(int(R), dim=0).contiguous()
        return new

    def dense_generator(self):
        """The hierarchy generator as dense matrices [B, S, S] (S = n_ado *
        M), column j = the right-hand side applied to unit vector j, from
        one application launch over all S unit vectors of every sample."""
        B, S = self.batch, self.size
        eye = torch.eye(S, dtype=torch.complex128, device=self.Hop.device)
        cols = _native.heom_apply(self.repeated(S), eye.repeat(B, 1).contiguous())
        return cols.reshape(B, S, S).transpose(1, 2).contiguous()


class HEOMModel(DynamicalModel):
    """HEOM for an identical Drude bath on every site with K Matsubara terms,
    truncated at ``level_cutoff`` (exclusive: ADOs of total level <
    level_cutoff), with the low-temperature correction (Ishizaki & Tanimura)
    and the optional rescaled ("modified") hierarchy."""
    system_operator = HEOMSpaceOperator

    def __init__(self, hamiltonian, rw_freq=None, hilbert_subspace='gef', unit_convert=1,
                 level_cutoff=3, K=1, low_temp_corr=True, modified_HEOM=False,
                 aki_temp_corr=False):
        super(HEOMModel, self).__init__(hamiltonian, rw_freq, hilbert_subspace, unit_convert)
        if aki_temp_corr:
            raise NotImplementedError('aki_temp_corr is not supported')
        if getattr(self.hamiltonian, 'n_vibrational_states', 1) != 1:
            raise NotImplementedError('the HEOM path takes electronic Hamiltonians '
                                      '(site occupations); vibronic ones are not supported')
        if modified_HEOM and not low_temp_corr:
            raise ValueError('modified_HEOM requires low_temp_corr')
        self.lspace_model = UnitaryModel(hamiltonian, rw_freq, hilbert_subspace, unit_convert)
        self.level_cutoff = level_cutoff
        self.K = K
        self.low_temp_corr = low_temp_corr
        self.modified_HEOM = modified_HEOM
        self.aki_temp_corr = aki_temp_corr
        self.ado_indices, self.mat_to_ind = ADO_mappings(self.hamiltonian.n_sites, K,
                                                         level_cutoff)
        self.ado_count = len(self.ado_indices)

    # ------------------------------------------------------------ structure
    def _bath(self):
        b = self.hamiltonian.bath
        return b.cutoff_freq, b.temperature, b.reorg_energy

    def _coefficients(self):
        gamma, T, lam = self._bath()
        nu = matsubara_frequencies(self.K, gamma, T)
        c = np.array(corr_func_coeffs(self.K, gamma, T, lam, nu), dtype=complex)
        nu_inf = matsubara_frequencies(self.K + 5000, gamma, T)
        c_inf = np.array(corr_func_coeffs(self.K + 5000, gamma, T, lam, nu_inf), dtype=complex)
        delta = np.sum((c_inf / nu_inf)[self.K + 1:])
        return nu, c, delta

    def _subspace_sites(self, liouville_subspace):
        n = self.hamiltonian.n_sites
        idx = self.lspace_model.liouville_subspace_index(liouville_subspace)
        _, N = _block_starts(self.hilbert_subspace, n)
        occ = _occupations(self.hilbert_subspace, n)
        return occ[idx % N], occ[idx // N], idx

    def tables(self, liouville_subspace, heisenberg_picture=False):
        """Host tables of the hierarchy on a Liouville subspace."""
        N, K = self.hamiltonian.n_sites, self.K
        uc = self.unit_convert
        nu, c, delta = self._coefficients()
        occ_r, occ_c, idx = self._subspace_sites(liouville_subspace)
        M = idx.size
        ados = np.array([a.flatten() for a in self.ado_indices])
        lookup = {tuple(a): i for i, a in enumerate(ados)}
        bins = N * (K + 1)
        W = [occ_r[:, j] - occ_c[:, j] for j in range(N)]
        W += [c[k] * occ_r[:, j] - np.conj(c[k]) * occ_c[:, j]
              for j in range(N) for k in range(K + 1)]
        W = np.array(W, dtype=complex)
        shift = -uc * (ados.reshape(-1, N, K + 1) @ nu).sum(axis=1)
        D = (occ_r + occ_c - 2 * occ_r * occ_c).sum(axis=1)
        dvec = -uc * delta * D if self.low_temp_corr else np.zeros(M, complex)
        edges = [[] for _ in range(len(ados))]
        # the same edges binned by (ADO, bin b, kind): kind 0 couples to n + e_b
        # with weight W_j, kind 1 to n - e_b with weight W_{N+b}; every
        # (ADO, bin, kind) holds at most one edge, also after transposition
        bsrc = -np.ones((len(ados), bins, 2), dtype=np.int32)
        bcoef = np.zeros((len(ados), bins, 2), dtype=complex)

        def binned(n, s, b, kind, cf):
            if heisenberg_picture:
                bsrc[s, b, kind], bcoef[s, b, kind] = n, cf
            else:
                bsrc[n, b, kind], bcoef[n, b, kind] = s, cf

        for n, m in enumerate(ados):
            for b in range(bins):
                j, k = divmod(b, K + 1)
                up = m.copy()
                up[b] += 1
                p = lookup.get(tuple(up))
                if p is not None:
                    cf = (np.sqrt((m[b] + 1) * abs(c[k])) if self.modified_HEOM else 1.0)
                    edges[n].append((p, j, -1j * uc * cf))
                    binned(n, p, b, 0, -1j * uc * cf)
                if m[b] > 0:
                    dn = m.copy()
                    dn[b] -= 1
                    q = lookup[tuple(dn)]
                    cf = (np.sqrt(m[b] / abs(c[k])) if self.modified_HEOM else m[b])
                    edges[n].append((q, N + j * (K + 1) + k, -1j * uc * cf))
                    binned(n, q, b, 1, -1j * uc * cf)
        occ = occ_r + occ_c
        nrel = max(1, int((occ > 0).sum(axis=1).max()))
        rel_site = -np.ones((M, nrel), dtype=np.int32)
        rel_occ = np.zeros((M, nrel, 2))
        for p_ in range(M):
            js = np.flatnonzero(occ[p_] > 0)
            rel_site[p_, :js.size] = js
            rel_occ[p_, :js.size, 0] = occ_r[p_, js]
            rel_occ[p_, :js.size, 1] = occ_c[p_, js]
        if heisenberg_picture:
            t_edges = [[] for _ in range(len(ados))]
            for n, lst in enumerate(edges):
                for (s, w, cf) in lst:
                    t_edges[s].append((n, w, cf))
            edges = t_edges
        E = max(1, max(len(l) for l in edges))
        esrc = -np.ones((len(ados), E), dtype=np.int32)
        ew = np.zeros((len(ados), E), dtype=np.int32)
        ecoef = np.zeros((len(ados), E), dtype=complex)
        for n, lst in enumerate(edges):
            for e, (s, w, cf) in enumerate(lst):
                esrc[n, e], ew[n, e], ecoef[n, e] = s, w, cf
        return dict(n_ado=len(ados), M=M, shift=shift, dvec=dvec.astype(complex), W=W, E=E,
                    esrc=esrc, ew=ew, ecoef=ecoef, idx=idx, nbins=bins, K1=K + 1,
                    bsrc=bsrc, bcoef=bcoef, ck=np.asarray(c[:K + 1], dtype=complex),
                    nrel=nrel, rel_site=rel_site, rel_occ=rel_occ)

    def _hilbert_tables(self, idx, N):
        """Device tables of the Hilbert-commutator form: the row and column
        state of every element, the element of every (row, column) pair (-1
        outside the subspace) and each state's block [first, end) of equal
        excitation number; checked here so that every element the kernel
        reads is in the subspace."""
        rows, cols = idx % N, idx // N
        pos = -np.ones((N, N), dtype=np.int32)
        pos[rows, cols] = np.arange(idx.size, dtype=np.int32)
        exc = _occupations(self.hilbert_subspace, self.hamiltonian.n_sites).sum(axis=1)
        blk = np.zeros((N, 2), dtype=np.int32)
        for a in range(N):
            same = np.flatnonzero(exc == exc[a])
            if not np.array_equal(same, np.arange(same[0], same[-1] + 1)):
                raise NotImplementedError('Hilbert states not grouped by excitation number')
            blk[a] = (same[0], same[-1] + 1)
        for a, b in zip(rows, cols):
            if np.any(pos[blk[a, 0]:blk[a, 1], b] < 0) or np.any(pos[a, blk[b, 0]:blk[b, 1]] < 0):
                raise ValueError('the subspace must consist of whole Liouville blocks')
        est = np.stack([rows, cols], axis=1).astype(np.int32)
        dev = _native.device()
        return dict(N=int(N), est=torch.as_tensor(est).to(dev), pos=torch.as_tensor(pos).to(dev),
                    blk=torch.as_tensor(blk).to(dev))

    def _single_block(self, liouville_subspace):
        """Hilbert manifolds (X, Y) when the subspace is the single block
        'XY'."""
        subs = liouville_subspace.split(',')
        if len(subs) == 1:
            return subs[0][0], subs[0][1]
        return None

    def HEOM_tensor(self, liouville_subspace, heisenberg_picture=False):
        """The hierarchy generator as a sparse matrix (without unit_convert),
        assembled from the propagation tables (heom.py:298-443 layout)."""
        tb = self.tables(liouville_subspace, heisenberg_picture)
        uc = self.unit_convert
        idx, M, n_ado = tb['idx'], tb['M'], tb['n_ado']
        H = self.hamiltonian.H(self.hilbert_subspace)
        Lh = -1j * super_commutator_matrix(H)[np.ix_(idx, idx)]
        if heisenberg_picture:
            Lh = Lh.T
        rows, cols, vals = [], [], []
        blk = coo_matrix(Lh)
        for n in range(n_ado):
            rows.append(blk.row + n * M)
            cols.append(blk.col + n * M)
            vals.append(blk.data)
            d = (tb['shift'][n] + tb['dvec']) / uc
            rows.append(np.arange(M) + n * M)
            cols.append(np.arange(M) + n * M)
            vals.append(d)
            for s, w, cf in zip(tb['esrc'][n], tb['ew'][n], tb['ecoef'][n]):
                if s < 0:
                    break
                rows.append(np.arange(M) + n * M)
                cols.append(np.arange(M) + s * M)
                vals.append(cf * tb['W'][w] / uc)
        L = coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                       shape=(n_ado * M, n_ado * M)).tocsr()
        L.eliminate_zeros()
        return L

    def _members(self):
        return getattr(self, '_batch', None) or [self.hamiltonian]

    _hamiltonians = _members

    def with_hamiltonians(self, hamiltonians):
        from ..utils import copy_with_new_cache
        new = copy_with_new_cache(self)
        new.hamiltonian = hamiltonians[0]
        new._batch = list(hamiltonians)
        return new

    def equation_of_motion(self, liouville_subspace, heisenberg_picture=False):
        key = (liouville_subspace, bool(heisenberg_picture))
        cache = self.__dict__.setdefault('_eoms', {})
        if key in cache:
            return cache[key]
        tb = self.tables(liouville_subspace, heisenberg_picture)
        dev = _native.device()
        XY = self._single_block(liouville_subspace)
        hams = self._members()
        uc = self.unit_convert
        hilbert = None
        if XY is not None and (tb['M'] > DENSE_BLOCK_MAX or XY[0] == XY[1]):
            # commutator form: H_X and H_Y blocks of every member
            blocks = []
            for Z in (XY[:1] if XY[0] == XY[1] else XY):     # square: H_Y = H_X
                sl = self.hamiltonian.hilbert_subspace_index(Z, self.hilbert_subspace)
                Hz = np.stack([h.H(self.hilbert_subspace)[sl, sl] for h in hams])
                if heisenberg_picture:
                    Hz = Hz.transpose(0, 2, 1)
                blocks.append(Hz.reshape(len(hams), -1))
            Hop = _native.to_device(-1j * uc * np.concatenate(blocks, axis=1))
            n_comm = int(round(np.sqrt(blocks[0].shape[1])))
        elif tb['M'] > DENSE_BLOCK_MAX:
            # several Liouville blocks beyond the dense form: the system part
            # as a commutator with the Hilbert-space Hamiltonian
            Hh = np.stack([h.H(self.hilbert_subspace) for h in hams])
            if heisenberg_picture:
                Hh = Hh.transpose(0, 2, 1)
            Hop = _native.to_device(-1j * uc * Hh.reshape(len(hams), -1))
            n_comm = 0
            hilbert = self._hilbert_tables(tb['idx'], Hh.shape[-1])
        else:
            lm = self.lspace_model.with_hamiltonians(hams) if len(hams) > 1 else self.lspace_model
            Hop = lm.liouvillian(liouville_subspace, heisenberg_picture).contiguous()
            n_comm = 0
        eom = DeviceHEOM(dict(
            n_ado=tb['n_ado'], M=tb['M'], n_comm=n_comm, Hop=Hop.contiguous(),
            shift=torch.as_tensor(tb['shift'], dtype=torch.float64).to(dev),
            dvec=_native.to_device(tb['dvec']), E=tb['E'],
            nbins=tb['nbins'], K1=tb['K1'], bsrc=torch.as_tensor(tb['bsrc']).to(dev),
            bcoef=_native.to_device(tb['bcoef']), ck=_native.to_device(tb['ck']),
            nrel=tb['nrel'], rel_site=torch.as_tensor(tb['rel_site']).to(dev),
            rel_occ=torch.as_tensor(tb['rel_occ'], dtype=torch.float64).to(dev),
            liouville_subspace=liouville_subspace, heisenberg_picture=heisenberg_picture,
            hilbert=hilbert))
        cache[key] = eom
        return eom

    # ---------------------------------------------------------- state maps
    def thermal_state(self, liouville_subspace):
        rho0 = self.lspace_model.thermal_state(liouville_subspace)
        rho = np.zeros(rho0.size * self.ado_count, dtype=complex)
        rho[:rho0.size] = rho0
        return rho

    def dipole_operator(self, liouv_subspace_map, polarization, transitions='-+'):
        op = self.hamiltonian.dipole_operator(self.hilbert_subspace, polarization, transitions)
        return HEOMSpaceOperator(op, liouv_subspace_map, self)

    def map_between_subspaces(self, state, from_subspace, to_subspace):
        fs, ts = [self.lspace_model.liouville_subspace_index(s).size
                  for s in (from_subspace, to_subspace)]
        s = np.asarray(state).reshape(self.ado_count, fs)
        return np.concatenate([self.lspace_model.map_between_subspaces(v, from_subspace,
                                                                       to_subspace) for v in s])

    def density_matrix_to_state_vector(self, rho0, liouville_subspace):
        s0 = self.lspace_model.density_matrix_to_state_vector(rho0, liouville_subspace)
        out = np.zeros(s0.size * self.ado_count, dtype=complex)
        out[:s0.size] = s0
        return out

    def state_vector_to_density_matrix(self, rhos):
        """Reduced density matrices (ADO 0) of hierarchy states.  (The
        reference slices with a float here, heom.py:293-294.)"""
        rhos = np.asarray(rhos)
        m = rhos.shape[-1] // self.ado_count
        return self.lspace_model.state_vector_to_density_matrix(rhos[..., :m])
