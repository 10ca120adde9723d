"""The HEOM propagation tables (host side, no GPU) assemble into exactly the
sparse generator the reference builds (golden CSR matrices written by
qspectra's HEOMModel.HEOM_tensor, heom.py:298-443), for the plain, modified
and no-low-temperature-correction hierarchies, square and multi-block
subspaces, and in the Heisenberg picture (transpose)."""
import numpy as np
import pytest
from numpy.testing import assert_equal

import qspectra_amd as qa
from qspectra_amd.dynamics import heom as H
from oracle import liouville as Lv
from tests.conftest import golden, golden_csr


def _ham(name, T):
    g = golden(name)
    return g, qa.ElectronicHamiltonian(g['in_H1'], bath=qa.DebyeBath(T, 35, 106),
                                       dipoles=g['in_dipoles'],
                                       energy_spread_extra=float(g['extra']))


def test_ado_enumeration_matches_reference():
    assert H.multichoose(3, 2) == Lv.multichoose(3, 2)
    for N, K, L in [(2, 1, 3), (7, 1, 3), (3, 2, 4)]:
        ados, m2i = H.ADO_mappings(N, K, L)
        ref, _ = Lv.ado_mappings(N, K, L)
        assert_equal(np.array([a.flatten() for a in ados]), np.array([a.flatten() for a in ref]))
        assert m2i(ados[-1]) == len(ados) - 1 and m2i(np.full((N, K + 1), 99)) is None


@pytest.mark.parametrize('name,kw', [('dimer', dict(level_cutoff=3, K=1)),
                                     ('dimer_mod', dict(level_cutoff=3, K=1, modified_HEOM=True)),
                                     ('dimer_nolt', dict(level_cutoff=3, K=2,
                                                         low_temp_corr=False))])
def test_heom_tensor_matches_reference(name, kw):
    gh = golden('heom')
    _, ham = _ham('redfield_dimer', qa.CM_K * 300)
    m = qa.HEOMModel(ham, rw_freq=12800.0, hilbert_subspace='ge', unit_convert=qa.CM_FS, **kw)
    assert_equal(np.array([a.flatten() for a in m.ado_indices]), gh[name + '_ados'])
    for s in ('ee', 'gg,ge,eg,ee', 'eg'):
        ref = golden_csr(gh, '%s_%s' % (name, s.replace(',', '_')))
        L = qa.CM_FS * m.HEOM_tensor(s)
        assert abs(L - ref).max() < 1e-13 * abs(ref).max(), (name, s)
        Lt = qa.CM_FS * m.HEOM_tensor(s, heisenberg_picture=True)
        assert abs(Lt - ref.T).max() < 1e-13 * abs(ref).max(), (name, s, 'heisenberg')


def test_heom_tensor_7site():
    gh = golden('heom')
    g, ham = _ham('redfield_7site', qa.CM_K * 300)
    m = qa.HEOMModel(ham, rw_freq=float(gh['agg7_rw_freq']), hilbert_subspace='ge',
                     unit_convert=qa.CM_FS, level_cutoff=3, K=1)
    ref = golden_csr(gh, 'agg7_ee')
    L = qa.CM_FS * m.HEOM_tensor('ee')
    assert abs(L - ref).max() < 1e-13 * abs(ref).max()
    # the workload hierarchy: depth 4, K = 1 -> C(14 + 4, 4) = 3060 ADOs, <= 28 edges
    m5 = qa.HEOMModel(ham, rw_freq=float(gh['agg7_rw_freq']), hilbert_subspace='ge',
                      unit_convert=qa.CM_FS, level_cutoff=5, K=1)
    assert m5.ado_count == 3060
    tb = m5.tables('ee')
    assert tb['E'] <= 28 and tb['M'] == 49


def _coupling_from_edges(tb):
    """Dense hierarchy coupling (neighbour terms only) from the packed edges."""
    n, M = tb['n_ado'], tb['M']
    C = np.zeros((n * M, n * M), complex)
    for a in range(n):
        for s, w, cf in zip(tb['esrc'][a], tb['ew'][a], tb['ecoef'][a]):
            if s < 0:
                break
            C[a * M:(a + 1) * M, s * M:(s + 1) * M] += np.diag(cf * tb['W'][w])
    return C


def _coupling_from_bins(tb):
    """The same from the binned tables and the per-element site lists, with the
    weights recomputed from the occupations as the GPU kernel does."""
    n, M, K1 = tb['n_ado'], tb['M'], tb['K1']
    C = np.zeros((n * M, n * M), complex)
    for a in range(n):
        for p in range(M):
            for t, j in enumerate(tb['rel_site'][p]):
                if j < 0:
                    continue
                o_r, o_c = tb['rel_occ'][p, t]
                for k in range(K1):
                    b = j * K1 + k
                    for kind, w in ((0, o_r - o_c),
                                    (1, tb['ck'][k] * o_r - np.conj(tb['ck'][k]) * o_c)):
                        s = tb['bsrc'][a, b, kind]
                        if s >= 0:
                            C[a * M + p, s * M + p] += tb['bcoef'][a, b, kind] * w
    return C


@pytest.mark.parametrize('kw', [dict(level_cutoff=3, K=1), dict(level_cutoff=3, K=2),
                                dict(level_cutoff=3, K=1, modified_HEOM=True)])
@pytest.mark.parametrize('subspace', ['ee', 'gg,ge,eg,ee', 'eg', 'fe', 'ff'])
@pytest.mark.parametrize('heisenberg', [False, True])
def test_binned_tables_equal_packed_edges(kw, subspace, heisenberg):
    """The GPU's binned neighbour tables (at most one edge per ADO, bin and
    kind; weights from the occupations of the element's row and column
    states, only for the sites they occupy) give the same coupling as the
    packed edge lists that assemble to the reference's generator."""
    _, ham = _ham('redfield_dimer', qa.CM_K * 300)
    hs = 'gef' if 'f' in subspace else 'ge'
    m = qa.HEOMModel(ham, rw_freq=12800.0, hilbert_subspace=hs, unit_convert=qa.CM_FS, **kw)
    tb = m.tables(subspace, heisenberg)
    A, B = _coupling_from_edges(tb), _coupling_from_bins(tb)
    assert np.abs(A - B).max() <= 1e-14 * np.abs(A).max()
    assert np.count_nonzero(A) == np.count_nonzero(B)
