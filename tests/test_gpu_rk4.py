"""Parity of the fused RK4 kernel (qsx_rk4_dense) and the Liouvillian apply
kernel with the oracle, called through the C ABI.

Liouvillians and dipole maps come from the oracle (which reproduces the
reference's to 1e-12, tests/test_oracle_liouville.py); the kernels are
checked against the oracle's fixed-step RK4 on the same schedule.
FP64 mode: 1e-12 relative (max-norm over the whole output).  MIXED mode:
2e-6 relative.
"""
import numpy as np
import pytest
import torch

from oracle import system as S
from oracle import liouville as Lv
from oracle import propagate as P
from tests.conftest import golden, rel_err

pytestmark = pytest.mark.gpu


def _agg7_model(secular=False):
    g = golden('redfield_7site')
    ham = S.ElectronicHamiltonian(g['in_H1'], bath=S.DebyeBath(S.CM_K * 77, 35, 106),
                                  dipoles=g['in_dipoles'], energy_spread_extra=float(g['extra']))
    return Lv.LiouvilleModel(ham, unit_convert=S.CM_FS, secular=secular)


def _dimer_model():
    g = golden('redfield_dimer')
    ham = S.ElectronicHamiltonian(g['in_H1'], bath=S.DebyeBath(S.CM_K * 77, 35, 106),
                                  dipoles=g['in_dipoles'], energy_spread_extra=float(g['extra']))
    return Lv.LiouvilleModel(ham, unit_convert=S.CM_FS)


def _run(L, y0, t, t0=None, proj=None, dt=None, mixed=False, B=None):
    from qspectra_amd import _native, engine
    dev = _native.device()
    nsub, h = engine.schedule_tensors(t, t0, dt, dev)
    L = np.asarray(L)
    if L.ndim == 2:
        L = L[None]
    y0 = np.asarray(y0, complex)
    if y0.ndim == 1:
        y0 = y0[None, None]
    elif y0.ndim == 2:
        y0 = y0[None]
    if mixed:
        d = np.ascontiguousarray(np.diagonal(L, axis1=1, axis2=2))
        Loff = L.copy()
        idx = np.arange(L.shape[-1])
        Loff[:, idx, idx] = 0
        Ld = _native.to_device(Loff, torch.complex64)
        dd = _native.to_device(d, torch.complex128)
        pj = None if proj is None else _native.to_device(np.asarray(proj)[None] if np.ndim(proj) == 2
                                                         else proj, torch.complex64)
        out, _, st = _native.rk4_dense(Ld, _native.to_device(y0), nsub, h, pj,
                                       _native.QSX_MIXED, dd)
    else:
        pj = None if proj is None else _native.to_device(np.asarray(proj)[None] if np.ndim(proj) == 2
                                                         else proj)
        out, _, st = _native.rk4_dense(_native.to_device(L), _native.to_device(y0), nsub, h, pj)
    torch.cuda.synchronize()
    assert int(st.item()) == 0
    return out.cpu().numpy()


def test_apply_dense_matches_numpy(gpu):
    from qspectra_amd import _native
    rs = np.random.RandomState(0)
    for B, R, M in [(1, 1, 2), (3, 5, 49), (2, 1, 147)]:
        L = rs.randn(B, M, M) + 1j * rs.randn(B, M, M)
        y = rs.randn(B, R, M) + 1j * rs.randn(B, R, M)
        out = _native.apply_dense(_native.to_device(L), _native.to_device(y)).cpu().numpy()
        assert rel_err(out, np.einsum('bij,brj->bri', L, y)) < 1e-14
    out = _native.apply_dense(_native.to_device(L[:1]), _native.to_device(y)).cpu().numpy()
    assert rel_err(out, np.einsum('ij,brj->bri', L[0], y)) < 1e-14


def test_dimer_absorption_trajectory_fp64(gpu):
    m = _dimer_model()
    Lh = m.evolve_matrix('eg', heisenberg_picture=True)
    V = m.dipole_operator('eg->gg', 'x', '-')
    y0 = -V.bra_vector
    t = np.arange(0, 2048, m.time_step)
    ref = P.integrate_rk4(Lh, y0, t)
    got = _run(Lh, y0, t)[0, 0]
    assert got.shape == ref.shape == (2048, 2)
    assert rel_err(got, ref) < 1e-12
    got = _run(Lh, y0, t, mixed=True)[0, 0]
    assert rel_err(got, ref) < 2e-6


def test_population_stage_multi_rhs_with_projection(gpu):
    """t2 stage of the 7-site third-order response: 128 RHS share the 'ee'
    Liouvillian; saves project with the 'ee->eg' / 'ee->fe' commutators."""
    m = _agg7_model()
    L = m.evolve_matrix('ee')
    rs = np.random.RandomState(1)
    y0 = rs.randn(128, 49) + 1j * rs.randn(128, 49)
    t2 = np.array([0., 100., 200., 500.])
    for sub in ('eg', 'fe'):
        V = m.dipole_operator('ee->' + sub, 'x', '+')
        Pm = V.commutator_matrix()
        ref = P.integrate_rk4(L, y0, t2, t0=0, save_matrix=Pm, dt=2.0)
        got = _run(L, y0, t2, t0=0, proj=Pm, dt=2.0)[0]
        assert got.shape == ref.shape
        assert rel_err(got, ref) < 1e-12, sub
        got = _run(L, y0, t2, t0=0, proj=Pm, dt=2.0, mixed=True)[0]
        assert rel_err(got, ref) < 2e-6, sub


def test_large_block_from_global_memory(gpu):
    """'fe' Heisenberg block (M = 147) does not fit LDS: read from L2/HBM."""
    m = _agg7_model()
    Lh = m.evolve_matrix('fe', heisenberg_picture=True)
    y0 = m.dipole_operator('fe->ee', 'x', '-').bra_vector
    t = np.arange(0, 254.5, 2.0)
    ref = P.integrate_rk4(Lh, y0, t)
    assert rel_err(_run(Lh, y0, t)[0, 0], ref) < 1e-12
    assert rel_err(_run(Lh, y0, t, mixed=True)[0, 0], ref) < 2e-6


def test_batched_disorder_members(gpu):
    m = _agg7_model()
    Ls, ys = [], []
    for n in range(16):
        mm = m.sample(n)
        Ls.append(mm.evolve_matrix('eg', heisenberg_picture=True))
        ys.append(-mm.dipole_operator('eg->gg', 'y', '-').bra_vector)
    Ls, ys = np.array(Ls), np.array(ys)[:, None, :]
    t = np.arange(0, 1024.5, 2.0)
    got = _run(Ls, ys, t)
    for n in range(16):
        ref = P.integrate_rk4(Ls[n], ys[n, 0], t)
        assert rel_err(got[n, 0], ref) < 1e-12, n


def test_non_finite_state_sets_status(gpu):
    from qspectra_amd import _native, engine
    L = np.eye(3) * (-1j)
    y0 = np.array([[[1.0, np.nan, 0]]], complex)
    nsub, h = engine.schedule_tensors(np.arange(4.0))
    out, _, st = _native.rk4_dense(_native.to_device(L[None]), _native.to_device(y0), nsub, h)
    torch.cuda.synchronize()
    assert int(st.item()) == 1


@pytest.mark.parametrize('M', [1, 2, 7, 13, 24, 49, 100, 147])
def test_register_stepper_all_shapes(gpu, M):
    """qsx_propagate in RK4 and propagator modes, FP64 and MIXED, for every
    row-segment shape (M = 147 keeps 19 rows in LDS)."""
    from qspectra_amd import _native, engine, lti
    rs = np.random.RandomState(M)
    L = (rs.randn(2, M, M) + 1j * rs.randn(2, M, M)) * 0.02 - 0.05j * np.eye(M)
    y0 = rs.randn(2, 3, M) + 1j * rs.randn(2, 3, M)
    t = np.arange(0, 40.0, 1.0)
    nsub, h = engine.schedule_tensors(t, 0.0, 0.5)
    ref = np.stack([P.integrate_rk4(L[b], y0[b], t, t0=0.0, dt=0.5) for b in range(2)])
    Ld = _native.to_device(L)
    y = _native.to_device(y0)
    out, _, st = _native.propagate_states(_native.QSX_MODE_RK4, Ld, y, nsub, h)
    assert int(st.item()) == 0
    assert rel_err(out.cpu().numpy(), ref) < 1e-12, 'rk4 fp64'
    off, dg = lti.split_diagonal(Ld)
    out, _, _ = _native.propagate_states(_native.QSX_MODE_RK4, off, y, nsub, h,
                                         _native.QSX_MIXED, dg)
    assert rel_err(out.cpu().numpy(), ref) < 2e-6, 'rk4 mixed'
    Pm = lti.rk4_propagator(Ld, 0.5)
    out, _, _ = _native.propagate_states(_native.QSX_MODE_PROPAGATOR, Pm, y, nsub, h)
    assert rel_err(out.cpu().numpy(), ref) < 1e-11, 'propagator fp64'
    off, dg = lti.split_diagonal(lti.rk4_increment(Ld, 0.5))
    out, _, _ = _native.propagate_states(_native.QSX_MODE_PROPAGATOR, off, y, nsub, h,
                                         _native.QSX_MIXED, dg)
    assert rel_err(out.cpu().numpy(), ref) < 2e-6, 'propagator mixed'


@pytest.mark.parametrize('M', [1, 7, 13, 24, 49, 100, 147])
def test_chunked_trajectory_one_launch_matches_two(gpu, M):
    """qsx_propagate_chunks (each work-group computes its own chunk start,
    then steps its chunk) against the chain of chunk starts followed by the
    chunks as right-hand sides (two qsx_propagate launches), FP64 and MIXED,
    shared and per-problem operators: the same products in the same order,
    so the same values; and against dense RK4 in FP64."""
    from qspectra_amd import _native, engine, lti
    rs = np.random.RandomState(100 + M)
    B, q, c, nc = 3, 2, 5, 4
    h = 0.5
    L = (rs.randn(B, M, M) + 1j * rs.randn(B, M, M)) * 0.02 - 0.05j * np.eye(M)
    y0 = rs.randn(B, 1, M) + 1j * rs.randn(B, 1, M)
    Ld = _native.to_device(L)
    y = _native.to_device(y0)
    Pm = lti.rk4_propagator(Ld, h)
    Pc = torch.linalg.matrix_power(Pm, q * c).contiguous()
    dev = y.device
    starts_n = torch.tensor([0] + [1] * (nc - 1), dtype=torch.int32, device=dev)
    starts_h = torch.full((nc,), h * q * c, dtype=torch.float64, device=dev)
    chunk_n = torch.tensor([0] + [q] * (c - 1), dtype=torch.int32, device=dev)
    chunk_h = torch.full((c,), h, dtype=torch.float64, device=dev)
    t = np.arange(nc * c) * (q * h)
    ref = np.stack([P.integrate_rk4(L[b], y0[b], t, t0=0.0, dt=h) for b in range(B)])
    for prec in (_native.QSX_FP64, _native.QSX_MIXED):
        if prec == _native.QSX_FP64:
            A, dg = Pm, None
        else:
            A, dg = lti.split_diagonal(lti.rk4_increment(Ld, h))
        for shared in (False, True):
            Pc_, A_ = (Pc[:1].contiguous(), A[:1].contiguous()) if shared else (Pc, A)
            dg_ = dg[:1].contiguous() if (shared and dg is not None) else dg
            st0, _, s0 = _native.propagate_states(_native.QSX_MODE_PROPAGATOR, Pc_, y,
                                                  starts_n, starts_h)
            starts = st0.reshape(B, nc, M).contiguous()
            two, _, s1 = _native.propagate_states(_native.QSX_MODE_PROPAGATOR, A_, starts,
                                                  chunk_n, chunk_h, prec, dg_)
            one, s2 = _native.propagate_chunks(Pc_, A_, y, nc, chunk_n, chunk_h, prec, dg_)
            assert int(s0.item()) == 0 and int(s1.item()) == 0 and int(s2.item()) == 0
            assert one.shape == (B, nc, c, M)
            assert rel_err(one.cpu().numpy(), two.cpu().numpy()) < 1e-14, (prec, shared)
            if prec == _native.QSX_FP64 and not shared:
                got = one.reshape(B, nc * c, M).cpu().numpy()
                assert rel_err(got, ref[:, 0]) < 1e-11


def test_zgemm_and_propagator(gpu):
    from qspectra_amd import _native, lti
    rs = np.random.RandomState(5)
    for (b, m, n, k) in [(1, 1, 1, 1), (3, 49, 128, 49), (2, 147, 49, 147), (1, 33, 65, 17)]:
        A = rs.randn(b, m, k) + 1j * rs.randn(b, m, k)
        B = rs.randn(b, k, n) + 1j * rs.randn(b, k, n)
        C0 = rs.randn(b, m, n) + 1j * rs.randn(b, m, n)
        C = _native.zgemm(_native.to_device(A), _native.to_device(B), _native.to_device(C0),
                          alpha=0.5 - 1j, beta=2.0, gamma=(0.25 if m == n else 0.0))
        ref = (0.5 - 1j) * A @ B + 2.0 * C0 + (0.25 * np.eye(m) if m == n else 0)
        assert rel_err(C.cpu().numpy(), ref) < 1e-14
        Cm = _native.zgemm(_native.to_device(A, torch.complex64),
                           _native.to_device(B, torch.complex64), precision=_native.QSX_MIXED)
        assert rel_err(Cm.cpu().numpy(), A @ B) < 2e-6
    A = rs.randn(1, 20, 20) + 1j * rs.randn(1, 20, 20)
    B = rs.randn(4, 20, 8) + 1j * rs.randn(4, 20, 8)
    C = _native.zgemm(_native.to_device(A), _native.to_device(B))
    assert rel_err(C.cpu().numpy(), A @ B) < 1e-14
    L = (rs.randn(2, 30, 30) + 1j * rs.randn(2, 30, 30)) * 0.1
    h = 0.3
    hL = h * L
    P4 = np.eye(30) + hL + hL @ hL / 2 + hL @ hL @ hL / 6 + hL @ hL @ hL @ hL / 24
    assert rel_err(lti.rk4_propagator(_native.to_device(L), h).cpu().numpy(), P4) < 1e-14
    assert rel_err(lti.rk4_increment(_native.to_device(L), h).cpu().numpy(),
                   P4 - np.eye(30)) < 1e-14


def test_zgemm_transposed_operands_and_tile_shapes(gpu):
    from qspectra_amd import _native
    rs = np.random.RandomState(9)
    # the last three shapes have >= 192 32x32 tiles: the FP64 matrix-core
    # path (one, three and five K chunks of 64, the last one partial)
    for (b, m, n, k) in [(1, 7, 255, 128), (3, 255, 7, 128), (2, 147, 147, 147),
                         (1, 300, 200, 40), (5, 1, 3, 2), (4, 255, 255, 154), (1, 520, 400, 33),
                         (3, 300, 260, 257)]:
        A = rs.randn(b, m, k) + 1j * rs.randn(b, m, k)
        B = rs.randn(b, k, n) + 1j * rs.randn(b, k, n)
        ref = A @ B
        for ta in (False, True):
            for tb in (False, True):
                Ad = _native.to_device(A.transpose(0, 2, 1).copy() if ta else A)
                Bd = _native.to_device(B.transpose(0, 2, 1).copy() if tb else B)
                C = _native.zgemm(Ad, Bd, transa=ta, transb=tb)
                assert rel_err(C.cpu().numpy(), ref) < 1e-14, (b, m, n, k, ta, tb)
