"""Full-size parity of the 2-D workload (7-site aggregate, non-secular
Redfield g/e/f, t1 = t3 = 128 points at 2 fs, t2 in {0, 100, 200, 500} fs,
GSB + ESE + ESA) and of the device-built propagation operators.

Tolerances: the device operators against numpy 1e-12 (P4, splits) and
1e-11 (powers up to 500, relative max-norm); the fused plan in FP64
against the oracle's RK4 on the same 1 fs step 1e-10; against the
reference's tight-tolerance zvode result on the full grid (fixture
tests/golden/spectra_2d_full.npz) 1e-6 in FP64 and MIXED, for the
response and for the 2-D spectrum (RK4 at 1 fs is 2.7e-7 from it).
"""
import numpy as np
import pytest
import torch

import qspectra_amd as qa
from qspectra_amd import _native, workloads as W
from oracle import system as S, liouville as Lv, propagate as P
from tests.conftest import golden, rel_err

pytestmark = pytest.mark.gpu

WORK = W.AGG7_2D


def _models(seed=1):
    H1, d, extra = W.aggregate7_hamiltonian(seed)
    m = qa.RedfieldModel(qa.ElectronicHamiltonian(H1, bath=qa.DebyeBath(S.CM_K * 77, 35, 106),
                                                  dipoles=d, energy_spread_extra=extra),
                         unit_convert=qa.CM_FS, secular=False)
    om = Lv.LiouvilleModel(S.ElectronicHamiltonian(H1, bath=S.DebyeBath(S.CM_K * 77, 35, 106),
                                                   dipoles=d, energy_spread_extra=extra),
                           unit_convert=S.CM_FS, secular=False)
    return m, om


def _oracle_response(om, dt=1.0):
    return P.third_order_response(om, WORK['coherence_time_max'],
                                  population_times=WORK['population_times'],
                                  geometry=WORK['geometry'], include_signal=WORK['include_signal'],
                                  dt=dt)[1]


def _plan(m, precision='fp64', fourier=False, chunk=None):
    from qspectra_amd.simulate.fused import ThirdOrderPlan
    return ThirdOrderPlan(m, WORK['coherence_time_max'],
                          population_times=WORK['population_times'], geometry=WORK['geometry'],
                          include_signal=WORK['include_signal'], dt=WORK['dt'],
                          precision=precision, fourier=fourier, chunk=chunk)


def _taylor4(X):
    M = X.shape[-1]
    I = np.eye(M)
    return I + X + X @ X / 2 + X @ X @ X / 6 + X @ X @ X @ X / 24


def test_rk4_propagator_powers_and_split(gpu):
    rs = np.random.RandomState(11)
    for M, B in [(1, 1), (7, 1), (7, 3), (49, 2), (147, 1)]:
        # random part of spectral radius ~0.07 beside a damped rotating
        # diagonal: P4(L)^500 neither overflows nor underflows
        L = ((rs.randn(B, M, M) + 1j * rs.randn(B, M, M)) * (0.05 / np.sqrt(M))
             - (0.08 + 0.3j) * np.eye(M))
        Ld = torch.as_tensor(L).cuda()
        ws = _native.propagator_workspace(M, B)
        for h, q in [(0.5, 1), (1.0, 2), (0.7, 3), (0.25, 4)]:
            Pd = torch.empty((B, M, M), dtype=torch.complex128, device='cuda')
            _native.rk4_propagator_into(Ld, h, q, Pd, ws)
            ref = np.stack([np.linalg.matrix_power(_taylor4(h * L[b]), q) for b in range(B)])
            assert rel_err(Pd.cpu().numpy(), ref) < 1e-12, (M, B, h, q)
        exps = [0, 1, 2, 5, 12, 16, 100, 500]
        out = torch.empty((B, len(exps), M, M), dtype=torch.complex128, device='cuda')
        _native.matrix_powers_into(Pd, exps, out, ws)
        Pn = Pd.cpu().numpy()
        for s, e in enumerate(exps):
            ref = np.stack([np.linalg.matrix_power(Pn[b], e) for b in range(B)])
            assert rel_err(out[:, s].cpu().numpy(), ref) < 1e-11, (M, e)
        off = torch.empty((B, M, M), dtype=torch.complex64, device='cuda')
        dg = torch.empty((B, M), dtype=torch.complex128, device='cuda')
        _native.split_increment_into(Pd, off, dg)
        o, g = off.cpu().numpy(), dg.cpu().numpy()
        for b in range(B):
            assert np.all(np.diag(o[b]) == 0)
            od = Pn[b] - np.diag(np.diag(Pn[b]))
            assert np.array_equal(o[b], od.astype(np.complex64))
            assert np.array_equal(g[b], np.diag(Pn[b]) - 1)
    # zero RK4 steps per application is rejected before any launch
    with pytest.raises(_native.NativeError):
        _native.rk4_propagator_into(Ld, 1.0, 0, Pd, ws)


def test_small_block_programs(gpu):
    """The one-launch programs of blocks of M <= 64 (Horner, powers, split;
    propagation of several states to a few save points) equal numpy."""
    rs = np.random.RandomState(12)
    for M, B in [(1, 1), (7, 2), (16, 1), (49, 2), (64, 1)]:
        L = ((rs.randn(B, M, M) + 1j * rs.randn(B, M, M)) * (0.05 / np.sqrt(M))
             - (0.08 + 0.3j) * np.eye(M))
        Ld = torch.as_tensor(L).cuda()
        assert _native.small_fits(M)
        for h, q, c in [(0.5, 1, 0), (1.0, 2, 16), (0.7, 3, 12), (0.25, 4, 5)]:
            Pd = torch.empty((B, M, M), dtype=torch.complex128, device='cuda')
            Pc = torch.empty_like(Pd)
            off = torch.empty((B, M, M), dtype=torch.complex64, device='cuda')
            dg = torch.empty((B, M), dtype=torch.complex128, device='cuda')
            _native.small_propagator_into(Ld, h, q, c, Pd, Pc if c else None, off, dg)
            ref = np.stack([np.linalg.matrix_power(_taylor4(h * L[b]), q) for b in range(B)])
            assert rel_err(Pd.cpu().numpy(), ref) < 1e-12, (M, h, q)
            if c:
                refc = np.stack([np.linalg.matrix_power(ref[b], c) for b in range(B)])
                assert rel_err(Pc.cpu().numpy(), refc) < 1e-11, (M, c)
            Pn = Pd.cpu().numpy()
            for b in range(B):
                od = Pn[b] - np.diag(np.diag(Pn[b]))
                assert np.array_equal(off[b].cpu().numpy(), od.astype(np.complex64))
                assert np.array_equal(dg[b].cpu().numpy(), np.diag(Pn[b]) - 1)
        for R, exps in [(7, [0, 100, 200, 500]), (3, [5]), (1, [0, 1, 2, 3, 4, 7, 64, 129])]:
            if not _native.small_fits(M, R, len(exps)):
                continue
            y0 = rs.randn(B, R, M) + 1j * rs.randn(B, R, M)
            out = torch.empty((B, len(exps), R, M), dtype=torch.complex128, device='cuda')
            _native.propagate_powers_into(Ld, 1.0, exps, torch.as_tensor(y0).cuda(), out)
            P1 = [_taylor4(L[b]) for b in range(B)]
            for s, e in enumerate(exps):
                ref = np.stack([y0[b] @ np.linalg.matrix_power(P1[b], e).T for b in range(B)])
                assert rel_err(out[:, s].cpu().numpy(), ref) < 1e-11, (M, R, e)
    assert not _native.small_fits(65) and not _native.small_fits(147)


def test_split_propagator_on_fp32_matrix_cores(gpu):
    """qsx_rk4_propagator_split (Horner on the differences from the exact
    diagonal recursion, FP32 MFMA GEMMs) against the FP64 propagator: the
    split reassembles P4(hL)^q to FP32 accuracy relative to |P| (1e-7), and
    a trajectory stepped with it stays within 1e-6 of the FP64 one."""
    m, _ = _models()
    for sub in ('fe', 'ee', 'eg'):
        L = np.ascontiguousarray(m.equation_of_motion(sub, heisenberg_picture=True)
                                 .L[0].cpu().numpy())
        for B, q in ((1, 1), (3, 2), (2, 4)):
            Lb = np.stack([L] * B) * np.linspace(1.0, 1.1, B)[:, None, None]
            M = L.shape[-1]
            Ld = torch.as_tensor(Lb).cuda()
            off = torch.empty((B, M, M), dtype=torch.complex64, device='cuda')
            dg = torch.empty((B, M), dtype=torch.complex128, device='cuda')
            _native.rk4_propagator_split_into(Ld, 1.0, q, off, dg,
                                              _native.split_propagator_workspace(M, B))
            o, g = off.cpu().numpy().astype(complex), dg.cpu().numpy()
            for b in range(B):
                P = np.linalg.matrix_power(_taylor4(Lb[b]), q)
                Pm = o[b] + np.diag(g[b] + 1)
                assert np.all(np.diag(o[b]) == 0)
                assert rel_err(Pm, P) < 1e-7, (sub, B, q, rel_err(Pm, P))
            # 127 save intervals from a uniform state, FP32 off-diagonal products
            y = yr = np.ones(M, complex) / np.sqrt(M)
            P = np.linalg.matrix_power(_taylor4(Lb[0]), q)
            o32 = off[0].cpu().numpy()
            err = 0.0
            for _ in range(127):
                yr = P @ yr
                y = y + (o32 @ y.astype(np.complex64)).astype(complex) + g[0] * y
                err = max(err, np.abs(y - yr).max() / np.abs(yr).max())
            assert err < 1e-6, (sub, err)
    with pytest.raises(_native.NativeError):        # q must be a power of two
        _native.rk4_propagator_split_into(Ld, 1.0, 3, off, dg,
                                          _native.split_propagator_workspace(M, B))


def test_full_grid_response_matches_oracle_and_reference(gpu):
    """The whole 128 x 4 x 128 response: FP64 plan vs the oracle's RK4 at the
    same 1 fs step, FP64 and MIXED vs the reference's tight zvode."""
    g = golden('spectra_2d_full')
    m, om = _models()
    ref_rk4 = _oracle_response(om)
    plan = _plan(m)
    X = plan.run()[0].cpu().numpy()
    plan.check_status()
    assert np.allclose(plan.t1, g['t1']) and np.allclose(plan.t2, g['t2'])
    assert X.shape == g['R3'].shape == (128, 4, 128)
    assert rel_err(X, ref_rk4) < 1e-10
    e64 = rel_err(X, g['R3'])
    assert e64 < 1e-6, e64
    Xm = _plan(m, 'mixed').run()[0].cpu().numpy()
    assert rel_err(Xm, g['R3']) < 1e-6
    assert rel_err(Xm, X) < 1e-7
    # the drop-in API (fused plan by default) on the same grid
    _, Xa = qa.third_order_response(m, WORK['coherence_time_max'],
                                    population_times=WORK['population_times'],
                                    geometry=WORK['geometry'],
                                    include_signal=WORK['include_signal'], dt=WORK['dt'])
    assert rel_err(Xa, ref_rk4) < 1e-10


def test_full_grid_2d_spectrum_matches_reference(gpu):
    g = golden('spectra_2d_full')
    m, _ = _models()
    for prec in ('fp64', 'mixed'):
        plan = _plan(m, prec, fourier=True).capture()
        f1, t2, f3 = plan.frequencies()
        assert np.allclose(f1, g['f1']) and np.allclose(f3, g['f3'])
        for _ in range(2):
            X = plan.run()[0].cpu().numpy()
            plan.check_status()
            assert X.shape == (255, 4, 255)
            assert rel_err(X, g['S2d']) < 1e-6, prec


def _swap_in(plan, model):
    """Writes the blocks of ``model`` into the plan's Liouvillians in place."""
    for (sub, heis), L in plan.liouvillians().items():
        new = model.equation_of_motion(sub, heisenberg_picture=heis).L
        assert new.shape == L.shape
        L.copy_(new)


def test_graph_replay_rebuilds_operators_from_liouvillians(gpu):
    """A captured plan rebuilds every propagator from the Liouvillian blocks
    on each replay: after the blocks are overwritten in place with those of
    another aggregate, the replay gives that aggregate's response (oracle,
    1e-10 FP64; 1e-6 MIXED), and writing the first blocks back restores the
    first response."""
    mA, omA = _models(1)
    mB, omB = _models(5)
    refA, refB = _oracle_response(omA), _oracle_response(omB)
    assert rel_err(refB, refA) > 1e-2                    # the two responses differ
    for prec, tol in (('fp64', 1e-10), ('mixed', 1e-6)):
        plan = _plan(mA, prec).capture()
        XA = plan.run()[0].cpu().numpy()
        assert rel_err(XA, refA) < tol
        saved = {k: L.clone() for k, L in plan.liouvillians().items()}
        _swap_in(plan, mB)
        XB = plan.run()[0].cpu().numpy()
        plan.check_status()
        assert rel_err(XB, refB) < tol, prec
        for k, L in plan.liouvillians().items():
            L.copy_(saved[k])
        XA2 = plan.run()[0].cpu().numpy()
        assert rel_err(XA2, XA) < 1e-13
