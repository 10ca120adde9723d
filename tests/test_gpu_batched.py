"""Batched absorption spectra of aggregates of uneven size against the oracle,
member by member (response.py:103-154 on each Hamiltonian)."""
import numpy as np
import pytest

import qspectra_amd as qa
from qspectra_amd import _native
from oracle import system as S
from oracle import liouville as Lv
from oracle import propagate as P
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

BATH = (S.CM_K * 77, 35.0, 106.0)


class _FixedStep(Lv.LiouvilleModel):
    time_step = 1.0


def _aggregates(sizes, seed=2):
    rng = np.random.RandomState(seed)
    out = []
    for n in sizes:
        J = rng.normal(0.0, 80.0, (n, n))
        H = (J + J.T) / 2
        H[np.diag_indices(n)] = 12500.0 + rng.normal(0.0, 100.0, n)
        out.append((H, rng.normal(size=(n, 3))))
    return out


def test_absorption_batch_uneven_sizes(gpu):
    from qspectra_amd.batched import absorption_batch
    systems = _aggregates([2, 3, 2, 5, 3, 4, 2])
    hams = [qa.ElectronicHamiltonian(H, bath=qa.DebyeBath(*BATH), dipoles=d)
            for H, d in systems]
    f, X = absorption_batch(hams, 512, rw_freq=12500.0, dt=1.0)
    fm, Xm = absorption_batch(hams, 512, rw_freq=12500.0, dt=1.0, precision='mixed')
    assert X.shape == (len(hams), len(f)) and np.allclose(f, fm)
    for i, (H, d) in enumerate(systems):
        om = _FixedStep(S.ElectronicHamiltonian(H, bath=S.DebyeBath(*BATH), dipoles=d),
                        rw_freq=12500.0, hilbert_subspace='ge', unit_convert=S.CM_FS)
        fo, Xo = P.absorption_spectra(om, 512, dt=1.0)
        assert np.allclose(f, fo)
        assert rel_err(X[i], Xo) < 1e-10, i
        assert rel_err(Xm[i], Xo) < 1e-6, i


def test_absorption_batch_disorder_members(gpu):
    """Static-disorder members evaluated in one batch equal the ensemble
    average of the drop-in API member by member."""
    from qspectra_amd.batched import absorption_batch
    H, d = _aggregates([4], seed=5)[0]
    ham = qa.ElectronicHamiltonian(H, bath=qa.DebyeBath(*BATH), dipoles=d, disorder=100,
                                   random_seed=7)
    members = list(ham.sample_ensemble(6))
    f, X = absorption_batch(members, 256, rw_freq=12500.0, dt=1.0)
    for i, h in enumerate(members[:3]):
        om = _FixedStep(S.ElectronicHamiltonian(h.H_1exc, bath=S.DebyeBath(*BATH),
                                                dipoles=h.dipoles),
                        rw_freq=12500.0, hilbert_subspace='ge',
                        unit_convert=S.CM_FS)
        _, Xo = P.absorption_spectra(om, 256, dt=1.0)
        assert rel_err(X[i], Xo) < 1e-10, i


def _plan_hams(sizes, seed):
    systems = _aggregates(sizes, seed=seed)
    return systems, [qa.ElectronicHamiltonian(H, bath=qa.DebyeBath(*BATH), dipoles=d)
                     for H, d in systems]


def test_ensemble_plan_matches_batch_and_oracle(gpu):
    """One captured graph (Redfield build, propagators, one bucketed launch,
    Fourier transform) over systems of 1..31 sites in mixed order: each
    replay, between eager work, equals absorption_batch (1e-12) and the
    oracle's member-by-member spectra (1e-10)."""
    import torch
    from qspectra_amd.batched import absorption_batch, AbsorptionEnsemblePlan
    sizes = [7, 2, 24, 7, 13, 31, 1, 7, 5, 24, 16, 15, 8, 3, 7, 2, 9]
    systems, hams = _plan_hams(sizes, 4)
    plan = AbsorptionEnsemblePlan(hams, 384, rw_freq=12500.0, dt=1.0)
    f, Xb = absorption_batch(hams, 384, rw_freq=12500.0, dt=1.0)
    assert np.allclose(plan.f, f)
    for rep in range(3):
        X = plan.run()
        torch.cuda.synchronize()
        plan.check_status()
        got = X.cpu().numpy()
        assert got.shape == Xb.shape
        for i in range(len(hams)):
            assert rel_err(got[i], Xb[i]) < 1e-12, (rep, i, sizes[i])
        torch.ones(1000, device='cuda').sum().item()          # eager work between replays
    for i in (0, 2, 5, 6):
        H, d = systems[i]
        om = _FixedStep(S.ElectronicHamiltonian(H, bath=S.DebyeBath(*BATH), dipoles=d),
                        rw_freq=12500.0, hilbert_subspace='ge', unit_convert=S.CM_FS)
        fo, Xo = P.absorption_spectra(om, 384, dt=1.0)
        assert np.allclose(f, fo)
        assert rel_err(got[i], Xo) < 1e-10, i


def test_ensemble_plan_per_bucket_builds(gpu):
    """Models the one-launch operator builder does not take (eigen-basis
    Redfield blocks here): one build per site count on the side streams and
    one bucketed launch at eight steps per pass; every replay equals
    absorption_batch (1e-12)."""
    import torch
    from qspectra_amd.batched import absorption_batch, AbsorptionEnsemblePlan
    sizes = [7, 2, 24, 7, 13, 31, 5, 9]
    _, hams = _plan_hams(sizes, 6)
    kw = dict(evolve_basis='eigen', secular=False)
    plan = AbsorptionEnsemblePlan(hams, 256, rw_freq=12500.0, dt=1.0, model_kwargs=kw)
    assert plan._fused is None
    f, Xb = absorption_batch(hams, 256, rw_freq=12500.0, dt=1.0, model_kwargs=kw)
    assert np.allclose(plan.f, f)
    for rep in range(2):
        got = plan.run().cpu().numpy()
        plan.check_status()
        for i in range(len(hams)):
            assert rel_err(got[i], Xb[i]) < 1e-12, (rep, i, sizes[i])
        torch.ones(1000, device='cuda').sum().item()


def test_ensemble_plan_rebuilds_from_new_hamiltonians(gpu):
    """New Hamiltonians written into a bucket's inputs in place: the replay
    rebuilds every operator and returns their spectra."""
    import torch
    from qspectra_amd.batched import absorption_batch, AbsorptionEnsemblePlan
    _, hams = _plan_hams([4, 6, 4, 4, 6], 8)
    _, new = _plan_hams([4, 6, 4, 4, 6], 9)
    plan = AbsorptionEnsemblePlan(hams, 256, rw_freq=12500.0, dt=1.0, correlation_decay_time=300.0)
    plan.run()
    for n in (4, 6):
        H1, _ = plan.hamiltonian_inputs(n)
        rows = [i for i in range(len(new)) if new[i].n_sites == n]
        H1.copy_(torch.as_tensor(np.stack([new[i].H_1exc for i in rows])))
    X = plan.run().cpu().numpy()
    # dipoles stay those of the plan's systems
    mixed = [qa.ElectronicHamiltonian(new[i].H_1exc, bath=qa.DebyeBath(*BATH),
                                      dipoles=hams[i].dipoles) for i in range(len(new))]
    _, Xb = absorption_batch(mixed, 256, rw_freq=12500.0, dt=1.0, correlation_decay_time=300.0)
    for i in range(len(new)):
        assert rel_err(X[i], Xb[i]) < 1e-12, i


def _pow2(x):
    p = 1
    while p < x:
        p *= 2
    return p


def _bucket_tables(sizes, J, lanes=None):
    """Wave table and offsets of qsx_response_buckets for problems of the
    given sizes (grouped by size, output rows in input order); lanes(M, J)
    gives the lanes per problem (the plan's choice by default)."""
    from qspectra_amd.batched import _group_size
    waves, offs, order = [], [None] * len(sizes), []
    offP = offY = 0
    for M in sorted(set(sizes)):
        idx = [i for i, m in enumerate(sizes) if m == M]
        G = (lanes or _group_size)(M, J)
        first = len(order)
        for w0 in range(0, len(idx), 64 // G):
            waves.append((first + w0, min(64 // G, len(idx) - w0), M, G))
        for i in idx:
            offs[len(order)] = (offP, offY, i)
            order.append(i)
            offP += M * M
            offY += M
    return waves, offs, order


def test_response_buckets_several_steps_per_pass(gpu):
    """qsx_response_buckets at J = 8 and 4 steps per pass (P^J and the
    projection rows (P^j)^T v written by qsx_response_bucket_weights) and
    at one step per pass, for problems of 1..31 states and a step count
    that is not a multiple of J, against the host's s_k = v . P^k y0
    (1e-12), in each of the kernel's forms: a problem per 16-lane row with
    the state in registers (M + J <= 16), two lanes per row, one lane per
    row."""
    import torch
    rs = np.random.RandomState(3)
    sizes = [7, 1, 31, 3, 12, 7, 13, 28, 29, 7, 2]
    nt = 203
    Ps = [np.eye(M) + 0.05 * (rs.randn(M, M) + 1j * rs.randn(M, M)) / np.sqrt(M) for M in sizes]
    y0s = [rs.randn(M) + 1j * rs.randn(M) for M in sizes]
    vs = [rs.randn(M) + 1j * rs.randn(M) for M in sizes]
    ref = np.zeros((len(sizes), nt), complex)
    for i, (P_, y, v) in enumerate(zip(Ps, y0s, vs)):
        for k in range(nt):
            ref[i, k] = v @ y
            y = P_ @ y
    one_lane = lambda M, J: _pow2(M + J)
    two_lanes = lambda M, J: 2 * _pow2(M + J) if 2 * _pow2(M + J) <= 64 else _pow2(M + J)
    for J, lanes in ((1, None), (4, None), (4, one_lane), (8, None), (8, one_lane),
                     (8, two_lanes)):
        waves, offs, order = _bucket_tables(sizes, J, lanes)
        dev = 'cuda'
        P1 = np.concatenate([Ps[i].ravel() for i in order])
        PJ = np.concatenate([np.linalg.matrix_power(Ps[i], J).ravel() for i in order])
        y0 = torch.as_tensor(np.concatenate([y0s[i] for i in order])).to(dev)
        v = torch.as_tensor(np.concatenate([vs[i] for i in order])).to(dev)
        wt = torch.as_tensor(np.asarray(waves, dtype=np.int32)).to(dev)
        ot = torch.as_tensor(np.asarray(offs, dtype=np.int64)).to(dev)
        W = torch.empty(J * y0.numel(), dtype=torch.complex128, device=dev)
        _native.response_bucket_weights_into(wt, ot, torch.as_tensor(P1).to(dev), v, J, W)
        Wh = W.cpu().numpy()
        for b, i in enumerate(order):
            M, oy = sizes[i], offs[b][1]
            for j in range(J):
                wj = np.linalg.matrix_power(Ps[i], j).T @ vs[i]
                assert rel_err(Wh[J * oy + j * M:J * oy + (j + 1) * M], wj) < 1e-13, (J, i, j)
        out = torch.empty((len(sizes), nt), dtype=torch.complex128, device=dev)
        status = torch.ones(len(sizes), dtype=torch.int32, device=dev)
        _native.response_buckets_into(wt, ot, torch.as_tensor(PJ).to(dev), y0, W, out, status, J=J)
        got = out.cpu().numpy()
        assert int(status.max().item()) == 0
        for i in range(len(sizes)):
            assert rel_err(got[i], ref[i]) < 1e-12, (J, i, sizes[i])


# baths of gamma / (2 pi T) = 0.315 and 28.6 (the fused builder's Matsubara
# sum in closed form, through digamma differences) and 0.0038 (the summed
# form, outside the closed form's window)
NARROW_WARM = (S.CM_K * 300, 35.0, 5.0)
BROAD_COLD = (S.CM_K * 4, 35.0, 500.0)


@pytest.mark.parametrize('secular,real,bath', [(False, False, BATH), (True, False, BATH),
                                               (False, True, BATH), (False, False, NARROW_WARM),
                                               (False, False, BROAD_COLD)])
def test_absorption_operators_match_block_builder(gpu, secular, real, bath):
    """qsx_absorption_operators (every member of every size in one launch)
    equals the per-size path: qsx_redfield_block's Heisenberg 'eg' block and
    qsx_small_propagator's P4(h L) and its fourth power (1e-12), for 1..31
    sites, secular or not, complex or real correlation function, with the
    Matsubara sum in closed form or summed; a launch whose max_n is below some
    sizes leaves those problems' outputs alone."""
    import torch
    sizes = [1, 2, 5, 7, 13, 24, 31]
    rw, dt, J = 12500.0, 1.0, 4
    table, H1s, rws, baths, ref1, refJ = [], [], [], [], [], []
    oh = op = 0
    for n in sizes:
        systems = _aggregates([n] * 3, seed=11 + n)
        hams = [qa.ElectronicHamiltonian(H, bath=qa.DebyeBath(*bath), dipoles=d)
                for H, d in systems]
        hams = [h.in_rotating_frame(rw) for h in hams]
        m = qa.RedfieldModel(hams[0], rw_freq=rw, hilbert_subspace='ge', unit_convert=qa.CM_FS,
                             secular=secular, discard_imag_corr=real).with_hamiltonians(hams)
        (H1, rv), rebuild = m.block_builder('eg', heisenberg_picture=True)
        L = torch.empty((3, n, n), dtype=torch.complex128, device='cuda')
        rebuild(L)
        P1 = torch.empty_like(L)
        PJ = torch.empty_like(L)
        _native.small_propagator_into(L, dt, 1, J, P=P1, Pc=PJ)
        ref1.append(P1.cpu().numpy())
        refJ.append(PJ.cpu().numpy())
        H1s.append(H1.cpu().numpy().ravel())
        rws.append(rv.cpu().numpy())
        flags = m._flags(True)
        for k in range(3):
            table.append((n, oh + k * n * n, op + k * n * n))
            baths.append(bath)
        oh += 3 * n * n
        op += 3 * n * n
    dev = 'cuda'
    args = (torch.as_tensor(np.asarray(table, dtype=np.int64)).to(dev),
            torch.as_tensor(np.concatenate(H1s)).to(dev),
            torch.as_tensor(np.concatenate(rws)).to(dev),
            torch.as_tensor(np.asarray(baths, dtype=float)).to(dev))
    # max_n below some sizes: those problems are skipped, the others built
    P1 = torch.zeros(op, dtype=torch.complex128, device=dev)
    PJ = torch.zeros(op, dtype=torch.complex128, device=dev)
    _native.absorption_operators_into(args[0], 7, *args[1:], flags, qa.CM_FS, dt, J, P1, PJ)
    P1, PJ = P1.cpu().numpy(), PJ.cpu().numpy()
    for (n, _, o), k in zip(table, range(len(table))):
        if n > 7:
            assert not P1[o:o + n * n].any() and not PJ[o:o + n * n].any()
        else:
            assert rel_err(P1[o:o + n * n].reshape(n, n), ref1[sizes.index(n)][k % 3]) < 1e-12
    P1 = torch.empty(op, dtype=torch.complex128, device=dev)
    PJ = torch.empty(op, dtype=torch.complex128, device=dev)
    _native.absorption_operators_into(args[0], 31, *args[1:], flags, qa.CM_FS, dt, J, P1, PJ)
    P1, PJ = P1.cpu().numpy(), PJ.cpu().numpy()
    row = 0
    for s, n in enumerate(sizes):
        for k in range(3):
            _, _, o = table[row]
            assert rel_err(P1[o:o + n * n].reshape(n, n), ref1[s][k]) < 1e-12, (n, k)
            assert rel_err(PJ[o:o + n * n].reshape(n, n), refJ[s][k]) < 1e-12, (n, k)
            row += 1
