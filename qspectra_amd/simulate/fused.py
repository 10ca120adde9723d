"""Fused execution plan for third-order response functions and 2-D spectra.

The reference evaluates every Liouville pathway independently
(response.py:269-337): t1 propagation, then all t1 branches propagated
through t2 as separate right-hand sides, then the dipole maps and the
contraction with the t3 Heisenberg trajectories; it notes the redundancy
itself (:292-295).  The response is a product of linear maps,

    R_p(t1, t2_s, t3) = S1(t1) . V1^T . (P2^T)^N_s . V2_p^T . G3_p(t3)^T

(S1: the t1 trajectory as rows; P2: the one-step RK4 propagator of the t2
block; N_s: RK4 steps up to t2_s; G3_p: the t3 Heisenberg trajectory of the
bra of pathway p), so a plan evaluates it in the cheapest association:

  * one t1 trajectory per distinct (s0, s1) pair (GSB, ESE and ESA of the
    '-++' geometry share it), computed by the register stepper;
  * the t2 stage acts on the M1 rows of V1^T instead of on the n1 t1
    branches: Q_s = V1^T (P2^N_s)^T, one batched GEMM with the RK4
    propagator powers;
  * W_s = Q_s V2_p^T for every pathway, written side by side so that all
    pathways contract with their t3 trajectories in one GEMM;
  * each distinct t3 Heisenberg trajectory once, each on a stream of its own
    (the largest block first), while the t2 and t1 factors run in order on
    the main stream;
  * the Fourier transforms of two_dimensional_spectra (simulate/utils.py:
    154-219) as products with the transform matrices F1, F3 (the FFT
    pipeline applied to unit vectors), folded into the t1 and t3 factors:
        X_s = [(F1 S1) W_s] . [F3 G3]^T
    with W_s = [sum_p W_s^p] side by side over the distinct t3 trajectories,
    so that only the last GEMM waits for the t3 streams.
Exact algebra: the result equals the pathway-by-pathway evaluation up to
rounding (the same RK4 schedule, only the order of the linear maps
differs).  The launch sequence can be captured into a HIP graph.

Every operator derived from a Liouvillian block is rebuilt from that block
by every launch sequence (and so by every graph replay): the one-step RK4
propagators P4(h L) (Horner, qsx_rk4_propagator), the save-interval
propagators P4(h L)^q of q RK4 steps (repeated squaring), the t2 powers and
the chunk-start powers (qsx_matrix_powers), and the complex64 / FP64 split
of MIXED precision (qsx_split_increment).  Only the dipole maps, the
initial state and the Fourier matrices -- none of which depends on the
Liouvillian -- are prepared once.  Overwriting the Liouvillian blocks in
place (``liouvillians()``) and replaying therefore computes the response of
the new generator.
"""
import numpy as np
import torch

from .. import _native
from ..engine import rk4_schedule
from .response import (_members, _op_matrices, _parse_pathways, THIRD_ORDER_PATHWAYS,
                       third_order_grids)
from ..schedules import _mode
from ..timing import capture as _capture


def transform_matrix(t, rw_freq, unit_convert, sign, device):
    """(f, F) with F [N, len(t)] the linear map x -> fourier_transform(t, x)
    (zero padding, ifftshift, FFT, fftshift, sign), obtained by applying the
    device FFT pipeline to the unit vectors."""
    from .utils import fourier_transform_device
    eye = torch.eye(len(t), dtype=torch.complex128, device=device)
    f, F = fourier_transform_device(t, eye, 0, rw_freq, unit_convert, sign=sign)
    return f, F.contiguous()


class PlanNotApplicable(ValueError):
    """The response cannot be evaluated by a fused plan (non-uniform t2 step,
    schedule 'rk4', pathways with different t1 propagations)."""


class _Trajectory(object):
    """A trajectory of one right-hand side per member saved at equally spaced
    points (the t1 or a t3 stage of a plan), in nc chunks of c saves.  build()
    enqueues the operators from the Liouvillian block L: the save-interval
    propagator P = P4(h L)^q (q RK4 steps of size h per save interval), the
    chunk-start power P^c and, for MIXED, the complex64 off-diagonal part of
    P with its FP64 diagonal minus one; run() the chunked stepping
    (qsx_propagate_chunks: work-group j reaches its start (P^c)^j y0, then
    takes c - 1 steps with P)."""
    kind = 'chunks'

    def __init__(self, eom, t, dt, precision, chunk):
        n, h = rk4_schedule(t, None, dt)
        if len(n) < 2 or n[0] != 0 or not np.all(n[1:] == n[1]) or n[1] < 1:
            raise PlanNotApplicable('the fused plan needs trajectories saved at equally '
                                    'spaced points from t = 0')
        hs = h[1:]
        if not np.all(np.abs(hs - hs[0]) <= 1e-12 * abs(hs[0])):
            raise PlanNotApplicable('the fused plan needs a uniform step')
        self.eom = eom
        self.L = eom.L
        self.q, self.h = int(n[1]), float(hs[0])
        self.nsave = len(t)
        self.rk4_steps = int(n.sum())
        if chunk is None:
            chunk = max(4, int(np.ceil(np.sqrt(self.nsave))))
        self.c = int(min(chunk, self.nsave))
        self.nc = -(-self.nsave // self.c)
        self.precision = _mode(precision)
        Bl, M, _ = self.L.shape
        dev = self.L.device
        self.M = M
        self.P = torch.empty((Bl, M, M), dtype=torch.complex128, device=dev)
        self.Pc = torch.empty((Bl, 1, M, M), dtype=torch.complex128, device=dev)
        self.ws = _native.propagator_workspace(M, Bl, dev)
        if self.precision == _native.QSX_MIXED:
            self.off = torch.empty((Bl, M, M), dtype=torch.complex64, device=dev)
            self.diag = torch.empty((Bl, M), dtype=torch.complex128, device=dev)
        else:
            self.off = self.diag = None
        self.chunk_n = torch.tensor([0] + [1] * (self.c - 1), dtype=torch.int32, device=dev)
        self.chunk_h = torch.full((self.c,), self.h * self.q, dtype=torch.float64, device=dev)
        # one chunk (no start powers): a large block in MIXED precision gets
        # its split straight from the FP32 matrix cores, without an FP64 P
        self.split_mfma = (self.nc == 1 and self.off is not None and not _native.small_fits(M)
                           and self.q & (self.q - 1) == 0)
        if self.split_mfma:
            self.ws = _native.split_propagator_workspace(M, Bl, dev)

    def products(self):
        """Propagator products the chunked stepping executes per member:
        nc chunks of c - 1 steps plus the nc - 1 products of the chain of
        chunk starts."""
        return self.nc * (self.c - 1) + (self.nc - 1)

    def build(self):
        if getattr(self, 'rebuild', None) is not None:
            self.rebuild()
        if self.split_mfma:
            _native.rk4_propagator_split_into(self.L, self.h, self.q, self.off, self.diag,
                                              self.ws)
            return
        if _native.small_fits(self.M):
            # one launch: Horner, the q-th and (q c)-th powers, the split
            _native.small_propagator_into(self.L, self.h, self.q,
                                          self.c if self.nc > 1 else 0, self.P,
                                          self.Pc if self.nc > 1 else None, self.off, self.diag)
            return
        _native.rk4_propagator_into(self.L, self.h, self.q, self.P, self.ws)
        if self.nc > 1:
            _native.matrix_powers_into(self.P, [self.c], self.Pc, self.ws)
        if self.off is not None:
            _native.split_increment_into(self.P, self.off, self.diag)

    def run(self, y0, statuses, status):
        B, _, M = y0.shape
        A = self.P if self.off is None else self.off
        if self.nc == 1:
            # one chunk: the register stepper, one product per save interval
            out, _, st = _native.propagate_states(_native.QSX_MODE_PROPAGATOR, A, y0,
                                                  self.chunk_n, self.chunk_h, self.precision,
                                                  self.diag, steps=self.c - 1, status=status)
            statuses.append(st)
            return out
        out, st = _native.propagate_chunks(self.Pc[:, 0] if self.nc > 1 else self.P, A, y0,
                                           self.nc, self.chunk_n, self.chunk_h, self.precision,
                                           self.diag, steps=self.c - 1, status=status)
        statuses.append(st)
        return out.reshape(B, 1, self.nc * self.c, M)[:, :, :self.nsave]


class ThirdOrderPlan(object):
    """Third-order response (fourier=False, run() -> [B, n1, n2, n3]) or 2-D
    spectrum (fourier=True, run() -> [B, N1, n2, N3], frequencies()) of a
    model; run() launches on the current stream, capture() records a graph.
    schedule 'rk4' (no propagator) is served by the pathway-by-pathway
    evaluation and not by a plan.

    ensemble_mean: the mean over the B members (the ensemble average of
    simulate/decorators.py:40-64) -> [1, ...]: contracted in the time domain,
        R_s = (1/B) sum_b S1_b (W_bs G3_b^T) = (1/B) [S1_1 .. S1_B] [Y_1s; ..; Y_Bs],
    one GEMM over all members per population time, then the two Fourier
    products once.
    rebuild: every launch first re-runs the Liouvillian builder (eigensystem,
    Redfield tensor, basis change: qsx_redfield_block) from the members'
    Hamiltonians (model.block_builder), so the whole path from H1 runs in
    the launch sequence.
    With chunk=None and at least 32 members the trajectories are stepped in
    one chunk per member (each member's chain fills a compute unit; no
    chunk-start powers), and MIXED blocks too large for one work-group take
    their split from the FP32 matrix cores (qsx_rk4_propagator_split)."""

    def __init__(self, model, coherence_time_max, population_time_max=None,
                 population_times=None, geometry='-++', polarization='xxxx',
                 include_signal=None, dt=None, precision='fp64', schedule='auto',
                 fourier=False, chunk=None, ensemble_mean=False, rebuild=False):
        if schedule == 'rk4':
            raise PlanNotApplicable("the fused plan needs the RK4 propagator; use "
                                    "plan='pathways' for schedule='rk4'")
        if hasattr(model, 'ado_count') or hasattr(model, 'oop_shape'):
            raise PlanNotApplicable('hierarchy (HEOM) and ZOFE models are evaluated pathway '
                                    'by pathway')
        self.model = model
        self.t1, self.t2, self.t3 = third_order_grids(model, coherence_time_max,
                                                      population_time_max, population_times)
        self.B = B = len(_members(model))
        self.precision = precision
        self.fourier = fourier
        self.ensemble_mean = ensemble_mean
        if chunk is None and B >= 32:
            chunk = max(len(self.t1), len(self.t3))          # one chunk per trajectory
        dt = dt if dt is not None else model.time_step
        rho0 = model.thermal_state('gg')
        trans = geometry + '-'
        paths = _parse_pathways(THIRD_ORDER_PATHWAYS[geometry], include_signal)
        dev = _native.device()
        n1, n2, n3 = len(self.t1), len(self.t2), len(self.t3)

        def dev_t(x):
            return torch.as_tensor(np.array(x)).to(dev, torch.complex128).contiguous()

        def per_batch(x, reps):
            """[B|1, ...] -> [B*reps | 1, ...] (member-major)."""
            if x.shape[0] == 1:
                return x
            return x.repeat_interleave(reps, dim=0).contiguous()

        # t2 schedule: cumulative RK4 steps at each population time
        n, h = rk4_schedule(self.t2, 0.0, dt)
        hs = h[n > 0]
        if hs.size and not np.all(np.abs(hs - hs[0]) <= 1e-12 * abs(hs[0])):
            raise PlanNotApplicable('the fused plan needs a uniform t2 step')
        self.N2 = np.cumsum(n)
        h2 = float(hs[0]) if hs.size else float(dt)

        self.t1_stages = {}      # (s0, s1) -> stage, initial states
        self.t3_stages = {}      # (s3, s4) -> stage, bra, column offset in G
        self.groups = {}         # (s0, s1, s2) -> Q operands, branches
        width = 0
        branches = []
        for path in paths:
            s = path.split('->')
            k1 = (s[0], s[1])
            if k1 not in self.t1_stages:
                V0 = _op_matrices(model, '%s->%s' % (s[0], s[1]), polarization[0], trans[0],
                                  'commutator')
                y1 = np.einsum('bij,j->bi', V0, rho0)[:, None, :]
                y1 = np.broadcast_to(y1, (B,) + y1.shape[1:])
                stage = _Trajectory(model.equation_of_motion(s[1]), self.t1, dt, precision,
                                    chunk)
                self.t1_stages[k1] = (stage, dev_t(y1))
            k12 = tuple(s[:3])
            if k12 not in self.groups:
                V1 = _op_matrices(model, '%s->%s' % (s[1], s[2]), polarization[1], trans[1],
                                  'commutator')                        # [B|1, M2, M1]
                eom2 = model.equation_of_motion(s[2])
                Bl, M2 = eom2.L.shape[0], eom2.L.shape[-1]
                if Bl not in (1, B):
                    raise PlanNotApplicable('Liouvillian batch does not match the members')
                # t2 operators, rebuilt by every launch: P2 = P4(h2 L2), its
                # powers P2^N_s at the population times, [B*n2, M2, M2]
                V1T = per_batch(dev_t(V1.transpose(0, 2, 1)), n2)           # [B*n2|1, M1, M2]
                g = dict(t1=k1, V1T=V1T, L=eom2.L, h=h2, rows=None)
                M1 = V1.shape[2]                                           # V1 [B|1, M2, M1]
                if _native.small_fits(M2, M1, n2):
                    # one launch: the rows of V1^T propagated to every population
                    # time (Q_s = V1^T (P2^N_s)^T, shared squarings)
                    rows = dev_t(V1.transpose(0, 2, 1))
                    g['rows'] = rows.expand(B, -1, -1).contiguous() if rows.shape[0] != B \
                        else rows
                else:
                    g['Pn'] = torch.empty((B, n2, M2, M2), dtype=torch.complex128, device=dev)
                    g['P'] = torch.empty((Bl, M2, M2), dtype=torch.complex128, device=dev)
                    g['ws'] = _native.propagator_workspace(M2, Bl, dev)
                self.groups[k12] = g
            k3 = tuple(s[3:])
            if k3 not in self.t3_stages:
                b3 = _op_matrices(model, '%s->%s' % (s[3], s[4]), polarization[3], trans[3],
                                  'bra')[:, None, :]
                b3 = np.broadcast_to(b3, (B,) + b3.shape[1:])
                stage = _Trajectory(model.equation_of_motion(s[3], heisenberg_picture=True),
                                    self.t3, dt, precision, chunk)
                self.t3_stages[k3] = dict(stage=stage, bra=dev_t(b3), col=None)
            V2 = _op_matrices(model, '%s->%s' % (s[2], s[3]), polarization[2], trans[2],
                              'commutator')                            # [B|1, M3, M2]
            V2T = per_batch(dev_t(V2.transpose(0, 2, 1)), n2)           # [B*n2|1, M2, M3]
            M3 = V2T.shape[-1]
            st3 = self.t3_stages[k3]
            first = st3['col'] is None
            if first:
                st3['col'] = width
                width += M3
            branches.append((k12, k3, V2T, st3['col'], M3, first))
        # pathways contract with their t3 trajectory through one column range
        # per distinct trajectory: per t2 group, the V2 maps of its pathways
        # are placed at their trajectory's columns (summed where pathways
        # share one), so that W = sum_g Q_g V2_g^T is one GEMM per group
        self.width = width
        self.branches = branches
        for k12, g in self.groups.items():
            mine = [br for br in branches if br[0] == k12]
            bsz = max(br[2].shape[0] for br in mine)
            M2 = mine[0][2].shape[1]
            full = torch.zeros((bsz, M2, width), dtype=torch.complex128, device=dev)
            for _, _, V2T, col, M3, _ in mine:
                full[:, :, col:col + M3] += V2T
            g['V2T'] = full.contiguous()
        M1s = set(int(g['V1T'].shape[1]) for g in self.groups.values())
        if len(M1s) != 1:
            raise PlanNotApplicable('pathways of one plan must share the t1 block')
        self.M1 = M1s.pop()
        if len(self.t1_stages) != 1:
            raise PlanNotApplicable('pathways of one plan must share the t1 propagation')

        # rebuild: each Liouvillian block is rebuilt on the stream that uses it
        for st, _ in self.t1_stages.values():
            st.rebuild = None
        for g in self.groups.values():
            g['rebuild'] = None
        for st in self.t3_stages.values():
            st['stage'].rebuild = None
        if rebuild:
            for k1, (st, _) in self.t1_stages.items():
                self.hamiltonian_inputs, fn = model.block_builder(k1[1], False)
                st.rebuild = (lambda f, L: lambda: f(L))(fn, st.L)
            for k12, g in self.groups.items():
                _, fn = model.block_builder(k12[2], False)
                g['rebuild'] = (lambda f, L: lambda: f(L))(fn, g['L'])
            for k3, st in self.t3_stages.items():
                _, fn = model.block_builder(k3[0], True)
                st['stage'].rebuild = (lambda f, L: lambda: f(L))(fn, st['stage'].L)
        rw, uc = model.rw_freq, model.unit_convert
        if fourier:
            self.f1, self.F1 = transform_matrix(self.t1, rw, uc, -1, dev)
            self.f3, self.F3 = transform_matrix(self.t3, rw, uc, 1, dev)
        else:
            self.f1, self.f3 = self.t1, self.t3
            self.F1 = torch.eye(n1, dtype=torch.complex128, device=dev)
            self.F3 = torch.eye(n3, dtype=torch.complex128, device=dev)
        self.F1, self.F3 = self.F1[None].contiguous(), self.F3[None].contiguous()
        N1, N3 = self.F1.shape[1], self.F3.shape[1]
        # device buffers reused by every run (fixed addresses for graph replay)
        self.W = torch.empty((B * n2, self.M1, width), dtype=torch.complex128, device=dev)
        self.Q = {k: torch.empty((B * n2, self.M1, g['V1T'].shape[2]), dtype=torch.complex128,
                                 device=dev) for k, g in self.groups.items()}
        if not ensemble_mean:
            # per-member spectra: t3 factors, t1 factors, products, result
            self.G = torch.empty((B, N3, width), dtype=torch.complex128, device=dev)
            self.U = torch.empty((B * n2, N1, width), dtype=torch.complex128, device=dev)
            self.S1F = torch.empty((B, N1, self.M1), dtype=torch.complex128, device=dev)
            self.X = torch.empty((B * n2, N1, N3), dtype=torch.complex128, device=dev)
        else:
            # time-domain contraction buffers: t3 trajectories, per-member
            # products Y_s, the stacked t1 trajectories, the mean response
            self.G3 = torch.empty((B, n3, width), dtype=torch.complex128, device=dev)
            self.Y = torch.empty((n2, B * self.M1, n3), dtype=torch.complex128, device=dev)
            self.S1cat = torch.empty((n1, B, self.M1), dtype=torch.complex128, device=dev)
            self.R = torch.empty((n2, n1, n3), dtype=torch.complex128, device=dev)
            self.RF = torch.empty((n2, N1, n3), dtype=torch.complex128, device=dev)
            self.XE = torch.empty((n2, N1, N3), dtype=torch.complex128, device=dev)
        # status slots of the stepper launches (t1, then each t3), cleared per run
        self.status = torch.zeros(1 + len(self.t3_stages), dtype=torch.int32, device=dev)
        self.steps = self._count_steps(paths)
        self._t3_list = list(self.t3_stages.values())
        self._t3_order = sorted(range(len(self._t3_list)),
                                key=lambda i: -self._t3_list[i]['stage'].M)
        self._sides = [torch.cuda.Stream() for _ in self.t3_stages]
        self._t2side = torch.cuda.Stream()
        # the largest t2 group runs on the t2 stream; the others propagate on
        # the main stream beside it and add to W after it
        self._t2_keys = sorted(self.groups, key=lambda k: -int(self.groups[k]['L'].shape[-1]))
        self._graph = None
        self._statuses = []

    def _count_steps(self, paths):
        """RK4 steps of the reference's pathway-by-pathway algorithm at the
        plan's step (one right-hand side per t1 branch in the t2 stage)."""
        n1 = len(self.t1)
        s1 = next(iter(self.t1_stages.values()))[0].rk4_steps
        s2 = int(self.N2[-1])
        s3 = {k: v['stage'].rk4_steps for k, v in self.t3_stages.items()}
        return sum(s1 + n1 * s2 + s3[tuple(p.split('->')[3:])] for p in paths)

    def liouvillians(self):
        """The Liouvillian blocks the plan reads, {(subspace, heisenberg): L}
        (device tensors [B|1, M, M]); writing into them in place changes the
        generator of the next run or replay."""
        out = {}
        for (s0, s1), (stage, _) in self.t1_stages.items():
            out[(s1, False)] = stage.L
        for (s0, s1, s2), g in self.groups.items():
            out[(s2, False)] = g['L']
        for (s3, s4), st in self.t3_stages.items():
            out[(s3, True)] = st['stage'].L
        return out

    def executed_products(self):
        """Per member: propagator products of the trajectories (chunked
        stepping and chunk-start chains) and GEMMs of the operator build and
        the contractions, by kind."""
        traj = [st for st, _ in self.t1_stages.values()] + \
            [st['stage'] for st in self.t3_stages.values()]
        return {'trajectory_products': {'M=%d' % t.M: t.products() for t in traj},
                'trajectories': len(traj)}

    # ------------------------------------------------------------------ run
    def _launch(self):
        B, n2 = self.B, len(self.t2)
        n1, n3 = len(self.t1), len(self.t3)
        main = torch.cuda.current_stream()
        self.status.zero_()
        statuses = []
        if self.ensemble_mean:
            return self._launch_ensemble(main, statuses)
        # each t3 Heisenberg trajectory on its own stream, transformed along t3
        # and written to its column range of G; the largest block (the
        # longest chain of the plan) is launched first
        for i, side in zip(self._t3_order, self._sides):
            st = self._t3_list[i]
            side.wait_stream(main)
            with torch.cuda.stream(side):
                M3 = st['bra'].shape[-1]
                st['stage'].build()
                G3 = st['stage'].run(st['bra'], statuses,
                                     self.status[1 + i:2 + i])[:, 0, :n3]      # [B, n3, M3]
                _native.zgemm(self.F3, G3, self.G[:, :, st['col']:st['col'] + M3])
        # t2 stage on the rows of V1^T and the V2 maps of every pathway (its
        # operator programs are the main stream's longest chain) on a stream of
        # its own, while the t1 trajectory, transformed along t1, runs on the
        # main stream
        t2side = self._t2side
        t2side.wait_stream(main)
        with torch.cuda.stream(t2side):
            self._t2_stage(B)
        stage1, y1 = next(iter(self.t1_stages.values()))
        stage1.build()
        S1 = stage1.run(y1, statuses, self.status[0:1])[:, 0, :n1]             # [B, n1, M1]
        _native.zgemm(self.F1, S1, self.S1F)
        S1F = self.S1F if B == 1 else self.S1F.repeat_interleave(n2, dim=0)
        for k12 in self._t2_keys[1:]:
            self._t2_propagate(B, k12)
        main.wait_stream(t2side)
        for k12 in self._t2_keys[1:]:
            self._t2_accumulate(self.groups[k12], self.Q[k12], 1.0)
        _native.zgemm(S1F, self.W, self.U)                                       # [B n2, N1, w]
        for side in self._sides:
            main.wait_stream(side)
        G = self.G if B == 1 else self.G.repeat_interleave(n2, dim=0)
        _native.zgemm(self.U, G, self.X, transb=True)                            # [B n2, N1, N3]
        self._statuses = statuses
        N1, N3 = self.X.shape[1], self.X.shape[2]
        return self.X.view(B, n2, N1, N3).permute(0, 2, 1, 3)

    def _launch_ensemble(self, main, statuses):
        """The ensemble-mean response: trajectories per member, then one
        contraction over all members per population time, then (fourier)
        the transforms of the mean."""
        B, n2 = self.B, len(self.t2)
        n1, n3, M1 = len(self.t1), len(self.t3), self.M1
        for i, side in zip(self._t3_order, self._sides):
            st = self._t3_list[i]
            side.wait_stream(main)
            with torch.cuda.stream(side):
                M3 = st['bra'].shape[-1]
                st['stage'].build()
                G3 = st['stage'].run(st['bra'], statuses,
                                     self.status[1 + i:2 + i])[:, 0, :n3]      # [B, n3, M3]
                self.G3[:, :, st['col']:st['col'] + M3].copy_(G3)
        t2side = self._t2side
        t2side.wait_stream(main)
        with torch.cuda.stream(t2side):
            self._t2_stage(B)
        stage1, y1 = next(iter(self.t1_stages.values()))
        stage1.build()
        S1 = stage1.run(y1, statuses, self.status[0:1])[:, 0, :n1]             # [B, n1, M1]
        self.S1cat.copy_(S1.permute(1, 0, 2))
        for k12 in self._t2_keys[1:]:
            self._t2_propagate(B, k12)
        main.wait_stream(t2side)
        for k12 in self._t2_keys[1:]:
            self._t2_accumulate(self.groups[k12], self.Q[k12], 1.0)
        for side in self._sides:
            main.wait_stream(side)
        W = self.W.view(B, n2, M1, -1)
        for s in range(n2):
            # Y_s[b] = W_bs G3_b^T  [M1, n3]
            _native.zgemm(W[:, s], self.G3, self.Y[s].view(B, M1, n3), transb=True)
        # R_s = (1/B) S1cat Y_s, one GEMM over all members per population time
        _native.zgemm(self.S1cat.view(1, n1, B * M1), self.Y, self.R, alpha=1.0 / B)
        self._statuses = statuses
        if not self.fourier:
            return self.R.permute(1, 0, 2)[None]
        _native.zgemm(self.F1, self.R, self.RF)                                 # [n2, N1, n3]
        _native.zgemm(self.RF, self.F3, self.XE, transb=True)                   # [n2, N1, N3]
        return self.XE.permute(1, 0, 2)[None]

    def _t2_stage(self, B):
        """W = Q_g V2_g^T for the largest t2 group g, Q_g the rows of V1^T
        propagated to every population time (the other groups add theirs
        after it, from the main stream)."""
        k12 = self._t2_keys[0]
        self._t2_propagate(B, k12)
        self._t2_accumulate(self.groups[k12], self.Q[k12], 0.0)

    def _t2_propagate(self, B, k12):
        """Q[k12]: the rows of the group's V1^T propagated to every
        population time (its block rebuilt first when the plan rebuilds)."""
        g = self.groups[k12]
        if g['rebuild'] is not None:
            g['rebuild']()
        if g['rows'] is not None:
            _native.propagate_powers_into(g['L'], g['h'], self.N2, g['rows'], self.Q[k12])
        else:
            _native.rk4_propagator_into(g['L'], g['h'], 1, g['P'], g['ws'])
            Pn = g['Pn']
            if g['P'].shape[0] == 1 and B > 1:
                _native.matrix_powers_into(g['P'], self.N2, Pn[:1], g['ws'])
                Pn[1:].copy_(Pn[:1].expand(B - 1, -1, -1, -1))
            else:
                _native.matrix_powers_into(g['P'], self.N2, Pn, g['ws'])
            _native.zgemm(g['V1T'], Pn.view(-1, *Pn.shape[2:]), self.Q[k12], transb=True)

    def _t2_accumulate(self, g, Q, beta):
        """W <- Q V2^T + beta W for one t2 group."""
        if g['V2T'].shape[0] == 1:
            # a shared V2 map: one tall GEMM over the stacked rows of every
            # member and population time
            W = self.W
            _native.zgemm(Q.view(1, -1, Q.shape[-1]), g['V2T'], W.view(1, -1, W.shape[-1]),
                          beta=beta)
        else:
            _native.zgemm(Q, g['V2T'], self.W, beta=beta)

    def capture(self):
        """Records the launch sequence into a HIP graph, replayed by run()."""
        self._launch()                      # warm up allocations and caches
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with _capture(g):
            self._out = self._launch()
        self._graph = g
        return self

    def run(self):
        """Device tensor [B, n1, n2, n3] (time domain) or [B, N1, n2, N3]
        (fourier=True); check_status() raises on a non-finite state."""
        if self._graph is not None:
            self._graph.replay()
            return self._out
        return self._launch()

    def frequencies(self):
        """(f1, t2, f3) of a fourier=True plan ((t1, t2, t3) otherwise)."""
        return self.f1, self.t2, self.f3

    def check_status(self):
        from .utils import IntegratorError
        for st in self._statuses:
            if int(st.item()) != 0:
                raise IntegratorError('integration failed: non-finite state in a response stage')


def plan_chunk(model, coherence_time_max, population_time_max=None, population_times=None,
               geometry='-++', polarization='xxxx', include_signal=None, dt=None,
               precision='fp64', schedule='auto', fourier=False):
    """The chunk length of the plan's time-chunked trajectory stages, chosen
    for the plan as a whole: the stages run concurrently on their streams,
    so the fastest schedule of each stage on its own is not the fastest
    plan.  Candidates (8, 12, 16 and sqrt(n1)) are timed as blocks of 20
    back-to-back replays of the captured plan, sampled in turn (9 samples
    each, the median decides); the decision is cached like the stage tuner's (per
    device and plan shape).  None when tuning is off or the schedule is not
    'auto' (each stage then takes its own default)."""
    from ..schedules import autotune, autotune_enabled
    if schedule != 'auto' or not autotune_enabled():
        return None
    if hasattr(model, 'ado_count') or hasattr(model, 'oop_shape'):
        return None                        # served pathway by pathway, not by a plan
    t1, t2, t3 = third_order_grids(model, coherence_time_max, population_time_max,
                                   population_times)
    members = _members(model)
    sizes = []
    paths = _parse_pathways(THIRD_ORDER_PATHWAYS[geometry], include_signal)
    for path in paths:
        s = path.split('->')
        sizes.append(model.equation_of_motion(s[1]).L.shape[-1])
        sizes.append(model.equation_of_motion(s[3], heisenberg_picture=True).L.shape[-1])
    key = 'plan M=%s B=%d n1=%d n2=%d n3=%d %s %s %s' % (
        ','.join(map(str, sizes)), len(members), len(t1), len(t2), len(t3), geometry,
        precision, 'fourier' if fourier else 'time')
    cands = sorted(set([8, 12, 16, max(4, int(np.ceil(np.sqrt(len(t1)))))]))
    cands = [('chunks', c) for c in cands if 2 * c < len(t1)]
    if not cands:
        return None
    built = {}

    def trial(cand):
        p = built.get(cand)
        if p is None:
            p = ThirdOrderPlan(model, coherence_time_max, population_time_max, population_times,
                               geometry, polarization, include_signal, dt, precision, schedule,
                               fourier, chunk=cand[1]).capture()
            built[cand] = p
        return p.run
    try:
        return autotune(key, cands, trial, reps=9, block=20)[1]
    except PlanNotApplicable:
        return None
