"""Fused execution plan for third-order response functions.

The reference evaluates every Liouville pathway independently
(response.py:309-336) and notes the redundancy itself (:292-295).  A plan
built here, once per model and grid:
  * merges pathways that share their first propagations (ESE and ESA share
    V0, V1, the t1 and the t2 stage; only V2 and t3 differ),
  * evaluates each distinct t3 Heisenberg propagation once (GSB and ESE
    share the 'eg' one) on a second stream, concurrently with t1 -> t2,
  * keeps every operator matrix, RK4 propagator and step schedule resident on
    the device, so a spectrum is a fixed sequence of kernel launches:
        register stepper (t1)  -> GEMM (V1 map)  -> register stepper (t2, all
        t1 branches as right-hand sides) -> GEMM (V2 map) -> GEMM (contract
        with the t3 bra trajectories)
  * can be captured once into a HIP graph and replayed.
Numerically identical to the pathway-by-pathway evaluation (same kernels and
operands, same summation of pathway contributions up to order).
"""
import numpy as np
import torch

from .. import _native
from ..engine import rk4_schedule
from .response import (_members, _op_matrices, _parse_pathways, THIRD_ORDER_PATHWAYS,
                       third_order_grids)
from .utils import _mode
from ..schedules import Stage as _Stage


class ThirdOrderPlan(object):
    def __init__(self, model, coherence_time_max, population_time_max=None,
                 population_times=None, geometry='-++', polarization='xxxx',
                 include_signal=None, dt=None, precision='fp64', schedule='auto'):
        self.model = model
        self.t1, self.t2, self.t3 = third_order_grids(model, coherence_time_max,
                                                      population_time_max, population_times)
        self.B = B = len(_members(model))
        self.precision = precision
        # dipole maps and contractions run in FP64 in every precision: the
        # contraction sums oscillating terms that cancel, and these GEMMs are
        # a small share of the work
        self.gemm_dt = torch.complex128
        dt = dt if dt is not None else model.time_step
        rho0 = model.thermal_state('gg')
        trans = geometry + '-'
        paths = _parse_pathways(THIRD_ORDER_PATHWAYS[geometry], include_signal)
        dev = _native.device()

        def dev_t(x, dtype=torch.complex128):
            return torch.as_tensor(np.array(x)).to(dev, dtype).contiguous()

        self.groups = {}      # (s0, s1, s2) -> dict(t1 stage, V1T, t2 stage, branches)
        self.t3_stages = {}   # (s3, s4) -> (stage, bra)
        for path in paths:
            s = path.split('->')
            key12 = tuple(s[:3])
            if key12 not in self.groups:
                V0 = _op_matrices(model, '%s->%s' % (s[0], s[1]), polarization[0], trans[0],
                                  'commutator')
                V1 = _op_matrices(model, '%s->%s' % (s[1], s[2]), polarization[1], trans[1],
                                  'commutator')
                y1 = np.einsum('bij,j->bi', V0, rho0)[:, None, :]
                y1 = np.broadcast_to(y1, (B,) + y1.shape[1:])
                self.groups[key12] = dict(
                    y1=dev_t(y1),
                    stage1=_Stage(model.equation_of_motion(s[1]), self.t1, None, dt, precision,
                                  schedule, 1),
                    V1T=dev_t(V1.transpose(0, 2, 1), self.gemm_dt),
                    stage2=_Stage(model.equation_of_motion(s[2]), self.t2, 0.0, dt, precision,
                                  schedule, len(self.t1)),
                    branches=[])
            key3 = tuple(s[3:])
            if key3 not in self.t3_stages:
                b3 = _op_matrices(model, '%s->%s' % (s[3], s[4]), polarization[3], trans[3],
                                  'bra')[:, None, :]
                b3 = np.broadcast_to(b3, (B,) + b3.shape[1:])
                self.t3_stages[key3] = (
                    _Stage(model.equation_of_motion(s[3], heisenberg_picture=True), self.t3,
                           None, dt, precision, schedule, 1), dev_t(b3))
            V2 = _op_matrices(model, '%s->%s' % (s[2], s[3]), polarization[2], trans[2],
                              'commutator')
            self.groups[key12]['branches'].append((dev_t(V2.transpose(0, 2, 1), self.gemm_dt),
                                                   key3))
        self.steps = sum(g['stage1'].steps + len(self.t1) * g['stage2'].steps * len(g['branches'])
                         + sum(self.t3_stages[k][0].steps for _, k in g['branches'])
                         for g in self.groups.values())
        self._side = torch.cuda.Stream()
        self._graph = None
        self._statuses = []

    # ------------------------------------------------------------------ run
    def _gemm(self, A, B, C=None, beta=0.0):
        return _native.zgemm(A.contiguous(), B, C, beta=beta)

    def _launch(self):
        B, n1, n2, n3 = self.B, len(self.t1), len(self.t2), len(self.t3)
        statuses = []
        main = torch.cuda.current_stream()
        self._side.wait_stream(main)
        G3 = {}
        with torch.cuda.stream(self._side):
            for key, (stage, bra) in self.t3_stages.items():
                G = stage.run(bra, statuses)                                  # [B, 1, n3, M3]
                G3[key] = G[:, 0].transpose(1, 2).to(self.gemm_dt).contiguous()   # [B, M3, n3]
                G3[key].record_stream(main)      # consumed on the main stream
        total = torch.zeros((B, n1 * n2, n3), dtype=torch.complex128, device=main.device)
        first = True
        for g in self.groups.values():
            s1 = g['stage1'].run(g['y1'], statuses)                           # [B, 1, n1, M1]
            Y1 = self._gemm(s1.reshape(B, n1, -1), g['V1T'])                  # [B, n1, M2]
            s2 = g['stage2'].run(Y1, statuses)                                # [B, n1, n2, M2]
            for V2T, key3 in g['branches']:
                Y2 = self._gemm(s2.reshape(B, n1 * n2, -1), V2T)              # [B, n1 n2, M3]
                if first:
                    main.wait_stream(self._side)
                    first = False
                self._gemm(Y2, G3[key3], total, beta=1.0)
        main.wait_stream(self._side)
        self._statuses = statuses
        return total.reshape(B, n1, n2, n3)

    def _launch_2d(self):
        """The response followed by the two Fourier transforms of
        two_dimensional_spectra (t1 with sign -1, then t3), on the device."""
        from .utils import fourier_transform_device
        m = self.model
        X = self._launch()
        _, X = fourier_transform_device(self.t1, X, 1, m.rw_freq, m.unit_convert, sign=-1)
        _, X = fourier_transform_device(self.t3, X, 3, m.rw_freq, m.unit_convert)
        return X

    def capture(self, two_dimensional=False):
        """Records the launch sequence (optionally with the 2-D Fourier
        transforms) into a HIP graph, replayed by run()."""
        fn = self._launch_2d if two_dimensional else self._launch
        fn()                                # warm up allocations, caches, FFT plans
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._out = fn()
        self._graph = g
        self._captured_2d = two_dimensional
        return self

    def run(self):
        """Device tensor [B, n1, n2, n3] (or the 2-D spectrum [B, N1, n2, N3]
        when captured with two_dimensional=True); check_status() raises on
        a non-finite state."""
        if self._graph is not None:
            self._graph.replay()
            return self._out
        return self._launch()

    def frequencies(self):
        """(f1, t2, f3) axes of the captured 2-D spectrum."""
        from .utils import fourier_transform_device
        m = self.model
        z1 = torch.zeros(len(self.t1), dtype=torch.complex128, device=self._side.device)
        f1, _ = fourier_transform_device(self.t1, z1, 0, m.rw_freq, m.unit_convert, sign=-1)
        z3 = torch.zeros(len(self.t3), dtype=torch.complex128, device=self._side.device)
        f3, _ = fourier_transform_device(self.t3, z3, 0, m.rw_freq, m.unit_convert)
        return f1, self.t2, f3

    def check_status(self):
        from .utils import IntegratorError
        for st in self._statuses:
            if int(st.item()) != 0:
                raise IntegratorError('integration failed: non-finite state in a response stage')
