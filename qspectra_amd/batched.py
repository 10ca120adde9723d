"""Batched spectra of many systems at once (static-disorder ensembles and
aggregates of uneven size).

The reference evaluates an ensemble member by member, rebuilding each
member's Liouvillian (decorators.py:51-63).  ``absorption_batch`` groups
systems by site count; each group is one batched GPU Liouvillian build
(eigensystems, C(w), Redfield blocks) and one propagation launch with every
member a batch entry.  All systems share the rotating frame ``rw_freq`` and
the step ``dt`` so their spectra share one frequency axis.
"""
import numpy as np
import torch

from . import _native
from .constants import CM_FS
from .dynamics.redfield import RedfieldModel
from .simulate.response import _batched_linear_response
from .simulate.utils import fourier_transform_device


def absorption_batch(hamiltonians, time_max, rw_freq, dt, model_cls=RedfieldModel,
                     model_kwargs=None, polarization='xx', precision='fp64', schedule='auto',
                     correlation_decay_time=None, unit_convert=CM_FS):
    """Absorption spectra of ``hamiltonians`` (any mix of site counts).

    Returns (f, X) with X [len(hamiltonians), len(f)], row i the spectrum of
    hamiltonians[i] (Re FT of -S(t), as absorption_spectra)."""
    kw = dict(hilbert_subspace='ge', unit_convert=unit_convert)
    kw.update(model_kwargs or {})
    groups = {}
    for i, h in enumerate(hamiltonians):
        groups.setdefault(h.n_sites, []).append(i)
    t = np.arange(0, time_max, dt)
    out = None
    f = None
    for n, idx in sorted(groups.items()):
        hams = [hamiltonians[i].in_rotating_frame(rw_freq) for i in idx]
        model = model_cls(hams[0], rw_freq=rw_freq, **kw).with_hamiltonians(hams)
        model.time_step_override = dt
        ticks, signal = _batched_linear_response(model, 'gg->eg->gg', time_max, None,
                                                 polarization, dt=dt, precision=precision,
                                                 schedule=schedule)
        x = torch.as_tensor(signal).to(_native.device())
        if correlation_decay_time is not None:
            x = x * torch.as_tensor(np.exp(-ticks / correlation_decay_time)).to(x.device)
        f, X = fourier_transform_device(ticks, -x, -1, rw_freq, unit_convert)
        X = X.real.cpu().numpy()
        if out is None:
            out = np.empty((len(hamiltonians), X.shape[-1]))
        out[idx] = X
    return f, out


# steps per round trip of qsx_response_buckets in the ensemble plan
STEPS_PER_PASS = 8


def _group_size(M, J=STEPS_PER_PASS):
    """Lanes per problem in qsx_response_buckets (M state rows and J
    projection rows): 16 when M + J <= 16 (the register-resident form, a
    problem per 16-lane row); otherwise p, the power of two at or above
    M + J, with two lanes per row (2p) when 2p fits the wavefront."""
    if J > 1 and M + J <= 16:
        return 16
    p = 2
    while p < M + J:
        p *= 2
    return 2 * p if J > 1 and 2 * p <= 64 else p


class AbsorptionEnsemblePlan(object):
    """Absorption spectra of many Hamiltonians of unequal site counts as one
    captured HIP graph, spectra left on the device.

    The Redfield 'eg' blocks (Heisenberg picture) rebuilt from the members'
    Hamiltonians and their one-step propagators P = P4(dt L) with P^8, the
    projection rows (P^j)^T v, j < 8 (qsx_response_bucket_weights), and a
    launch that propagates the members eight steps per pass and records
    their linear-response signals (qsx_response_buckets).  For a Debye bath
    the operators come from qsx_absorption_operators and this chain runs
    once per size class (sites <= 8, 16, 24, 31), the classes concurrent on
    side streams; otherwise each site count (bucket) is built with
    qsx_redfield_block and qsx_small_propagator and one launch propagates
    every bucket.  The Fourier transform of all signals gives the spectra.  The values are those of absorption_batch
    (absorption_spectra member by member, response.py:103-154) with FP64
    propagation at the step dt.

        plan = AbsorptionEnsemblePlan(hams, 4096.0, rw_freq=12450.0, dt=1.0)
        X = plan.run()          # [len(hams), len(plan.f)] device tensor
        plan.check_status()

    Writing new Hamiltonians into plan.hamiltonian_inputs(n)[0] in place
    and replaying gives their spectra (the graph rebuilds every operator)."""

    def __init__(self, hamiltonians, time_max, rw_freq, dt, model_cls=RedfieldModel,
                 model_kwargs=None, polarization='xx', correlation_decay_time=None,
                 unit_convert=CM_FS, graph=True):
        from .simulate.response import _op_matrices
        lib = _native.require_native()
        max_m = int(lib.qsx_response_buckets_max_states())
        dev = _native.device()
        kw = dict(hilbert_subspace='ge', unit_convert=unit_convert)
        kw.update(model_kwargs or {})
        groups = {}
        for i, h in enumerate(hamiltonians):
            groups.setdefault(h.n_sites, []).append(i)
        self.t = np.arange(0, time_max, dt)
        nt = len(self.t)
        self.dt, self.rw_freq, self.unit_convert = float(dt), float(rw_freq), unit_convert
        self.decay = correlation_decay_time
        self.buckets = []
        order, waves, offs, y0s, vs = [], [], [], [], []
        offP = offY = 0
        for n, idx in sorted(groups.items()):
            hams = [hamiltonians[i].in_rotating_frame(rw_freq) for i in idx]
            model = model_cls(hams[0], rw_freq=rw_freq, **kw).with_hamiltonians(hams)
            model.time_step_override = dt
            inputs, rebuild = model.block_builder('eg', heisenberg_picture=True)
            B = len(idx)
            M = n
            if M > max_m:
                raise ValueError('AbsorptionEnsemblePlan handles up to %d sites per system '
                                 '(use absorption_batch for larger ones)' % max_m)
            rho0 = model.thermal_state('gg')
            V0 = _op_matrices(model, 'gg->eg', polarization[0], '+', 'commutator')
            bra = _op_matrices(model, 'eg->gg', polarization[1], '-', 'bra')
            v = np.einsum('bij,j->bi', V0, rho0)
            y0s.append(np.broadcast_to(-bra, (B, M)).reshape(-1))
            vs.append(np.broadcast_to(v, (B, M)).reshape(-1))
            G = _group_size(M)
            per = 64 // G
            first = len(order)
            for w0 in range(0, B, per):
                waves.append((first + w0, min(per, B - w0), M, G))
            for k in range(B):
                offs.append((offP + k * M * M, offY + k * M, idx[k]))
            self.buckets.append(dict(n=n, B=B, M=M, inputs=inputs, rebuild=rebuild,
                                     L=torch.empty((B, M, M), dtype=torch.complex128,
                                                   device=dev),
                                     p0=offP, model=model))
            order.extend(idx)
            offP += B * M * M
            offY += B * M
        self.order = np.asarray(order)
        self.B = len(order)
        J = self.J = STEPS_PER_PASS
        self.P = torch.empty(offP, dtype=torch.complex128, device=dev)     # P^J
        self.P1 = torch.empty(offP, dtype=torch.complex128, device=dev)    # P
        for bk in self.buckets:
            sl = slice(bk['p0'], bk['p0'] + bk['B'] * bk['M'] ** 2)
            bk['P'] = self.P[sl].view(bk['B'], bk['M'], bk['M'])
            bk['P1'] = self.P1[sl].view(bk['B'], bk['M'], bk['M'])
        self.y0 = _native.to_device(np.concatenate(y0s), dev=dev)
        self.v = _native.to_device(np.concatenate(vs), dev=dev)
        self.W = torch.empty(J * offY, dtype=torch.complex128, device=dev)
        self.waves = torch.as_tensor(np.asarray(waves, dtype=np.int32)).to(dev)
        self.offsets = torch.as_tensor(np.asarray(offs, dtype=np.int64)).to(dev)
        self.signal = torch.empty((self.B, nt), dtype=torch.complex128, device=dev)
        self.status = torch.zeros(self.B, dtype=torch.int32, device=dev)
        if correlation_decay_time is not None:
            self._decay = torch.as_tensor(np.exp(-self.t / correlation_decay_time)).to(dev)
        self._check_tables(waves, offs, offP, offY)
        self._fused = self._fused_builds(offs, waves, offP, dev)
        # the buckets' builds are independent: largest first, round-robin over
        # side streams, joined before the propagation launch
        self._build_order = sorted(range(len(self.buckets)),
                                   key=lambda k: -self.buckets[k]['B'] * self.buckets[k]['M'] ** 3)
        self._sides = [torch.cuda.Stream() for _ in range(min(4, len(self.buckets)))]
        self.f = None
        self._graph = None
        self._out = None
        if graph:
            self.capture()

    # largest site count of each launch of qsx_absorption_operators
    FUSED_CLASSES = (8, 16, 24, 31)

    def _fused_builds(self, offs, waves, nP, dev):
        """Redfield models of a Debye bath in the site basis: the operators of
        every bucket come from qsx_absorption_operators, one launch per size
        class reading flat H1 / rw buffers, and each bucket's
        hamiltonian_inputs become views into them.  Each class then runs its
        own projection-row and propagation launches over its wavefronts, so
        the classes proceed independently on the side streams.  Returns the
        launch arguments, or None (other models and baths keep one build per
        bucket)."""
        from .bath import debye_parameters, is_debye_form
        from .dynamics.redfield import RedfieldModel
        lib = _native.require_native()
        flags = set()
        for bk in self.buckets:
            m = bk['model']
            if (type(m) is not RedfieldModel or m.evolve_basis != 'site'
                    or not is_debye_form(m.hamiltonian.bath) or bk['inputs'][0].is_complex()
                    or bk['n'] > int(lib.qsx_absorption_operators_max_sites())):
                return None
            flags.add(m._flags(True))
        if len(flags) != 1:
            return None
        nH = sum(bk['B'] * bk['n'] ** 2 for bk in self.buckets)
        H1 = torch.empty(nH, dtype=torch.float64, device=dev)
        rw = torch.empty(self.B, dtype=torch.float64, device=dev)
        table, bath = [], []
        oh = ob = 0
        for bk in self.buckets:
            B, n = bk['B'], bk['n']
            Hv = H1[oh:oh + B * n * n].view(B, n, n)
            rv = rw[ob:ob + B]
            Hv.copy_(bk['inputs'][0])
            rv.copy_(bk['inputs'][1])
            bk['inputs'] = (Hv, rv)
            tb = debye_parameters(bk['model'].hamiltonian.bath)
            for k in range(B):
                table.append((n, oh + k * n * n, offs[ob + k][0]))
                bath.append(tb)
            oh += B * n * n
            ob += B
        for n, h, pp in table:
            assert 0 <= h and h + n * n <= nH and 0 <= pp and pp + n * n <= nP
        # one launch per size class (the table is in ascending n): each sizes
        # its LDS for its largest n
        tab = torch.as_tensor(np.asarray(table, dtype=np.int64)).to(dev)
        bt = torch.as_tensor(np.asarray(bath, dtype=np.float64)).to(dev)
        sizes = np.asarray([r[0] for r in table])
        wsizes = np.asarray([w[2] for w in waves])      # ascending M = n
        launches = []
        lo = wlo = 0
        for top in self.FUSED_CLASSES:
            hi = int(np.searchsorted(sizes, top, side='right'))
            whi = int(np.searchsorted(wsizes, top, side='right'))
            if hi > lo:
                launches.append((top, tab[lo:hi], rw[lo:hi], bt[lo:hi], self.waves[wlo:whi]))
            lo, wlo = hi, whi
        assert lo == len(table) and wlo == len(waves)
        return dict(launches=launches, H1=H1, flags=flags.pop(),
                    unit=self.buckets[0]['model'].unit_convert)

    @staticmethod
    def _check_tables(waves, offs, nP, nY, J=STEPS_PER_PASS):
        """Host check that every wavefront's problems and every offset stay
        inside the packed buffers the kernel reads, and that the output rows
        are a permutation."""
        B = len(offs)
        assert sorted(o[2] for o in offs) == list(range(B))
        for first, count, M, G in waves:
            assert G * count <= 64 and G >= M + J and 0 <= first and first + count <= B
            for b in range(first, first + count):
                p, y, _ = offs[b]
                assert 0 <= p and p + M * M <= nP and 0 <= y and y + M <= nY

    def hamiltonian_inputs(self, n):
        """(H1 [B, n, n], rw [B]) device inputs of the n-site bucket."""
        for bk in self.buckets:
            if bk['n'] == n:
                return bk['inputs']
        raise KeyError(n)

    def _launch(self, sequential=False):
        """The plan's launches; ``sequential`` puts them all on the current
        stream (one after the other, for per-launch timing)."""
        fz = self._fused
        main = torch.cuda.current_stream()
        sides = [main] if sequential else self._sides
        for side in sides:
            side.wait_stream(main)
        if fz is not None:
            # per size class (largest sites first, concurrent on the side
            # streams): the operators, the projection rows and the signals of
            # its wavefronts, which write disjoint rows of W, signal and status
            for j, (top, table, rw, bath, waves) in enumerate(reversed(fz['launches'])):
                with torch.cuda.stream(sides[j % len(sides)]):
                    _native.absorption_operators_into(table, top, fz['H1'], rw, bath,
                                                      fz['flags'], fz['unit'], self.dt, self.J,
                                                      self.P1, self.P)
                    _native.response_bucket_weights_into(waves, self.offsets, self.P1, self.v,
                                                         self.J, self.W)
                    _native.response_buckets_into(waves, self.offsets, self.P, self.y0, self.W,
                                                  self.signal, self.status, J=self.J)
            for side in sides:
                main.wait_stream(side)
        else:
            for j, k in enumerate(self._build_order):
                bk = self.buckets[k]
                with torch.cuda.stream(sides[j % len(sides)]):
                    bk['rebuild'](bk['L'])
                    _native.small_propagator_into(bk['L'], self.dt, 1, self.J, P=bk['P1'],
                                                  Pc=bk['P'])
            for side in sides:
                main.wait_stream(side)
            _native.response_bucket_weights_into(self.waves, self.offsets, self.P1, self.v,
                                                 self.J, self.W)
            _native.response_buckets_into(self.waves, self.offsets, self.P, self.y0, self.W,
                                          self.signal, self.status, J=self.J)
        # the signal rows are in input order; the response's sign goes into
        # the transform's scale
        x = self.signal if self.decay is None else self.signal * self._decay
        f, X = fourier_transform_device(self.t, x, -1, self.rw_freq, self.unit_convert,
                                        scale=-1.0)
        self.f = f
        return X.real

    def capture(self):
        """Records the whole plan into one HIP graph, replayed by run()."""
        from .timing import capture as _capture
        self._launch()                      # warm up allocations and the FFT plan
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with _capture(g):
            self._out = self._launch()
        self._graph = g
        return self

    def run(self):
        """Spectra [len(hamiltonians), len(f)] (device, input order)."""
        if self._graph is not None:
            self._graph.replay()
            return self._out
        return self._launch()

    def check_status(self):
        from .simulate.utils import IntegratorError
        if int(self.status.max().item()) != 0:
            raise IntegratorError('integration failed: non-finite linear-response signal')
