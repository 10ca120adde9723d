"""Gradients of Redfield absorption spectra, third-order responses and 2-D
spectra with respect to site energies, couplings and the reorganization
energy (reverse mode on the GPU; RedfieldAbsorption, RedfieldThirdOrder).

The reference computes spectra only (response.py:103-154); this module makes
the same computation differentiable:

    theta = (E [N], J [N(N-1)/2], lam)
      -> H_n = diag(E) + J + diag(delta_n)       (member n of the static-
         disorder ensemble; delta_n is drawn as ElectronicHamiltonian.sample(n)
         draws it, hamiltonian.py:552-578, and does not depend on theta)
      -> L_n  Redfield 'eg' block in the site basis (redfield.py:9-104)
      -> P_n = P4(h L_n^T)                        (Heisenberg picture, one RK4 step)
      -> x_n(t_k) = (P_n^k y_0) . v               (linear response, response.py:15-43)
      -> S(w) = Re FT[-mean_n x_n]                (absorption, response.py:103-154)

Values come from the product kernels (qsx_redfield_block, qsx_zgemm,
qsx_propagate); the reverse pass uses
  * qsx_propagate_adjoint (checkpointed recomputation of the forward states)
    for the time propagation,
  * the closed-form vector-Jacobian product of the Horner polynomial P4 (on
    qsx_zgemm),
  * the hand-written reverse mode of the Redfield block
    (qsx_redfield_block_vjp: basis change, Redfield tensor, C(w) and the
    symmetric eigenproblem) for dL/dtheta.
"""
import numpy as np
import torch

from . import _native
from .constants import CM_FS
from .hamiltonian import ElectronicHamiltonian, diagonal_gaussian_disorder
from .bath import DebyeBath
from .dynamics.redfield import RedfieldModel
from .simulate.response import _op_matrices
from .simulate.utils import fourier_transform_device
from .utils import check_random_state


# --------------------------------------------------------- time propagation
class _Propagate(torch.autograd.Function):
    """states [B, R, nsave, M] of y <- A y (A = P, the one-step propagator),
    with nsub[s] steps before save s; reverse mode by qsx_propagate_adjoint."""

    @staticmethod
    def forward(ctx, A, y0, nsub_host, ckpt):
        dev = y0.device
        nsub = torch.as_tensor(np.asarray(nsub_host, dtype=np.int32)).to(dev)
        h = torch.ones(len(nsub_host), dtype=torch.float64, device=dev)
        total = int(np.sum(nsub_host))
        states, _, status = _native.propagate_states(
            _native.QSX_MODE_PROPAGATOR, A.contiguous(), y0.contiguous(), nsub, h,
            steps=total)
        ctx.save_for_backward(A, y0, nsub)
        ctx.total, ctx.ckpt = total, ckpt
        ctx.mark_non_differentiable(status)
        return states, status

    @staticmethod
    def backward(ctx, gstates, _gstatus):
        A, y0, nsub = ctx.saved_tensors
        gA, gy0, status = _native.propagate_adjoint(A, y0, nsub, ctx.total,
                                                    gstates.contiguous(), ctx.ckpt)
        if int(status.item()) != 0:
            raise FloatingPointError('adjoint propagation failed (status %d)'
                                     % int(status.item()))
        gA = gA.sum(dim=1)
        if A.shape[0] == 1 and gA.shape[0] > 1:
            gA = gA.sum(dim=0, keepdim=True)
        return gA, gy0, None, None


class _RK4Propagator(torch.autograd.Function):
    """P = P4(h L) = I + hL + (hL)^2/2 + (hL)^3/6 + (hL)^4/24 for L [B, M, M].

    Reverse mode: with X = h L^H and G = dLoss/dP,
        dLoss/dL = h (T1 + T2/2 + T3/6 + T4/24),  T1 = G,
        T_{m+1} = X T_m + G X^m
    (T_m = sum_i X^i G X^(m-1-i), the adjoint of d(L^m) = sum_i L^i dL L^(m-1-i))."""

    @staticmethod
    def forward(ctx, L, h):
        from .lti import rk4_propagator
        L = L.contiguous()
        ctx.save_for_backward(L)
        ctx.h = float(h)
        return rk4_propagator(L, ctx.h)

    @staticmethod
    def backward(ctx, G):
        (L,) = ctx.saved_tensors
        h = ctx.h
        G = G.contiguous()
        X = (L.conj().transpose(-1, -2) * h).contiguous()
        T = G
        Xm = X
        acc = G.clone()
        for m, fact in ((2, 2.0), (3, 6.0), (4, 24.0)):
            T = _native.zgemm(X, T)
            T = _native.zgemm(G, Xm, T, beta=1.0)
            acc = acc + T / fact
            if m < 4:
                Xm = _native.zgemm(Xm, X)
        return acc * h, None


def rk4_propagator(L, h):
    """Differentiable P4(h L) for L [B, M, M] complex128 on the device."""
    return _RK4Propagator.apply(L, h)


def propagate(A, y0, nsub, ckpt=64):
    """Differentiable propagation with the one-step propagator A [B|1, M, M]:
    states [B, R, len(nsub), M] (nsub: host step counts per save interval).
    The reverse pass recomputes the forward states from checkpoints every
    ckpt steps (blocks up to M = 160)."""
    states, status = _Propagate.apply(A, y0, nsub, int(ckpt))
    if int(status.item()) != 0:
        raise FloatingPointError('propagation produced a non-finite state')
    return states


# ------------------------------------------------------------ Redfield block
class _RedfieldBlock(torch.autograd.Function):
    """Redfield block XY of one-exciton Hamiltonians H1 [B, n, n] (float64)
    and a Debye bath: values from qsx_redfield_block, cotangents of H1 and of
    the reorganization energy from its hand-written reverse mode
    (qsx_redfield_block_vjp: basis change, Redfield tensor, C(w) and the
    symmetric eigenproblem's VJP on the GPU)."""

    @staticmethod
    def forward(ctx, H1, lam, rw, T, gamma, X, Y, flags, unit_convert):
        H1d = H1.detach().contiguous()
        lam_v = float(lam)
        ctx.args = (rw, T, lam_v, gamma, X, Y, flags, unit_convert)
        ctx.lam_is_tensor = torch.is_tensor(lam)
        ctx.save_for_backward(H1d)
        return _native.redfield_block(H1d, rw, T, lam_v, gamma, X, Y, flags, unit_convert)

    @staticmethod
    def backward(ctx, gL):
        (H1,) = ctx.saved_tensors
        rw, T, lam, gamma, X, Y, flags, unit_convert = ctx.args
        gH1, glam = _native.redfield_block_vjp(H1, rw, T, lam, gamma, X, Y, flags, unit_convert,
                                               gL.contiguous())
        g_lam = glam.sum() if ctx.lam_is_tensor else None
        return gH1, g_lam, None, None, None, None, None, None, None


def redfield_block(H1, lam, gamma, T, rw, X, Y, flags, unit_convert=CM_FS):
    """Differentiable Redfield block (see _RedfieldBlock); rw [B] float64
    tensor, X, Y Hilbert block indices (0 g, 1 e, 2 f), flags the QSX_* bits
    of qsx_redfield_block."""
    return _RedfieldBlock.apply(H1, lam, rw, float(T), float(gamma), int(X), int(Y),
                                int(flags), float(unit_convert))


# ------------------------------------------------------------ spectra
def site_hamiltonian(E, J):
    """H1 = diag(E) + J with J the upper triangle (i < j, row-major order)
    mirrored below the diagonal; E [N], J [N(N-1)/2] tensors."""
    N = E.shape[0]
    iu = torch.triu_indices(N, N, offset=1, device=E.device)
    H = torch.zeros((N, N), dtype=torch.float64, device=E.device)
    H = H.index_put((iu[0], iu[1]), J).index_put((iu[1], iu[0]), J)
    return H + torch.diag(E)


def disorder_shifts(n_sites, fwhm, ensemble_size, random_seed=0):
    """Diagonal static disorder of ensemble members 0..ensemble_size-1 as
    ElectronicHamiltonian.sample_ensemble draws it: member n uses
    RandomState(list(atleast_1d(random_seed)) + [n])."""
    out = np.zeros((ensemble_size, n_sites))
    if fwhm is None:
        return out
    draw = diagonal_gaussian_disorder(fwhm, n_sites)
    for n in range(ensemble_size):
        rs = check_random_state(list(np.atleast_1d(random_seed)) + [n])
        out[n] = np.diag(draw(rs))
    return out


class RedfieldAbsorption(object):
    """Differentiable absorption spectrum of a Redfield aggregate ('ge'
    Hilbert space, 'eg' coherence), optionally averaged over a static-
    disorder ensemble evaluated as one batch.

        spec = RedfieldAbsorption(dipoles, cutoff_freq, temperature, rw_freq,
                                  time_max, dt, disorder=100 / GAUSSIAN_SD_FWHM,
                                  ensemble_size=512)
        f, S = spec(E, J, lam)        # S: torch tensor, differentiable

    E [N] site energies, J [N(N-1)/2] couplings (upper triangle), lam the
    reorganization energy (cm^-1), each a float64 tensor on the device.
    rw_freq is fixed (so the frequency axis does not move with theta).  The
    forward values equal absorption_spectra(model, time_max, dt=dt,
    ensemble_size=...) of the RedfieldModel with the same parameters."""

    def __init__(self, dipoles, cutoff_freq, temperature, rw_freq, time_max, dt,
                 disorder=None, ensemble_size=None, random_seed=0, secular=False,
                 discard_imag_corr=False, polarization='xx', unit_convert=CM_FS, ckpt=64,
                 correlation_decay_time=None):
        self.dipoles = np.asarray(dipoles, dtype=float)
        self.N = self.dipoles.shape[0]
        self.gamma, self.T = float(cutoff_freq), float(temperature)
        self.rw_freq, self.dt = float(rw_freq), float(dt)
        self.secular, self.discard_imag_corr = secular, discard_imag_corr
        self.polarization, self.unit_convert, self.ckpt = polarization, unit_convert, ckpt
        self.size = ensemble_size or 1
        self.shifts = disorder_shifts(self.N, disorder if ensemble_size else None, self.size,
                                      random_seed)
        self.t = np.arange(0, time_max, dt)
        self.decay = correlation_decay_time
        dev = _native.device()
        self.dev = dev
        # dipole maps in the site basis do not depend on theta
        ham = ElectronicHamiltonian(np.eye(self.N), bath=DebyeBath(self.T, 1.0, self.gamma),
                                    dipoles=self.dipoles)
        model = RedfieldModel(ham, rw_freq=self.rw_freq, hilbert_subspace='ge',
                              unit_convert=unit_convert)
        V0 = _op_matrices(model, 'gg->eg', polarization[0], '+', 'commutator')[0]   # [M, 1]
        bra = _op_matrices(model, 'eg->gg', polarization[1], '-', 'bra')[0]          # [M]
        self.v = torch.as_tensor(V0[:, 0] * model.thermal_state('gg')[0]).to(dev)
        self.y0 = torch.as_tensor(-bra).to(dev).reshape(1, 1, -1).expand(
            self.size, 1, -1).contiguous()
        self.nsub = np.ones(len(self.t), dtype=np.int64)
        self.nsub[0] = 0

    def hamiltonians(self, E, J):
        H = site_hamiltonian(E, J)
        return H[None] + torch.diag_embed(torch.as_tensor(self.shifts).to(H.device))

    def liouvillian(self, H1, lam):
        """Heisenberg 'eg' blocks [B, M, M] (qsx_redfield_block; reverse mode
        qsx_redfield_block_vjp)."""
        flags = ((_native.QSX_SECULAR if self.secular else 0)
                 | (_native.QSX_REAL_CORR if self.discard_imag_corr else 0)
                 | _native.QSX_HEISENBERG)
        rw = torch.full((H1.shape[0],), self.rw_freq, dtype=torch.float64, device=H1.device)
        return redfield_block(H1.contiguous(), lam, self.gamma, self.T, rw, 1, 0, flags,
                              self.unit_convert)

    def signal(self, E, J, lam):
        """Ensemble-mean linear response x(t) [len(t)] (complex128)."""
        H1 = self.hamiltonians(E, J)
        L = self.liouvillian(H1, lam)
        P = rk4_propagator(L, self.dt)
        states = propagate(P, self.y0, self.nsub, self.ckpt)       # [B, 1, nt, M]
        x = torch.einsum('bkm,m->bk', states[:, 0], self.v)
        return x.mean(dim=0)

    def __call__(self, E, J, lam):
        x = self.signal(E, J, lam)
        if self.decay is not None:
            x = x * torch.as_tensor(np.exp(-self.t / self.decay)).to(x.device)
        f, X = fourier_transform_device(self.t, -x, -1, self.rw_freq, self.unit_convert)
        return f, X.real


_BLOCK_INDEX = {'g': 0, 'e': 1, 'f': 2}


class RedfieldThirdOrder(object):
    """Differentiable third-order response R(t1, t2, t3) and 2-D spectrum of
    a Redfield aggregate ('gef' Hilbert space), optionally averaged over a
    static-disorder ensemble evaluated as one batch.

        tds = RedfieldThirdOrder(dipoles, cutoff_freq, temperature, rw_freq,
                                 coherence_time_max, population_times, dt)
        R = tds.response(E, J, lam)        # [n1, n2, n3] complex128
        (f1, t2, f3), X = tds(E, J, lam)   # the 2-D spectrum

    Each Liouville pathway (response.py:269-337, the pathways of
    THIRD_ORDER_PATHWAYS[geometry]) is evaluated as the reference orders it:
    the coherence rho1 = V0 rho0 propagated over t1, V1 applied at every t1,
    those states propagated to the population times, V2 applied, and the
    contraction with the t3 bra propagated in the Heisenberg picture.  Every
    block is the differentiable Redfield block (qsx_redfield_block, reverse
    mode qsx_redfield_block_vjp), every propagation the differentiable
    one-step-propagator recursion (qsx_propagate, reverse mode
    qsx_propagate_adjoint; blocks up to 160 elements, so up to seven sites).
    The forward values equal third_order_response / two_dimensional_spectra
    of the RedfieldModel with the same parameters at the step dt
    (two_dimensional_spectra: response.py:430-455, the t1 transform with
    sign -1, the t3 transform with sign +1)."""

    def __init__(self, dipoles, cutoff_freq, temperature, rw_freq, coherence_time_max,
                 population_times, dt, geometry='-++', polarization='xxxx', include_signal=None,
                 disorder=None, ensemble_size=None, random_seed=0, secular=False,
                 discard_imag_corr=False, unit_convert=CM_FS, ckpt=64):
        from .simulate.response import THIRD_ORDER_PATHWAYS, _parse_pathways
        self.dipoles = np.asarray(dipoles, dtype=float)
        self.N = self.dipoles.shape[0]
        self.gamma, self.T = float(cutoff_freq), float(temperature)
        self.rw_freq, self.dt = float(rw_freq), float(dt)
        self.secular, self.discard_imag_corr = secular, discard_imag_corr
        self.unit_convert, self.ckpt = unit_convert, ckpt
        self.size = ensemble_size or 1
        self.shifts = disorder_shifts(self.N, disorder if ensemble_size else None, self.size,
                                      random_seed)
        self.t1 = np.arange(0, coherence_time_max, dt)
        self.t3 = self.t1
        self.t2 = np.asarray(population_times, dtype=float)
        steps2 = np.round(self.t2 / dt).astype(np.int64)
        if np.any(np.abs(steps2 * dt - self.t2) > 1e-9 * max(1.0, dt)) or \
                np.any(np.diff(steps2) < 0) or steps2[0] < 0:
            raise ValueError('population times must be ascending multiples of dt')
        self.nsub1 = np.ones(len(self.t1), dtype=np.int64)
        self.nsub1[0] = 0
        self.nsub2 = np.diff(np.concatenate([[0], steps2]))
        dev = _native.device()
        self.dev = dev
        # dipole maps and the ground state in the site basis do not depend on
        # theta
        ham = ElectronicHamiltonian(np.eye(self.N), bath=DebyeBath(self.T, 1.0, self.gamma),
                                    dipoles=self.dipoles)
        model = RedfieldModel(ham, rw_freq=self.rw_freq, hilbert_subspace='gef',
                              unit_convert=unit_convert)
        rho0 = model.thermal_state('gg')
        trans = geometry + '-'
        dv = lambda a: torch.as_tensor(np.array(a, dtype=np.complex128)).to(dev)
        self.paths = []
        for path in _parse_pathways(THIRD_ORDER_PATHWAYS[geometry], include_signal):
            s = path.split('->')
            V = [_op_matrices(model, '%s->%s' % (s[i], s[i + 1]), polarization[i], trans[i],
                              'commutator')[0] for i in range(3)]
            b3 = _op_matrices(model, '%s->%s' % (s[3], s[4]), polarization[3], trans[3],
                              'bra')[0]
            y1 = (V[0] @ rho0)[None, None, :]
            self.paths.append(dict(
                blocks=(s[1], s[2], s[3]),
                y1=dv(np.broadcast_to(y1, (self.size, 1, y1.shape[-1]))),
                V1T=dv(V[1].T), V2T=dv(V[2].T),
                b3=dv(np.broadcast_to(b3[None, None, :], (self.size, 1, b3.shape[-1])))))

    def hamiltonians(self, E, J):
        H = site_hamiltonian(E, J)
        return H[None] + torch.diag_embed(torch.as_tensor(self.shifts).to(H.device))

    def _propagator(self, cache, H1, lam, block, heisenberg):
        """P4(dt L) of Liouville block `block` ('eg', 'fe', ...) of every member."""
        key = (block, heisenberg)
        if key not in cache:
            flags = ((_native.QSX_SECULAR if self.secular else 0)
                     | (_native.QSX_REAL_CORR if self.discard_imag_corr else 0)
                     | (_native.QSX_HEISENBERG if heisenberg else 0))
            rw = torch.full((H1.shape[0],), self.rw_freq, dtype=torch.float64,
                            device=H1.device)
            L = redfield_block(H1.contiguous(), lam, self.gamma, self.T, rw,
                               _BLOCK_INDEX[block[0]], _BLOCK_INDEX[block[1]], flags,
                               self.unit_convert)
            cache[key] = rk4_propagator(L, self.dt)
        return cache[key]

    def response(self, E, J, lam):
        """Ensemble-mean third-order response [len(t1), len(t2), len(t3)]."""
        H1 = self.hamiltonians(E, J)
        B, n1, n2, n3 = self.size, len(self.t1), len(self.t2), len(self.t3)
        cache, total = {}, None
        for p in self.paths:
            s1, s2, s3 = p['blocks']
            P1 = self._propagator(cache, H1, lam, s1, False)
            P2 = self._propagator(cache, H1, lam, s2, False)
            P3 = self._propagator(cache, H1, lam, s3, True)
            S1 = propagate(P1, p['y1'], self.nsub1, self.ckpt)[:, 0]           # [B, n1, M1]
            y2 = torch.matmul(S1, p['V1T'])                                    # [B, n1, M2]
            S2 = propagate(P2, y2.contiguous(), self.nsub2, self.ckpt)          # [B, n1, n2, M2]
            y3 = torch.matmul(S2, p['V2T'])                                    # [B, n1, n2, M3]
            G3 = propagate(P3, p['b3'], self.nsub1, self.ckpt)[:, 0]           # [B, n3, M3]
            R = torch.matmul(y3.reshape(B, n1 * n2, -1),
                             G3.transpose(1, 2).contiguous()).reshape(B, n1, n2, n3)
            total = R if total is None else total + R
        return total.mean(dim=0)

    def __call__(self, E, J, lam):
        R = self.response(E, J, lam)
        f1, X = fourier_transform_device(self.t1, R, 0, self.rw_freq, self.unit_convert,
                                         sign=-1)
        f3, X = fourier_transform_device(self.t3, X, 2, self.rw_freq, self.unit_convert)
        return (f1, self.t2, f3), X
