"""Propagation of density matrices without and with classical fields
(qspectra/simulate/eom.py: simulate_dynamics, simulate_with_fields,
simulate_pump)."""
import numpy as np
import torch

from .. import _native
from ..engine import rk4_schedule, schedule_tensors
from ..lti import split_diagonal
from ..pulse import sample_field, field_rate
from .decorators import optional_ensemble_average, optional_2nd_order_isotropic_average
from .utils import (propagate, propagate_heom, propagate_zofe, zofe_projection, IntegratorError,
                    ZVODE_OPTIONS)
from ..dynamics.liouville_space import LinearMap
from ..dynamics.heom import DeviceHEOM
from ..dynamics.zofe import DeviceZOFE


@optional_ensemble_average
def _simulate_dynamics(dynamical_model, initial_state, duration, times, liouville_subspace,
                       save_func, **integrate_kwargs):
    model = dynamical_model
    B = len(model._hamiltonians()) if hasattr(model, '_hamiltonians') else 1
    eom = model.equation_of_motion(liouville_subspace)
    rho = np.asarray(initial_state)
    if rho.ndim == 1:
        # a wavefunction psi becomes outer(conj(psi), psi), as in the reference
        rho = np.outer(rho.conj(), rho)
    y0 = model.density_matrix_to_state_vector(rho, liouville_subspace)
    t = np.arange(0, duration, model.time_step) if times is None else np.asarray(times)
    kw = {k: integrate_kwargs[k] for k in ('dt', 'precision', 'rtol', 'schedule')
          if k in integrate_kwargs}
    if isinstance(eom, DeviceZOFE):
        # rho and the auxiliary operators; density matrices from rho
        B = eom.batch
        y = _native.to_device(np.broadcast_to(y0, (B, 1, y0.shape[-1])).copy())
        proj, scalar = (zofe_projection(eom, save_func) if isinstance(save_func, LinearMap)
                        else (None, False))
        out = propagate_zofe(eom, y, t, dt=kw.get('dt'), proj=proj,
                             rtol=kw.get('rtol', 1e-6))[:, 0].cpu().numpy()   # [B, nt, *]
        if proj is not None:
            return t, out[..., 0] if scalar else out
        if save_func is not None:
            return t, np.stack([np.asarray([save_func(v) for v in ob]) for ob in out])
        return t, np.stack([model.state_vector_to_density_matrix(ob) for ob in out])
    if isinstance(eom, DeviceHEOM):
        # hierarchy: the reduced density matrix is ADO 0 of the state
        B = eom.batch
        y = _native.to_device(np.broadcast_to(y0, (B, y0.shape[-1])).copy())
        ado0 = propagate_heom(eom, y, t, dt=kw.get('dt'), precision=kw.get('precision', 'fp64'),
                              rtol=kw.get('rtol', 1e-6)).cpu().numpy()
        if save_func is not None:
            return t, np.stack([np.asarray([save_func(v) for v in ob]) for ob in ado0])
        lm = model.lspace_model
        return t, np.stack([lm.state_vector_to_density_matrix(ob) for ob in ado0])
    proj = None
    if isinstance(save_func, LinearMap):
        proj = _native.to_device(np.atleast_2d(save_func.matrix)[None])
    y = _native.to_device(np.broadcast_to(y0, (B, 1, y0.shape[-1])).copy())
    out = propagate(eom, y, t, proj=proj, **kw)[:, 0].cpu().numpy()    # [B, nt, K]
    if save_func is not None and not isinstance(save_func, LinearMap):
        out = np.stack([np.asarray([save_func(v) for v in ob]) for ob in out])
        return t, out
    if save_func is not None:
        return t, out
    return t, np.stack([model.state_vector_to_density_matrix(ob) for ob in out])


def simulate_dynamics(dynamical_model, initial_state, duration=None, times=None,
                      liouville_subspace='ee', save_func=None, ensemble_size=None,
                      ensemble_random_orientations=False, **integrate_kwargs):
    """Field-free dynamics from ``initial_state`` (a density matrix, or a
    wavefunction psi taken as outer(conj psi, psi)); returns (t, rho(t))."""
    return _simulate_dynamics(dynamical_model, initial_state, duration, times,
                              liouville_subspace, save_func, ensemble_size=ensemble_size,
                              ensemble_random_orientations=ensemble_random_orientations,
                              **integrate_kwargs)


# ------------------------------------------------------------ with fields
def stage_times(t, t0, n, h):
    """Start time of every RK4 step of the schedule (n, h) from t0 to the
    output times t: interval s starts at t[s-1] (t0 for the first) and its
    k-th step at that start + k h[s]."""
    starts = np.concatenate([[float(t0)], np.asarray(t, dtype=float)[:-1]])
    n = np.asarray(n, dtype=np.int64)
    first = np.concatenate([[0], np.cumsum(n)[:-1]])
    k = np.arange(int(n.sum())) - np.repeat(first, n)
    return np.repeat(starts, n) + k * np.repeat(h, n), np.repeat(h, n)


def field_samples(terms, rw_freq, tg, hg):
    """F [nF, 3, P] complex128: f_i = -i E_i (E_i conjugated for a '+'
    interaction) at t_g, t_g + h/2 and t_g + h of every step g."""
    grid = np.stack([tg, tg + hg / 2, tg + hg], axis=1)          # [nF, 3]
    F = np.empty(grid.shape + (len(terms),), dtype=complex)
    for i, (pulse, trans) in enumerate(terms):
        E = sample_field(pulse, grid, rw_freq)
        F[..., i] = -1j * (np.conj(E) if trans == '+' else E)
    return F


def _field_rate_bound(L, C, terms, rw_freq):
    """Rate bound of the driven equation: the largest absolute row sum of L
    plus, for each field term, its peak amplitude times that of C_i, plus
    the fastest relative rate of change of the fields."""
    rho = float(torch.max(torch.sum(torch.abs(L), dim=-1)).item())
    rate = 0.0
    for i, (pulse, _) in enumerate(terms):
        amp, r = field_rate(pulse, rw_freq)
        rho += amp * float(np.abs(C[:, i]).sum(axis=-1).max())
        rate = max(rate, r)
    return rho + rate


def _field_step(L, C, terms, rw_freq, span, rtol):
    """RK4 step for the driven equation with estimated global error <= rtol
    over ``span`` for any state, from the rate bound (as step_size)."""
    rho = _field_rate_bound(L, C, terms, rw_freq)
    if rho == 0 or span <= 0:
        return None
    return (120.0 * rtol / (rho * span)) ** 0.25 / rho


def _heom_with_fields(model, eom, pulses, geometry, polarization, time_extra, times,
                      liouville_subspace, save_func, dt, precision, rtol):
    """simulate_with_fields of a HEOM model (eom.py:79-103): the dipole
    commutators act on every ADO (heom.py:22-58), the hierarchy in its dense
    form (several Liouville blocks, M <= 64); (t, full hierarchy states
    [B, len(t), n_ado * M]) as the reference's integrate returns them."""
    from .response import _members, _op_matrices
    from .utils import calibrate_step, heom_rate
    pairs = list(zip(polarization, geometry))
    terms = list(zip(pulses, geometry))[:len(pairs)]
    pairs = pairs[:len(terms)]
    if not terms:
        raise ValueError('at least one pulse is needed')
    if len(terms) > _native.MAX_FIELD_TERMS:
        raise ValueError('at most %d field terms are supported' % _native.MAX_FIELD_TERMS)
    hams = _members(model)
    B, M = len(hams), eom.M
    dev = eom.Hop.device
    mats = [_op_matrices(model, liouville_subspace, pol, trans, 'commutator')
            for pol, trans in pairs]                                   # each [B|1, M, M]
    Bc = max(m.shape[0] for m in mats)
    C = np.stack([np.broadcast_to(m, (Bc, M, M)) for m in mats], axis=1)   # [B|1, P, M, M]
    Cd = _native.to_device(np.ascontiguousarray(C), dev=dev)
    rho0 = (model.thermal_states(liouville_subspace) if hasattr(model, 'thermal_states')
            else np.asarray(model.thermal_state(liouville_subspace))[None])
    y0 = _native.to_device(np.ascontiguousarray(np.broadcast_to(rho0, (B, eom.size))), dev=dev)
    if eom.n_comm != 0:
        # one Liouville block: a dipole commutator maps it out of itself, so
        # its restriction (the field term) vanishes and the hierarchy evolves
        # freely, as the reference's integrate of eom + 0 does
        if np.any(C != 0):
            raise NotImplementedError('field terms inside one Liouville block')
        t0 = min(p.t_init for p in pulses)
        tf = max(p.t_final for p in pulses)
        t = (np.arange(t0, tf + time_extra, model.time_step) if times is None
             else tf + np.asarray(times, dtype=float))
        t = np.atleast_1d(np.asarray(t, dtype=float))
        out = propagate_heom(eom, y0, t, t0, dt, precision, full_states=True, rtol=rtol)
        if save_func is not None:
            arr = out.cpu().numpy()
            return t, np.stack([np.asarray([save_func(v) for v in ob]) for ob in arr])
        return t, out
    t0 = min(p.t_init for p in pulses)
    tf = max(p.t_final for p in pulses)
    t = (np.arange(t0, tf + time_extra, model.time_step) if times is None
         else tf + np.asarray(times, dtype=float))
    t = np.atleast_1d(np.asarray(t, dtype=float))
    if np.any(t < t0):
        raise ValueError('output times must not precede the first pulse (t=%g)' % t0)
    rw_freq = model.rw_freq
    mode = _native.QSX_FP64 if precision in ('fp64', 'double', None) else None
    if precision == 'mixed':
        mode = _native.QSX_MIXED
    if mode is None:
        raise ValueError("precision must be 'fp64' or 'mixed'")

    def run(y, tp, t_start, h):
        ni, hi = rk4_schedule(tp, t_start, h)
        tg, hg = stage_times(tp, t_start, ni, hi)
        if int(ni.sum()) == 0:
            return y, 0
        F = _native.to_device(field_samples(terms, rw_freq, tg, hg)[None], dev=dev)
        _, yf, st = _native.heom_propagate_driven(eom, Cd, F, y, ni, hi, mode, want_final=True)
        return yf, int(st.item())

    if dt is None:
        span = float(t[-1]) - float(t0)
        rho = heom_rate(eom)
        field = 0.0
        for i, (pulse, _) in enumerate(terms):
            amp, r = field_rate(pulse, rw_freq)
            rho += amp * float(np.abs(C[:, i]).sum(axis=-1).max())
            field = max(field, r)
        rho += field
        if rho > 0 and span > 0:
            h_safe = (120.0 * rtol / (rho * span)) ** 0.25 / rho

            def probe(H, n_p):
                ends = []
                for hp in (H, H / 2):
                    yf, st = run(y0, np.array([t0 + n_p * H]), t0, hp)
                    if st != 0:
                        return None
                    ends.append(yf)
                return ends
            dt = calibrate_step(probe, h_safe, 2.0 / rho, span, rtol,
                                window=float(tf) - float(t0))
        gaps = np.diff(np.concatenate([[t0], t]))
        gaps = gaps[gaps > 0]
        if gaps.size:
            dt = gaps.min() if dt is None else min(dt, float(gaps.min()))
    states, y, start = [], y0, float(t0)
    for tk in t:
        y, st = run(y, np.array([tk]), start, dt)
        if st != 0:
            raise IntegratorError('integration failed: non-finite state between t=%g and t=%g'
                                  % (start, tk))
        states.append(y)
        start = float(tk)
    out = torch.stack(states, dim=1)                                   # [B, nt, S]
    if save_func is not None:
        arr = out.cpu().numpy()
        return t, np.stack([np.asarray([save_func(v) for v in ob]) for ob in arr])
    return t, out


def _simulate_with_fields(dynamical_model, pulses, geometry, polarization, time_extra, times,
                          liouville_subspace, save_func, **integrate_kwargs):
    """(t, states [B, len(t), ...]) for a (possibly batched) model; the
    terms pair pulses, geometry and polarization in order (as zip does)."""
    from .response import _members, _op_matrices
    model = dynamical_model
    dt = integrate_kwargs.pop('dt', None)
    precision = integrate_kwargs.pop('precision', 'fp64')
    rtol = integrate_kwargs.pop('rtol', 1e-6)
    unknown = set(integrate_kwargs) - set(ZVODE_OPTIONS)
    if unknown:
        raise TypeError('unexpected integrate options: %s' % sorted(unknown))
    eom = model.equation_of_motion(liouville_subspace)
    if isinstance(eom, DeviceHEOM):
        return _heom_with_fields(model, eom, pulses, geometry, polarization, time_extra, times,
                                 liouville_subspace, save_func, dt, precision, rtol)
    if isinstance(eom, DeviceZOFE):
        return _zofe_with_fields(model, eom, pulses, geometry, polarization, time_extra, times,
                                 save_func, dt, rtol)
    pairs = list(zip(polarization, geometry))
    terms = list(zip(pulses, geometry))[:len(pairs)]
    pairs = pairs[:len(terms)]
    if not terms:
        raise ValueError('at least one pulse is needed')
    if len(terms) > _native.MAX_FIELD_TERMS:
        raise ValueError('at most %d field terms are supported' % _native.MAX_FIELD_TERMS)
    M = eom.size
    if M > _native.MAX_DRIVEN_BLOCK:
        raise ValueError('field-driven propagation supports Liouville subspaces of at most %d '
                         'states (%r has %d)' % (_native.MAX_DRIVEN_BLOCK, liouville_subspace, M))
    hams = _members(model)
    B = len(hams)
    mats = [_op_matrices(model, liouville_subspace, pol, trans, 'commutator')
            for pol, trans in pairs]                                   # each [B|1, M, M]
    Bc = max(m.shape[0] for m in mats)
    C = np.stack([np.broadcast_to(m, (Bc, M, M)) for m in mats], axis=1)   # [B|1, P, M, M]

    rho0 = (model.thermal_states(liouville_subspace) if hasattr(model, 'thermal_states')
            else np.asarray(model.thermal_state(liouville_subspace))[None])
    t0 = min(p.t_init for p in pulses)
    tf = max(p.t_final for p in pulses)
    t = (np.arange(t0, tf + time_extra, model.time_step) if times is None
         else tf + np.asarray(times, dtype=float))
    t = np.atleast_1d(np.asarray(t, dtype=float))
    if np.any(t < t0):
        raise ValueError('output times must not precede the first pulse (t=%g)' % t0)
    rw_freq = model.rw_freq
    dev = eom.L.device
    y0 = _native.to_device(rho0[:, None, :], dev=dev)                  # [B, 1, M]
    col, coef = _native.ell_pack(C)
    col = torch.as_tensor(col, device=dev)
    if dt is None:
        span = float(t[-1]) - float(t0)
        h_safe = _field_step(eom.L, C, terms, rw_freq, span, rtol)
        coef64 = _native.to_device(coef, dev=dev)

        def probe(H, n_p):
            # n_p steps of H and 2 n_p of H / 2 from the thermal states, the
            # fields sampled on each schedule (FP64)
            ends = []
            for hp in (H, H / 2):
                tp = np.array([t0 + n_p * H])
                ni, hi = rk4_schedule(tp, t0, hp)
                Fp = field_samples(terms, rw_freq, *stage_times(tp, t0, ni, hi))
                nsub_p, hs_p = schedule_tensors(tp, t0, hp, dev)
                o, st = _native.propagate_driven(eom.L, col, coef64,
                                                 _native.to_device(Fp[None], dev=dev), y0,
                                                 nsub_p, hs_p, _native.QSX_FP64)
                if int(st.item()) != 0:
                    return None
                ends.append(o[:, :, -1])
            return ends
        from .utils import calibrate_step
        rho_rate = _field_rate_bound(eom.L, C, terms, rw_freq)
        dt = calibrate_step(probe, h_safe, 2.0 / rho_rate if rho_rate > 0 else span, span,
                            rtol, window=float(tf) - float(t0))
        gaps = np.diff(np.concatenate([[t0], t]))
        gaps = gaps[gaps > 0]
        if gaps.size:
            dt = gaps.min() if dt is None else min(dt, float(gaps.min()))
    n, h = rk4_schedule(t, t0, dt)
    tg, hg = stage_times(t, t0, n, h)
    F = field_samples(terms, rw_freq, tg, hg)
    if F.shape[0] == 0:
        F = np.zeros((1, 3, len(terms)), dtype=complex)                # nothing to step
    nsub, hs = schedule_tensors(t, t0, dt, dev)
    Ft = _native.to_device(F[None], dev=dev)
    if precision in ('fp64', 'double', None):
        out, status = _native.propagate_driven(eom.L, col, _native.to_device(coef, dev=dev), Ft,
                                               y0, nsub, hs, _native.QSX_FP64)
    elif precision == 'mixed':
        Lm, diag = split_diagonal(eom.L)
        cm = _native.to_device(coef, dtype=torch.complex64, dev=dev)
        out, status = _native.propagate_driven(Lm, col, cm, Ft, y0, nsub, hs, _native.QSX_MIXED,
                                               diag=diag)
    else:
        raise ValueError("precision must be 'fp64' or 'mixed'")
    if int(status.item()) != 0:
        raise IntegratorError('integration failed: non-finite state between t=%g and t=%g'
                              % (t0, t[-1]))
    states = out[:, 0]                                                 # [B, nt, M]
    if isinstance(save_func, LinearMap):
        mat = np.asarray(save_func.matrix)
        proj = _native.to_device(np.atleast_2d(mat), dev=dev)
        saved = _native.zgemm(states.contiguous(), proj[None], transb=True)
        return t, saved[..., 0] if mat.ndim == 1 else saved
    if save_func is not None:
        states = states.cpu().numpy()
        return t, np.stack([np.asarray([save_func(v) for v in ob]) for ob in states])
    return t, states                       # device tensor: ensembles average on the GPU


def _zofe_with_fields(model, eom, pulses, geometry, polarization, time_extra, times,
                      save_func, dt, rtol):
    """simulate_with_fields of a ZOFE model: the field terms act through
    the ZOFE commutator (on rho and on every auxiliary operator) inside the
    ZOFE kernel."""
    from .response import _members
    pairs = list(zip(polarization, geometry))
    terms = list(zip(pulses, geometry))[:len(pairs)]
    pairs = pairs[:len(terms)]
    if not terms:
        raise ValueError('at least one pulse is needed')
    if len(terms) > _native.MAX_FIELD_TERMS:
        raise ValueError('at most %d field terms are supported' % _native.MAX_FIELD_TERMS)
    hams = _members(model)
    V = np.stack([np.stack([np.asarray(h.dipole_operator(model.hilbert_subspace, pol, trans),
                                       dtype=complex) for pol, trans in pairs])
                  for h in hams])                                      # [B, P, N, N]
    if all(np.array_equal(V[0], v) for v in V[1:]):
        V = V[:1]
    rho0 = np.stack([model.density_matrix_to_state_vector(
        h.thermal_state(model.hilbert_subspace)) for h in hams])
    t0 = min(p.t_init for p in pulses)
    tf = max(p.t_final for p in pulses)
    t = (np.arange(t0, tf + time_extra, model.time_step) if times is None
         else tf + np.asarray(times, dtype=float))
    t = np.atleast_1d(np.asarray(t, dtype=float))
    if np.any(t < t0):
        raise ValueError('output times must not precede the first pulse (t=%g)' % t0)
    rw_freq = model.rw_freq
    rate = 0.0
    for i, (pulse, _) in enumerate(terms):
        amp, r = field_rate(pulse, rw_freq)
        rate += amp * float(np.abs(V[:, i]).sum(axis=-1).max()) * 2 + r
    dev = eom.H.device
    Vd = _native.to_device(V, dev=dev)

    def fields(t_out, t_start, step):
        n_p, h_p = rk4_schedule(t_out, t_start, step)
        F_p = field_samples(terms, rw_freq, *stage_times(t_out, t_start, n_p, h_p))
        return Vd, _native.to_device(F_p[None], dev=dev)

    y0 = _native.to_device(rho0[:, None, :], dev=dev)
    if dt is None:
        from .utils import zofe_step
        dt = zofe_step(eom, float(t[-1]) - float(t0), rtol, rate, y0=y0, fields=fields,
                       t0=float(t0), window=float(tf) - float(t0))
        gaps = np.diff(np.concatenate([[t0], t]))
        gaps = gaps[gaps > 0]
        if gaps.size:
            dt = gaps.min() if dt is None else min(dt, float(gaps.min()))
    n, h = rk4_schedule(t, t0, dt)
    tg, hg = stage_times(t, t0, n, h)
    F = field_samples(terms, rw_freq, tg, hg)
    if F.shape[0] == 0:
        F = np.zeros((1, 3, len(terms)), dtype=complex)
    proj, scalar = (zofe_projection(eom, save_func) if isinstance(save_func, LinearMap)
                    else (None, False))
    out = propagate_zofe(eom, y0, t, t0, dt, proj=proj, V=Vd,
                         F=_native.to_device(F[None], dev=dev))[:, 0]      # [B, nt, *]
    if proj is not None:
        return t, out[..., 0] if scalar else out
    if save_func is not None:
        states = out.cpu().numpy()
        return t, np.stack([np.asarray([save_func(v) for v in ob]) for ob in states])
    return t, out


def simulate_with_fields(dynamical_model, pulses, geometry='-+', polarization='xx',
                         time_extra=0, times=None, liouville_subspace='gg,ge,eg,ee',
                         save_func=None, ensemble_size=None, ensemble_random_orientations=False,
                         **integrate_kwargs):
    """Time evolution under classical pulses in the rotating wave
    approximation (qspectra/simulate/eom.py:105-159): from the thermal
    state at the start of the first pulse, y' = L y + sum_i f_i(t) [V_i, y]
    with f_i = -i E_i (conjugated for '+').  Output times: every model time
    step from the first pulse's start to the last pulse's end + time_extra,
    or the last pulse's end + ``times``.  Returns (t, states) -- state
    vectors, or save_func of them.  Keyword options: dt (RK4 step), rtol
    (accuracy of the default step), precision ('fp64' | 'mixed')."""
    return optional_ensemble_average(_simulate_with_fields)(
        dynamical_model, pulses, geometry, polarization, time_extra, times, liouville_subspace,
        save_func, ensemble_size=ensemble_size,
        ensemble_random_orientations=ensemble_random_orientations, **integrate_kwargs)


def simulate_pump(dynamical_model, pump, polarization='x', time_extra=0, times=None,
                  liouville_subspace='gg,ge,eg,ee', save_func=None, ensemble_size=None,
                  ensemble_random_orientations=False, exact_isotropic_average=False,
                  **integrate_kwargs):
    """Time evolution under one pump pulse acting twice ('-+', the same
    polarization), qspectra/simulate/eom.py:162-227; exact_isotropic_average
    averages over orientations to second order in the field."""
    return optional_ensemble_average(optional_2nd_order_isotropic_average(_simulate_with_fields))(
        dynamical_model, [pump, pump], '-+', [polarization, polarization], time_extra, times,
        liouville_subspace, save_func, ensemble_size=ensemble_size,
        ensemble_random_orientations=ensemble_random_orientations,
        exact_isotropic_average=exact_isotropic_average, **integrate_kwargs)
