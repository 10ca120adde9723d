"""Linear and third-order response functions (qspectra/simulate/response.py
API), computed on the GPU.

Each Liouville-space pathway is propagated exactly as the reference
structures it (response.py:15-43, 269-337) -- Schroedinger picture over t1
and t2, Heisenberg picture over t3 -- but every stage is one fused RK4
launch: all t1 branches of the t2 stage are right-hand sides of one launch,
all ensemble members are batch entries of the same launch, and the dipole
maps applied at each output time run inside the kernel.  Signals stay on
the device until the end.

Keyword options in addition to the reference's: ``dt`` (RK4 step),
``precision`` ('fp64' | 'mixed'), ``rtol`` (default-step accuracy),
``schedule`` ('auto' | 'rk4' | 'propagator'), and for third-order responses
``plan`` ('fused', the default, see fused.py; or 'pathways', one pathway at a
time as the reference does).
"""
import copy

import numpy as np
import torch

from .. import _native
from .decorators import (optional_ensemble_average, optional_2nd_order_isotropic_average,
                         optional_4th_order_isotropic_average)
from .utils import propagate, propagate_heom, propagate_zofe, fourier_transform
from ..dynamics.heom import DeviceHEOM
from ..utils import ZeroArray

_PROPAGATE_KW = ('dt', 'precision', 'rtol', 'schedule')


def _propagate_kw(kwargs):
    kw = {k: kwargs[k] for k in _PROPAGATE_KW if k in kwargs}
    # zvode tolerances other than rtol have no RK4 meaning; rtol sets the step
    return kw


def _members(model):
    return model._hamiltonians() if hasattr(model, '_hamiltonians') else [model.hamiltonian]


def _member_model(model, ham):
    m = copy.copy(model)
    m.__dict__.pop('_imemoize__cache', None)
    m.__dict__.pop('_batch', None)
    m.hamiltonian = ham
    return m


def _op_matrices(model, liouv_map, polarization, transition, what):
    """Per-member matrices of a dipole map: what = 'commutator' -> [B|1, K, M],
    'bra' -> [B|1, M].  In the site basis the map depends on the member only
    through its dipoles, so it is evaluated once per distinct dipole set (one
    shared copy when all members have the same dipoles)."""
    hams = _members(model)
    site = model.evolve_basis == 'site' if hasattr(model, 'evolve_basis') else True
    if site:
        keys, first = [], {}
        for i, h in enumerate(hams):
            d = np.ascontiguousarray(h.dipoles)
            k = (d.shape, d.tobytes())
            first.setdefault(k, i)
            keys.append(first[k])
        uniq = sorted(set(keys))
    else:
        keys = list(range(len(hams)))
        uniq = keys
    mats = {}
    for i in uniq:
        op = _member_model(model, hams[i]).dipole_operator(liouv_map, polarization, transition)
        op = getattr(op, 'lspace_op', op)       # hierarchy operators act per ADO
        mats[i] = op.commutator.matrix if what == 'commutator' else op.bra_vector
    if len(uniq) == 1:
        return np.stack([mats[uniq[0]]])
    return np.stack([mats[k] for k in keys])


def _dev(x):
    return _native.to_device(x)


def _expand(x, B):
    return x if x.shape[0] == B else x.expand((B,) + tuple(x.shape[1:])).contiguous()


def _zofe_ops(model, polarization, transition):
    """Dipole operator matrices of every member of a ZOFE model, [B|1, N, N]
    (one shared copy when the members' dipoles agree)."""
    hams = _members(model)
    ops = np.stack([np.asarray(h.dipole_operator(model.hilbert_subspace, polarization,
                                                 transition), dtype=complex) for h in hams])
    return ops[:1] if all(np.array_equal(ops[0], o) for o in ops[1:]) else ops


def _zofe_commutator(model, states, V):
    """The ZOFE commutator [V, X] on rho and on every O[p, s] of device
    states [B, ..., size] (reference layout), V [B|1, N, N] on the device:
    batched matrix products on the GPU."""
    P, S, N, _ = model.oop_shape
    B = states.shape[0]
    lead = states.shape[1:-1]
    x = states.reshape(B, -1, states.shape[-1])
    n2 = N * N
    # column-major blocks: rho = vec^T, O[p, s, i, j] = v[p + P (s + S (i + N j))]
    rho = x[..., :n2].reshape(B, -1, N, N).transpose(-1, -2)
    oop = x[..., n2:].reshape(B, -1, N, N, S, P).permute(0, 1, 5, 4, 3, 2)   # [B, R, P, S, i, j]
    Vb = V.reshape(V.shape[0], 1, N, N)
    crho = torch.matmul(Vb, rho) - torch.matmul(rho, Vb)
    Vo = V.reshape(V.shape[0], 1, 1, 1, N, N)
    coop = torch.matmul(Vo, oop) - torch.matmul(oop, Vo)
    out = torch.cat([crho.transpose(-1, -2).reshape(B, x.shape[1], n2),
                     coop.permute(0, 1, 5, 4, 3, 2).reshape(B, x.shape[1], -1)], dim=-1)
    return out.reshape((B,) + tuple(lead) + (states.shape[-1],))


def _zofe_expectation_rows(V):
    """tr(V rho) as rows on vec(rho): V[i, j] at index j + N i, [B|1, 1, N*N]."""
    return _dev(V.reshape(V.shape[0], 1, -1))


def _zofe_linear_response(model, liouv_space_path, time_max, initial_state, polarization, kw):
    """Linear response of a ZOFE model in the Schroedinger picture with the
    expectation value as save function (response.py:33-40): (t, S [B, ..., nt])."""
    subspaces = liouv_space_path.split('->')
    B = len(_members(model))
    if initial_state is None:
        initial_state = np.stack([model.density_matrix_to_state_vector(
            h.thermal_state(model.hilbert_subspace)) for h in _members(model)])
    else:
        initial_state = np.broadcast_to(np.asarray(initial_state),
                                        (B,) + np.shape(initial_state))
    t = np.arange(0, time_max, model.time_step)
    V0 = _dev(_zofe_ops(model, polarization[0], '+'))
    V1 = _zofe_ops(model, polarization[1], '-')
    init = _dev(np.ascontiguousarray(initial_state))
    lead = init.shape[1:-1]
    y0 = _zofe_commutator(model, init, V0).reshape(B, -1, model.state_size).contiguous()
    eom = model.equation_of_motion(subspaces[1])
    out = propagate_zofe(eom, y0, t, dt=kw.get('dt'), proj=_zofe_expectation_rows(V1),
                         rtol=kw.get('rtol', 1e-6))                      # [B, R, nt, 1]
    return t, (-out[..., 0]).reshape((B,) + tuple(lead) + (len(t),))


def _batched_linear_response(dynamical_model, liouv_space_path, time_max, initial_state=None,
                             polarization='xx', **integrate_kwargs):
    """Linear response of a (possibly batched) model: (t, S [B, ..., nt])."""
    model = dynamical_model
    if hasattr(model, 'oop_shape'):
        return _zofe_linear_response(model, liouv_space_path, time_max, initial_state,
                                     polarization, _propagate_kw(integrate_kwargs))
    kw = _propagate_kw(integrate_kwargs)
    subspaces = liouv_space_path.split('->')
    B = len(_members(model))
    if initial_state is None:
        initial_state = model.thermal_state(subspaces[0])
    init = np.asarray(initial_state)
    t = np.arange(0, time_max, model.time_step)
    signal = ZeroArray()
    for sim_subspace in subspaces[1].split(','):
        V0 = _op_matrices(model, '%s->%s' % (subspaces[0], subspaces[1]), polarization[0], '+',
                          'commutator')                                  # [B|1, M1, M0]
        bra = _op_matrices(model, '%s->%s' % (subspaces[1], subspaces[2]), polarization[1], '-',
                           'bra')                                        # [B|1, M1]
        V_rho0 = np.einsum('bij,...j->b...i', V0, init)                 # [B|1, ..., M1]
        lead = V_rho0.shape[1:-1]
        proj = _dev(V_rho0.reshape(V_rho0.shape[0], -1, V_rho0.shape[-1]))   # [B|1, K, M1]
        eom = model.equation_of_motion(sim_subspace, heisenberg_picture=True)
        if isinstance(eom, DeviceHEOM):
            # the bra and the initial state live on ADO 0 (heom.py:22-58)
            y0 = np.zeros((B, eom.size), dtype=complex)
            y0[:, :bra.shape[-1]] = -bra
            ado0 = propagate_heom(eom, _dev(y0), t, dt=kw.get('dt'),
                                  precision=kw.get('precision', 'fp64'),
                                  rtol=kw.get('rtol', 1e-6))               # [B, nt, M1]
            out = _native.zgemm(ado0.contiguous(), proj, transb=True)    # [B, nt, K]
            s = out.permute(0, 2, 1).reshape((B,) + lead + (len(t),))
        else:
            y0 = _expand(_dev(-bra[:, None, :]), B)                      # [B, 1, M1]
            out = propagate(eom, y0, t, proj=proj, **kw)                 # [B, 1, nt, K]
            s = out[:, 0].permute(0, 2, 1).reshape((B,) + lead + (len(t),))
        signal += s
    return t, signal


_linear_response = optional_ensemble_average(
    optional_2nd_order_isotropic_average(_batched_linear_response))


def linear_response(dynamical_model, liouv_space_path, time_max, initial_state=None,
                    polarization='xx', ensemble_size=None, ensemble_random_orientations=False,
                    exact_isotropic_average=False, **integrate_kwargs):
    """Linear response function S(t) for a path 'ab->cd->ef' (rotating
    wave approximation); returns (t, S)."""
    return _linear_response(dynamical_model, liouv_space_path, time_max, initial_state,
                            polarization, ensemble_size=ensemble_size,
                            ensemble_random_orientations=ensemble_random_orientations,
                            exact_isotropic_average=exact_isotropic_average,
                            **integrate_kwargs)


def absorption_spectra(dynamical_model, time_max, correlation_decay_time=None,
                       polarization='xx', ensemble_size=None, ensemble_random_orientations=False,
                       exact_isotropic_average=False, **integrate_kwargs):
    """Absorption spectrum: Re FT[-S(t)] of the 'gg->eg->gg' response,
    optionally damped by exp(-t/correlation_decay_time); returns (f, X)."""
    t, x = linear_response(dynamical_model, 'gg->eg->gg', time_max, polarization=polarization,
                           ensemble_size=ensemble_size,
                           ensemble_random_orientations=ensemble_random_orientations,
                           exact_isotropic_average=exact_isotropic_average, **integrate_kwargs)
    if correlation_decay_time is not None:
        x = x * np.exp(-t / correlation_decay_time)
    f, X = fourier_transform(t, -x, rw_freq=dynamical_model.rw_freq,
                             unit_convert=dynamical_model.unit_convert)
    return f, X.real


PUMP_PROBE_PATHWAYS = {'GSB': 'gg->eg->gg', 'ESE': 'ee->eg->gg', 'ESA': 'ee->fe->ee'}

def impulsive_probe(dynamical_model, state, time_max, polarization='xx',
                    initial_liouv_subspace='gg,ge,eg,ee', include_signal='GSB,ESE,ESA',
                    ensemble_size=None, ensemble_random_orientations=False,
                    exact_isotropic_average=False, **integrate_kwargs):
    """Probe the second-order part of ``state`` (one or several state
    vectors in ``initial_liouv_subspace``) with an impulsive probe pulse:
    for each selected pump-probe pathway (GSB, ESE, ESA) the part of
    state - thermal state in its first subspace is the initial state of a
    linear response; the summed signal is Fourier transformed.  Returns
    (f, signal).  Several states are right-hand sides of the same launches."""
    model = dynamical_model
    initial_state = np.asarray(state) - model.thermal_state(initial_liouv_subspace)
    total = ZeroArray()
    t = None
    for path in _parse_pathways(PUMP_PROBE_PATHWAYS, include_signal):
        sub0 = path.split('->')[0]
        portion = np.apply_along_axis(
            lambda v: model.map_between_subspaces(v, initial_liouv_subspace, sub0), -1,
            initial_state)
        t, signal = linear_response(model, path, time_max, portion, polarization,
                                    ensemble_size=ensemble_size,
                                    ensemble_random_orientations=ensemble_random_orientations,
                                    exact_isotropic_average=exact_isotropic_average,
                                    **integrate_kwargs)
        total += signal
    return fourier_transform(t, total, rw_freq=model.rw_freq, unit_convert=model.unit_convert)


# Abramavicius, Palmieri, Voronine, Sanda & Mukamel, Chem. Rev. 109, 2350
# (2009), Figs. 4-6 (the pathway table of response.py:252-264).
THIRD_ORDER_PATHWAYS = {
    '-++': {'ESE': 'gg->ge->ee->eg->gg', 'GSB': 'gg->ge->gg->eg->gg',
            'ESA': 'gg->ge->ee->fe->ee'},
    '+-+': {'ESE': 'gg->eg->ee->eg->gg', 'GSB': 'gg->eg->gg->eg->gg',
            'ESA': 'gg->eg->ee->fe->ee'},
    '++-': {'ESA1': 'gg->eg->fg->fe->ee', 'ESA2': 'gg->eg->fg->eg->gg'},
}


def _parse_pathways(possible_pathways, include_signal):
    selected = [v for k, v in possible_pathways.items()
                if include_signal is None or k in include_signal]
    if not selected:
        raise ValueError('at least one Liouville space pathway must be selected, i.e., '
                         'include_signal must include at least one of %r'
                         % list(possible_pathways.keys()))
    return selected


def third_order_grids(model, coherence_time_max, population_time_max=None,
                      population_times=None):
    dt = model.time_step
    t1 = np.arange(0, coherence_time_max, dt)
    t2 = (np.arange(0, population_time_max, dt) if population_times is None
          else np.asarray(population_times, dtype=float))
    t3 = np.arange(0, coherence_time_max, dt)
    return t1, t2, t3


@optional_ensemble_average
@optional_4th_order_isotropic_average
def _third_order_response(dynamical_model, coherence_time_max, population_time_max,
                          population_times, geometry, polarization, include_signal,
                          **integrate_kwargs):
    model = dynamical_model
    kw = _propagate_kw(integrate_kwargs)
    plan = integrate_kwargs.get('plan', 'fused')
    if plan not in ('fused', 'pathways'):
        raise ValueError("plan must be 'fused' or 'pathways'")
    p = None
    if plan == 'fused':
        from .fused import ThirdOrderPlan, PlanNotApplicable, plan_chunk
        args = (model, coherence_time_max, population_time_max, population_times, geometry,
                polarization, include_signal, kw.get('dt'), kw.get('precision', 'fp64'),
                kw.get('schedule', 'auto'))
        try:
            p = ThirdOrderPlan(*args, chunk=plan_chunk(*args))
        except PlanNotApplicable:
            p = None
    if p is None:
        return _pathway_by_pathway(model, coherence_time_max, population_time_max,
                                   population_times, geometry, polarization, include_signal, kw)
    X = p.run()
    p.check_status()
    return (p.t1, p.t2, p.t3), X


# hierarchies with at most this many elements (ADOs x Liouville block) are
# propagated through their dense generator by the Liouville-space engine;
# larger ones by the HEOM stage kernel with hierarchy states as batch entries
HEOM_DENSE_MAX = 512


def _hierarchy_maps(model, mats, what):
    """Per-ADO dipole maps of a hierarchy model as hierarchy matrices:
    commutators kron(I_ado, V) [B|1, n_ado*Mto, n_ado*Mfrom] (ADO-major
    states), bras zero-padded beyond ADO 0 [B|1, n_ado*M]."""
    n = model.ado_count
    if what == 'commutator':
        return np.stack([np.kron(np.eye(n), m) for m in mats])
    out = np.zeros((mats.shape[0], n * mats.shape[-1]), dtype=complex)
    out[:, :mats.shape[-1]] = mats
    return out


def _heom_pathways(model, t1, t2, t3, paths, geometry, polarization, kw):
    """Third-order response of a hierarchy model (response.py:269-337 with
    the HEOM state maps of heom.py:14-58): the whole hierarchy is carried
    through t1 and t2 and mapped by the dipoles ADO by ADO; t3 runs in the
    Heisenberg picture from the bra on ADO 0; the signal contracts over the
    whole hierarchy.  Small hierarchies use the dense generator and the
    Liouville-space engine, large ones the HEOM stage kernel."""
    B = len(_members(model))
    n_ado = model.ado_count
    rho0 = model.thermal_state('gg')
    dt, prec, rtol = kw.get('dt'), kw.get('precision', 'fp64'), kw.get('rtol', 1e-6)
    trans = geometry + '-'
    total = None
    for path in paths:
        s = path.split('->')
        V = [_op_matrices(model, '%s->%s' % (s[i], s[i + 1]), polarization[i], trans[i],
                          'commutator') for i in range(3)]
        b3 = _op_matrices(model, '%s->%s' % (s[3], s[4]), polarization[3], trans[3], 'bra')
        eoms = [model.equation_of_motion(s[1]), model.equation_of_motion(s[2]),
                model.equation_of_motion(s[3], heisenberg_picture=True)]
        if max(e.size for e in eoms) <= HEOM_DENSE_MAX:
            from ..dynamics.liouville_space import DeviceEOM
            dense = [DeviceEOM(e.dense_generator(), e.liouville_subspace, e.heisenberg_picture)
                     for e in eoms]
            Vh = [_hierarchy_maps(model, v, 'commutator') for v in V]
            bh = _hierarchy_maps(model, b3, 'bra')
            V_rho0 = _expand(_dev(np.einsum('bij,j->bi', Vh[0], rho0)[:, None, :]), B)
            V_rho1 = propagate(dense[0], V_rho0, t1, proj=_dev(Vh[1]), **kw)
            V_rho1 = V_rho1.reshape(B, len(t1), -1)
            V_rho2 = propagate(dense[1], V_rho1, t2, t0=0, proj=_dev(Vh[2]), **kw)
            G3 = propagate(dense[2], _expand(_dev(bh[:, None, :]), B), t3, **kw)[:, 0]
        else:
            def per_ado(x, Vm):          # x [B, ..., n_ado*M] -> [B, ..., n_ado*Mto]
                lead = x.shape[1:-1]
                xs = x.reshape(B, -1, Vm.shape[-1])
                return _native.zgemm(xs, _dev(Vm), transb=True).reshape((B,) + lead + (-1,))
            M0 = V[0].shape[-1]
            y0 = torch.zeros((B, eoms[0].size), dtype=torch.complex128, device=_native.device())
            y0[:, :V[0].shape[1]] = _expand(_dev(np.einsum('bij,j->bi', V[0], rho0[:M0])), B)
            st1 = propagate_heom(eoms[0], y0, t1, dt=dt, precision=prec, full_states=True,
                                 rtol=rtol)
            X1 = per_ado(st1, V[1]).reshape(B * len(t1), -1)
            st2 = propagate_heom(eoms[1].repeated(len(t1)), X1, t2, t0=0, dt=dt, precision=prec,
                                 full_states=True, rtol=rtol)
            V_rho2 = per_ado(st2.reshape(B, len(t1), len(t2), -1), V[2])
            g0 = torch.zeros((B, eoms[2].size), dtype=torch.complex128, device=_native.device())
            g0[:, :b3.shape[-1]] = _expand(_dev(b3), B)
            G3 = propagate_heom(eoms[2], g0, t3, dt=dt, precision=prec, full_states=True,
                                rtol=rtol)
        S3 = G3.shape[-1]
        prod = _native.zgemm(V_rho2.reshape(B, -1, S3), G3, transb=True)
        prod = prod.reshape(B, len(t1), len(t2), len(t3))
        total = prod if total is None else total + prod
    return (t1, t2, t3), total


def _zofe_pathways(model, t1, t2, t3, paths, geometry, polarization, kw):
    """Third-order response of a ZOFE model: the Schroedinger picture in
    every interval (no Heisenberg picture, response.py:315-332), the
    commutators applied on the GPU between intervals, all (t1, t2) states
    right-hand sides of the t3 launch with the expectation value saved."""
    B = len(_members(model))
    dt, rtol = kw.get('dt'), kw.get('rtol', 1e-6)
    rho0 = np.stack([model.density_matrix_to_state_vector(
        h.thermal_state(model.hilbert_subspace)) for h in _members(model)])
    trans = geometry + '-'
    size = model.state_size
    total = None
    for path in paths:
        s = path.split('->')
        V = [_dev(_zofe_ops(model, polarization[i], trans[i])) for i in range(3)]
        V3 = _zofe_ops(model, polarization[3], trans[3])
        eoms = [model.equation_of_motion(x) for x in s[1:4]]
        y0 = _zofe_commutator(model, _dev(rho0[:, None, :]), V[0])           # [B, 1, size]
        st1 = propagate_zofe(eoms[0], y0.contiguous(), t1, dt=dt, rtol=rtol)   # [B, 1, n1, size]
        X1 = _zofe_commutator(model, st1[:, 0], V[1]).contiguous()           # [B, n1, size]
        st2 = propagate_zofe(eoms[1], X1, t2, t0=0, dt=dt, rtol=rtol)       # [B, n1, n2, size]
        X2 = _zofe_commutator(model, st2, V[2]).reshape(B, -1, size).contiguous()
        sig = propagate_zofe(eoms[2], X2, t3, dt=dt, proj=_zofe_expectation_rows(V3),
                             rtol=rtol)[..., 0]                              # [B, n1*n2, n3]
        sig = sig.reshape(B, len(t1), len(t2), len(t3))
        total = sig if total is None else total + sig
    return (t1, t2, t3), total


def _pathway_by_pathway(model, coherence_time_max, population_time_max, population_times,
                        geometry, polarization, include_signal, kw):
    """Each pathway evaluated on its own, in the reference's order."""
    t1, t2, t3 = third_order_grids(model, coherence_time_max, population_time_max,
                                   population_times)
    if hasattr(model, 'ado_count'):
        return _heom_pathways(model, t1, t2, t3,
                              _parse_pathways(THIRD_ORDER_PATHWAYS[geometry], include_signal),
                              geometry, polarization, kw)
    if hasattr(model, 'oop_shape'):
        return _zofe_pathways(model, t1, t2, t3,
                              _parse_pathways(THIRD_ORDER_PATHWAYS[geometry], include_signal),
                              geometry, polarization, kw)
    B = len(_members(model))
    rho0 = model.thermal_state('gg')
    total = None
    for path in _parse_pathways(THIRD_ORDER_PATHWAYS[geometry], include_signal):
        s = path.split('->')
        trans = geometry + '-'
        V0 = _op_matrices(model, '%s->%s' % (s[0], s[1]), polarization[0], trans[0], 'commutator')
        V1 = _op_matrices(model, '%s->%s' % (s[1], s[2]), polarization[1], trans[1], 'commutator')
        V2 = _op_matrices(model, '%s->%s' % (s[2], s[3]), polarization[2], trans[2], 'commutator')
        b3 = _op_matrices(model, '%s->%s' % (s[3], s[4]), polarization[3], trans[3], 'bra')
        eom1 = model.equation_of_motion(s[1])
        eom2 = model.equation_of_motion(s[2])
        eom3 = model.equation_of_motion(s[3], heisenberg_picture=True)
        V_rho0 = _expand(_dev(np.einsum('bij,j->bi', V0, rho0)[:, None, :]), B)   # [B,1,M1]
        V_rho1 = propagate(eom1, V_rho0, t1, proj=_dev(V1), **kw)              # [B,1,n1,M2]
        V_rho1 = V_rho1.reshape(B, len(t1), -1)
        V_rho2 = propagate(eom2, V_rho1, t2, t0=0, proj=_dev(V2), **kw)        # [B,n1,n2,M3]
        G3 = propagate(eom3, _expand(_dev(b3[:, None, :]), B), t3, **kw)      # [B,1,n3,M3]
        # signal[b, a, c, d] = sum_i G3[b, d, i] * V_rho2[b, a, c, i]
        M3 = G3.shape[-1]
        prod = _native.zgemm(V_rho2.reshape(B, -1, M3), G3[:, 0], transb=True)   # [B, n1*n2, n3]
        prod = prod.reshape(B, len(t1), len(t2), len(t3))
        total = prod if total is None else total + prod
    return (t1, t2, t3), total


def third_order_response(dynamical_model, coherence_time_max, population_time_max=None,
                         population_times=None, geometry='-++', polarization='xxxx',
                         include_signal=None, ensemble_size=None,
                         ensemble_random_orientations=False, exact_isotropic_average=False,
                         **integrate_kwargs):
    """Third-order response R(t1, t2, t3) summed over Liouville pathways
    (photon echo '-++', non-rephasing '+-+', double quantum '++-');
    returns ((t1, t2, t3), signal[len(t1), len(t2), len(t3)])."""
    return _third_order_response(dynamical_model, coherence_time_max, population_time_max,
                                 population_times, geometry, polarization, include_signal,
                                 ensemble_size=ensemble_size,
                                 ensemble_random_orientations=ensemble_random_orientations,
                                 exact_isotropic_average=exact_isotropic_average,
                                 **integrate_kwargs)


def two_dimensional_spectra(dynamical_model, coherence_time_max, population_time_max=None,
                            population_times=None, geometry='-++', polarization='xxxx',
                            include_signal=None, ensemble_size=None,
                            ensemble_random_orientations=False, exact_isotropic_average=False,
                            **integrate_kwargs):
    """2-D spectrum: the third-order response Fourier transformed along t1
    (sign -1) and t3; returns ((f1, t2, f3), X)."""
    (t1, t2, t3), X = third_order_response(
        dynamical_model, coherence_time_max, population_time_max, population_times, geometry,
        polarization, include_signal, ensemble_size, ensemble_random_orientations,
        exact_isotropic_average, **integrate_kwargs)
    rw, uc = dynamical_model.rw_freq, dynamical_model.unit_convert
    f1, X = fourier_transform(t1, X, 0, rw_freq=rw, sign=-1, unit_convert=uc)
    f3, X = fourier_transform(t3, X, 2, rw_freq=rw, unit_convert=uc)
    return (f1, t2, f3), X
