"""Oracle: time integration, Fourier transform and response-function drivers.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

* ``integrate_zvode`` restates qspectra/simulate/utils.py:16-109 literally
  (scipy ``ode('zvode')`` per output time, serial loop over rows of a 2-D y0).
  It is the "reference CPU path" of the benchmark.
* ``integrate_rk4`` is the same driver with a fixed-step classical RK4 in
  place of zvode (the product's discretisation), vectorised over leading axes.
* ``integrate_rk4_driven`` is classical RK4 for the field-driven equation
  y' = L y + sum_i f_i(t) C_i y on the same schedule (stage times t,
  t + h/2, t + h of every step), the discretisation of the driven kernel.
* the response drivers restate qspectra/simulate/response.py and the
  decorators of qspectra/simulate/decorators.py; simulate_with_fields and
  simulate_pump restate qspectra/simulate/eom.py:79-227.
"""
import numpy as np
import scipy.integrate
from scipy.fftpack import fft, fftfreq, fftshift, ifftshift

from .system import (check_polarizations, invariant_weights_4th_order,
                     invariant_polarizations, FOURTH_ORDER_INVARIANTS)


class IntegratorError(Exception):
    pass


# ------------------------------------------------------------- integrators
def integrate_zvode(f, y0, t, t0=None, save_func=None, **kwargs):
    """simulate/utils.py:16-50 and :103-109 (rows of 2-D y0 looped serially)."""
    y0 = np.asarray(y0)
    if y0.ndim > 1:
        out = [integrate_zvode(f, yi, t, t0, save_func, **kwargs) for yi in y0]
        return np.array(out)
    if t0 is None:
        t0 = t[0]
    solver = scipy.integrate.ode(f)
    solver.set_integrator('zvode', **kwargs)
    solver.set_initial_value(y0, t0)
    save = save_func if save_func is not None else (lambda x: x)
    ys = save(y0)
    y = np.empty([len(t)] + list(np.shape(ys)), dtype=np.asarray(ys).dtype)
    i0 = 0
    if t[0] == t0:
        y[0] = ys
        i0 = 1
    for i in range(i0, len(t)):
        if not solver.successful():
            raise IntegratorError('integration failed at time {}'.format(t[i]))
        y[i] = save(solver.integrate(t[i]))
    return y


def rk4_schedule(t, t0=None, dt=None):
    """Fixed-step schedule used by the RK4 oracle.

    Interval i runs from the previous output time (t0 for the first) to
    t[i]; it is split into n_i equal steps of length h_i <= dt (up to a
    1e-6 relative tolerance).  dt defaults to the smallest nonzero interval.
    """
    t = np.asarray(t, float)
    if t0 is None:
        t0 = t[0]
    prev = np.concatenate([[t0], t[:-1]])
    dts = t - prev
    if dt is None:
        pos = dts[dts > 0]
        dt = pos.min() if pos.size else 1.0
    k = dts / dt
    kr = np.round(k)
    n = np.where(np.abs(k - kr) <= 1e-6 * np.maximum(1.0, k), kr, np.ceil(k)).astype(int)
    n = np.where(dts == 0, 0, np.maximum(n, 1))
    h = np.where(n > 0, dts / np.maximum(n, 1), 0.0)
    return n, h


def rk4_step(L, y, h):
    """Classical RK4 on y' = L y; y has shape [..., M] (L applied on the
    last axis, L may be dense ndarray or scipy sparse)."""
    def f(v):
        return (L @ v.reshape(-1, v.shape[-1]).T).T.reshape(v.shape)
    k1 = f(y)
    k2 = f(y + (h / 2) * k1)
    k3 = f(y + (h / 2) * k2)
    k4 = f(y + h * k3)
    return y + (h / 6) * (k1 + 2 * k2 + 2 * k3 + k4)


def integrate_rk4(L, y0, t, t0=None, save_matrix=None, dt=None):
    """Fixed-step RK4 counterpart of ``integrate``.

    Output layout follows simulate/utils.py:103-109: for y0 of shape
    [..., M] the result is [..., len(t), K] (rows first), with K = M or the
    row count of ``save_matrix`` (the matrix of a linear save_func).
    """
    y = np.array(y0, dtype=complex)
    n, h = rk4_schedule(t, t0, dt)
    out = []
    for ni, hi in zip(n, h):
        for _ in range(ni):
            y = rk4_step(L, y, hi)
        out.append(y if save_matrix is None else y @ np.asarray(save_matrix).T)
    return np.stack(out, axis=-2)


def integrate_rk4_f(f, y0, t, t0=None, dt=None, save_func=None):
    """Classical RK4 on y' = f(t, y) (any right-hand side, e.g. the
    nonlinear ZOFE equation) on the same schedule as integrate_rk4; rows
    of a multi-dimensional y0 are integrated one by one, as
    simulate/utils.py:103-109 does.  Returns [..., len(t), *save_shape]."""
    y0 = np.asarray(y0)
    if y0.ndim > 1:
        return np.array([integrate_rk4_f(f, yi, t, t0, dt, save_func) for yi in y0])
    t = np.asarray(t, float)
    if t0 is None:
        t0 = t[0]
    n, h = rk4_schedule(t, t0, dt)
    starts = np.concatenate([[t0], t[:-1]])
    save = save_func if save_func is not None else (lambda v: v)
    y = np.array(y0, dtype=complex)
    out = []
    for s, (ns, hs) in enumerate(zip(n, h)):
        for k in range(ns):
            tk = starts[s] + k * hs
            k1 = f(tk, y)
            k2 = f(tk + hs / 2, y + (hs / 2) * k1)
            k3 = f(tk + hs / 2, y + (hs / 2) * k2)
            k4 = f(tk + hs, y + hs * k3)
            y = y + (hs / 6) * (k1 + 2 * k2 + 2 * k3 + k4)
        out.append(save(y))
    return np.array(out)


def _integrate_f(kind, dt, zvode_kw):
    """integrate(f, y0, t, t0, save_func) for a right-hand-side function:
    zvode (simulate/utils.py:16-109, rows looped) or the fixed-step RK4."""
    def run(f, y0, t, t0=None, save_func=None):
        if kind == 'rk4':
            return integrate_rk4_f(f, y0, t, t0, dt, save_func)
        return integrate_zvode(f, y0, t, t0, save_func, **zvode_kw)
    return run


def integrate_rk4_driven(L, fields, y0, t, t0=None, dt=None):
    """RK4 on y' = L y + sum_i f_i(t) (C_i y), fields = [(f_i, C_i)] with
    f_i a function of time; step k of interval s starts at t_{s-1} + k h_s
    (t0 for s = 0).  Returns [len(t), M]."""
    t = np.asarray(t, float)
    if t0 is None:
        t0 = t[0]
    n, h = rk4_schedule(t, t0, dt)
    starts = np.concatenate([[t0], t[:-1]])

    def f(tt, v):
        d = L.dot(v)
        for fi, Ci in fields:
            d = d + fi(tt) * Ci.dot(v)
        return d
    y = np.array(y0, dtype=complex)
    out = []
    for s, (ns, hs) in enumerate(zip(n, h)):
        for k in range(ns):
            tk = starts[s] + k * hs
            k1 = f(tk, y)
            k2 = f(tk + hs / 2, y + (hs / 2) * k1)
            k3 = f(tk + hs / 2, y + (hs / 2) * k2)
            k4 = f(tk + hs, y + hs * k3)
            y = y + (hs / 6) * (k1 + 2 * k2 + 2 * k3 + k4)
        out.append(y)
    return np.array(out)


# ------------------------------------------------------------ Fourier tools
def is_constant(x, atol=1e-7, positive=None):
    """simulate/utils.py:122-125."""
    x = np.asarray(x)
    return (np.max(np.abs(x - x[0])) < atol
            and (np.all((x > 0) == positive) if positive is not None else True))


def symmetrize(t, x, axis=-1):
    """simulate/utils.py:128-151."""
    if not is_constant(np.diff(t), positive=True):
        raise ValueError('sample times must differ by a positive constant')
    t = np.asarray(t)
    x = np.asarray(x)
    T = max(t[-1], -t[0])
    dt = t[1] - t[0]
    n_plus = int((T - t[-1]) / dt) + 1
    n_minus = int((T + t[0]) / dt) + 1
    t_sym = np.concatenate([t[0] - dt * np.arange(1, n_minus)[::-1], t,
                            t[-1] + dt * np.arange(1, n_plus)])
    ax = axis % x.ndim
    shape = list(x.shape)
    shape[ax] = t_sym.size
    x_sym = np.zeros(shape, dtype=x.dtype)
    start, end = (np.searchsorted(t_sym, ti) for ti in (t[0], t[-1]))
    sl = [slice(None)] * x.ndim
    sl[ax] = slice(start, end + 1)
    x_sym[tuple(sl)] = x
    return t_sym, x_sym


def fourier_transform(t, x, axis=-1, rw_freq=0, unit_convert=1, sign=1,
                      convention='angular'):
    """simulate/utils.py:154-219."""
    if convention == 'angular':
        unit_convert = unit_convert / (2 * np.pi)
    t, x = symmetrize(t, x, axis)
    N = x.shape[axis]
    dt = t[1] - t[0]
    f = fftshift(fftfreq(N, dt * unit_convert))
    X = fftshift(fft(ifftshift(x * dt, axes=axis), axis=axis), axes=axis)
    if sign == 1:
        f = -f[::-1]
        sl = [slice(None)] * X.ndim
        sl[axis] = slice(None, None, -1)
        X = X[tuple(sl)]
    return f + rw_freq, X


# --------------------------------------------------------- response drivers
def _integrator(kind, dt, zvode_kw):
    """Returns integrate(L, y0, t, t0, save_matrix) for 'rk4' or 'zvode'."""
    def run(L, y0, t, t0=None, save_matrix=None):
        if kind == 'rk4':
            return integrate_rk4(L, y0, t, t0, save_matrix, dt)
        save = None if save_matrix is None else (lambda y: np.asarray(save_matrix).dot(y))
        return integrate_zvode(lambda tt, y: L.dot(y), y0, t, t0, save, **zvode_kw)
    return run


def linear_response(model, path, time_max, polarization='xx', initial_state=None,
                    integrator='rk4', dt=None, **zvode_kw):
    """response.py:15-43 (Heisenberg branch; the Liouville models implement it)."""
    run = _integrator(integrator, dt, zvode_kw)
    subs = path.split('->')
    if initial_state is None:
        initial_state = model.thermal_state(subs[0])
    t = np.arange(0, time_max, model.time_step)
    signal = 0
    for sim in subs[1].split(','):
        V = [model.dipole_operator('%s->%s' % (a, b), p, tr)
             for a, b, p, tr in zip(subs[:-1], subs[1:], polarization, '+-')]
        V_rho0 = np.apply_along_axis(V[0].commutator, -1, initial_state)
        if not hasattr(model, 'evolve_matrix'):
            # no Heisenberg picture (ZOFE): Schroedinger picture with the
            # expectation value as save function (response.py:33-40)
            runf = _integrate_f(integrator, dt, zvode_kw)
            signal = signal - runf(model.equation_of_motion(sim), V_rho0, t,
                                   save_func=V[1].expectation_value)
            continue
        L = model.evolve_matrix(sim, heisenberg_picture=True)
        V_Gt = run(L, -V[1].bra_vector, t)
        signal = signal + np.tensordot(V_rho0, V_Gt, (-1, -1))
    return t, signal


def absorption_spectra(model, time_max, correlation_decay_time=None,
                       polarization='xx', **kw):
    """response.py:103-154 (returns Re FT of -S)."""
    t, x = linear_response(model, 'gg->eg->gg', time_max, polarization, **kw)
    if correlation_decay_time is not None:
        x = x * np.exp(-t / correlation_decay_time)
    f, X = fourier_transform(t, -x, rw_freq=model.rw_freq, unit_convert=model.unit_convert)
    return f, X.real


THIRD_ORDER_PATHWAYS = {
    '-++': {'ESE': 'gg->ge->ee->eg->gg', 'GSB': 'gg->ge->gg->eg->gg',
            'ESA': 'gg->ge->ee->fe->ee'},
    '+-+': {'ESE': 'gg->eg->ee->eg->gg', 'GSB': 'gg->eg->gg->eg->gg',
            'ESA': 'gg->eg->ee->fe->ee'},
    '++-': {'ESA1': 'gg->eg->fg->fe->ee', 'ESA2': 'gg->eg->fg->eg->gg'},
}
"""response.py:252-264 (Abramavicius et al., Chem. Rev. 109, 2350)."""


PUMP_PROBE_PATHWAYS = {'GSB': 'gg->eg->gg', 'ESE': 'ee->eg->gg', 'ESA': 'ee->fe->ee'}


def impulsive_probe(model, state, time_max, polarization='xx',
                    initial_liouv_subspace='gg,ge,eg,ee', include_signal='GSB,ESE,ESA', **kw):
    """response.py:157-231: the second-order part of ``state`` (state minus
    the thermal state), mapped to the first subspace of each pump-probe
    pathway, as the initial state of a linear response; summed over the
    pathways and Fourier transformed."""
    initial_state = np.asarray(state) - model.thermal_state(initial_liouv_subspace)
    total = 0
    t = None
    for path in parse_pathways(PUMP_PROBE_PATHWAYS, include_signal):
        sub0 = path.split('->')[0]
        portion = np.apply_along_axis(
            lambda v: model.map_between_subspaces(v, initial_liouv_subspace, sub0), -1,
            initial_state)
        t, sig = linear_response(model, path, time_max, polarization, portion, **kw)
        total = total + sig
    return fourier_transform(t, total, rw_freq=model.rw_freq, unit_convert=model.unit_convert)


def parse_pathways(possible, include_signal):
    """response.py:162-171 (dict order of the table)."""
    sel = [v for k, v in possible.items() if include_signal is None or k in include_signal]
    if not sel:
        raise ValueError('at least one Liouville space pathway must be selected')
    return sel


def third_order_response(model, coherence_time_max, population_time_max=None,
                         population_times=None, geometry='-++', polarization='xxxx',
                         include_signal=None, integrator='rk4', dt=None, **zvode_kw):
    """response.py:269-337 (Schroedinger t1,t2 + Heisenberg t3)."""
    run = _integrator(integrator, dt, zvode_kw)
    runf = _integrate_f(integrator, dt, zvode_kw)
    grid = model.time_step
    t1 = np.arange(0, coherence_time_max, grid)
    t2 = (np.arange(0, population_time_max, grid) if population_times is None
          else np.asarray(population_times))
    t3 = np.arange(0, coherence_time_max, grid)
    rho0 = model.thermal_state('gg')
    total = 0
    for path in parse_pathways(THIRD_ORDER_PATHWAYS[geometry], include_signal):
        subs = path.split('->')
        V = [model.dipole_operator('%s->%s' % (a, b), p, tr)
             for a, b, p, tr in zip(subs[:-1], subs[1:], polarization, geometry + '-')]
        if not hasattr(model, 'evolve_matrix'):
            # Schroedinger picture throughout (response.py:315-332 fallback)
            eom = [model.equation_of_motion(sub) for sub in subs[1:-1]]
            V_rho1 = runf(eom[0], V[0].commutator(rho0), t1, save_func=V[1].commutator)
            V_rho2 = runf(eom[1], V_rho1, t2, t0=0, save_func=V[2].commutator)
            total = total + runf(eom[2], V_rho2, t3, save_func=V[3].expectation_value)
            continue
        L = [model.evolve_matrix(s) for s in subs[1:-1]]
        V_rho0 = V[0].commutator(rho0)
        V_rho1 = run(L[0], V_rho0, t1, save_matrix=V[1].commutator_matrix())
        V_rho2 = run(L[1], V_rho1, t2, t0=0, save_matrix=V[2].commutator_matrix())
        Lh = model.evolve_matrix(subs[3], heisenberg_picture=True)
        V_Gt3 = run(Lh, V[3].bra_vector, t3)
        total = total + np.einsum('ci,abi', V_Gt3, V_rho2)
    return (t1, t2, t3), total


def two_dimensional_spectra(model, coherence_time_max, population_time_max=None,
                            population_times=None, geometry='-++', polarization='xxxx',
                            include_signal=None, **kw):
    """response.py:430-455."""
    (t1, t2, t3), X = third_order_response(model, coherence_time_max, population_time_max,
                                           population_times, geometry, polarization,
                                           include_signal, **kw)
    f1, X = fourier_transform(t1, X, 0, rw_freq=model.rw_freq, sign=-1,
                              unit_convert=model.unit_convert)
    f3, X = fourier_transform(t3, X, 2, rw_freq=model.rw_freq,
                              unit_convert=model.unit_convert)
    return (f1, t2, f3), X


def simulate_dynamics(model, initial_state, duration=None, times=None,
                      liouville_subspace='ee', integrator='rk4', dt=None, **zvode_kw):
    """simulate/eom.py:11-26 (psi -> outer(conj psi, psi), then reshape 'F')."""
    rho = np.asarray(initial_state)
    if rho.ndim == 1:
        rho = np.outer(rho.conj(), rho)
    y0 = model.density_matrix_to_state_vector(rho, liouville_subspace)
    t = np.arange(0, duration, model.time_step) if times is None else times
    if not hasattr(model, 'evolve_matrix'):
        runf = _integrate_f(integrator, dt, zvode_kw)
        states = runf(model.equation_of_motion(liouville_subspace), y0, t)
        return t, model.state_vector_to_density_matrix(states)
    run = _integrator(integrator, dt, zvode_kw)
    states = run(model.evolve_matrix(liouville_subspace), y0, t)
    N = int(np.sqrt(states.shape[-1]))
    return t, states.reshape(-1, N, N, order='F')


def simulate_with_fields(model, pulses, geometry='-+', polarization='xx', time_extra=0,
                         times=None, liouville_subspace='gg,ge,eg,ee', save_func=None,
                         integrator='rk4', dt=None, **zvode_kw):
    """simulate/eom.py:79-103: thermal state at t0 = min t_init, derivative
    L y + sum (-1j E (conj for '+'))[V, y], outputs on arange(t0, tf +
    time_extra, time_step) or tf + times; states (or save_func of them)."""
    V = [model.dipole_operator(liouville_subspace, p, tr)
         for p, tr in zip(polarization, geometry)]
    info = list(zip(pulses, geometry, V))

    def coef(pulse, trans):
        def fi(tt):
            E = pulse(tt, model.rw_freq)
            return -1j * (np.conj(E) if trans == '+' else E)
        return fi
    rho0 = model.thermal_state(liouville_subspace)
    t0 = min(p.t_init for p in pulses)
    tf = max(p.t_final for p in pulses)
    t = (np.arange(t0, tf + time_extra, model.time_step) if times is None
         else tf + np.asarray(times))
    if not hasattr(model, 'evolve_matrix'):
        # a right-hand-side function (ZOFE): eom.py:86-93 literally
        eom = model.equation_of_motion(liouville_subspace)
        terms = [(coef(p, tr), Vi) for p, tr, Vi in info]

        def f_fields(tt, y):
            d = eom(tt, y)
            for fi, Vi in terms:
                d = d + fi(tt) * Vi.commutator(y)
            return d
        runf = _integrate_f(integrator, dt, zvode_kw)
        return t, runf(f_fields, rho0, t, t0, save_func)
    L = model.evolve_matrix(liouville_subspace)
    fields = [(coef(p, tr), Vi.commutator_matrix()) for p, tr, Vi in info]
    if integrator == 'rk4':
        states = integrate_rk4_driven(L, fields, rho0, t, t0, dt)
        if save_func is not None:
            states = np.array([save_func(v) for v in states])
        return t, states

    def rhs(tt, y):
        d = L.dot(y)
        for fi, Ci in fields:
            d = d + fi(tt) * Ci.dot(y)
        return d
    return t, integrate_zvode(rhs, rho0, t, t0, save_func, **zvode_kw)


def simulate_pump(model, pump, polarization='x', time_extra=0, times=None,
                  liouville_subspace='gg,ge,eg,ee', save_func=None,
                  exact_isotropic_average=False, **kw):
    """simulate/eom.py:162-227: [pump, pump], geometry '-+', the same
    polarization twice; optional 2nd-order isotropic average."""
    def run(polarization):
        return simulate_with_fields(model, [pump, pump], '-+', polarization, time_extra,
                                    times, liouville_subspace, save_func, **kw)
    pols = [polarization, polarization]
    if exact_isotropic_average:
        return isotropic_average_2nd(run, pols)
    return run(pols)


# ------------------------------------------------------------------ decorators
def ensemble_average(func, model, ensemble_size, *args, **kwargs):
    """decorators.py:40-64: serial mean over model.sample(n), n < size."""
    total = 0
    for n in range(ensemble_size):
        ticks, sig = func(model.sample(n), *args, **kwargs)
        total = total + sig
    return ticks, total / ensemble_size


def isotropic_average_2nd(func, polarization, **kwargs):
    """decorators.py:99-125."""
    p = check_polarizations(polarization, 2)
    w = np.dot(*p)
    total = 0
    for pol in ['xx', 'yy', 'zz']:
        t, s = func(polarization=pol, **kwargs)
        total = total + s
    return t, total * (w / 3.0)


def isotropic_average_4th(func, polarization, **kwargs):
    """decorators.py:67-96 (weights <= 1e-8 skipped, signals cached)."""
    weights = invariant_weights_4th_order(polarization)
    cache = {}
    total = 0
    t = None
    for inv, w in zip(FOURTH_ORDER_INVARIANTS, weights):
        if w > 1e-8:
            for p in invariant_polarizations(inv):
                if p not in cache:
                    t, cache[p] = func(polarization=p, **kwargs)
                total = total + w * cache[p]
    return t, total
