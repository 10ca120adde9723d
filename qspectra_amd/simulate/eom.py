"""Field-free propagation of density matrices (qspectra/simulate/eom.py,
simulate_dynamics)."""
import numpy as np

from .. import _native
from .decorators import optional_ensemble_average
from .utils import propagate, propagate_heom
from ..dynamics.liouville_space import LinearMap
from ..dynamics.heom import DeviceHEOM


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
    if isinstance(eom, DeviceHEOM):
        # hierarchy: the reduced density matrix is ADO 0 of the state
        B = eom.batch
        y = _native.to_device(np.broadcast_to(y0, (B, y0.shape[-1])).copy())
        ado0 = propagate_heom(eom, y, t, dt=kw.get('dt'),
                              precision=kw.get('precision', 'fp64')).cpu().numpy()
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
