"""qspectra_amd: MI355X-native density-matrix propagation for nonlinear
spectroscopy, a drop-in for the propagation path of qspectra.

The Python layer mirrors qspectra's DynamicalModel / simulate API; the
arithmetic runs in hand-written HIP kernels for gfx950 (libqsx.so) reached
through a C ABI (include/qsx.h).  There is no CPU fallback.
"""
from ._native import (NativeUnavailable, NativeError, require_native, load_library,
                      QSX_FP64, QSX_MIXED)

__version__ = '0.1.0'
from .bath import DebyeBath, ArbitraryBath, UncoupledBath, PseudomodeBath
from .constants import CM_FS, CM_K, GAUSSIAN_SD_FWHM
from .dynamics.liouville_space import (n_excitations, matrix_to_ket_vec, ket_vec_to_matrix,
                                       matrix_to_bra_vec)
from .dynamics.redfield import RedfieldModel
from .dynamics.unitary import UnitaryModel
from .dynamics.heom import HEOMModel
from .dynamics.zofe import ZOFEModel
from .hamiltonian import Hamiltonian, ElectronicHamiltonian, VibronicHamiltonian
from .operator_tools import (unit_vec, basis_transform_operator, basis_transform_vector,
                             all_states)
from .polarization import (polarization_vector, check_polarizations,
                           invariant_weights_4th_order, invariant_polarizations,
                           FOURTH_ORDER_INVARIANTS, MAGIC_ANGLE)
from .pulse import Pulse, CustomPulse, GaussianPulse
from .simulate.eom import simulate_dynamics, simulate_with_fields, simulate_pump
from .simulate.response import (linear_response, absorption_spectra, third_order_response,
                                two_dimensional_spectra, impulsive_probe, PUMP_PROBE_PATHWAYS,
                                THIRD_ORDER_PATHWAYS)
from .simulate.utils import fourier_transform, integrate, bound_signal, IntegratorError

__all__ = ['DebyeBath', 'ArbitraryBath', 'UncoupledBath', 'PseudomodeBath', 'ZOFEModel', 'CM_FS', 'CM_K', 'GAUSSIAN_SD_FWHM',
           'Hamiltonian', 'ElectronicHamiltonian', 'VibronicHamiltonian', 'n_excitations', 'matrix_to_ket_vec',
           'ket_vec_to_matrix', 'matrix_to_bra_vec', 'unit_vec', 'basis_transform_operator',
           'basis_transform_vector', 'all_states', 'polarization_vector',
           'check_polarizations', 'invariant_weights_4th_order', 'invariant_polarizations',
           'FOURTH_ORDER_INVARIANTS', 'MAGIC_ANGLE', 'RedfieldModel', 'UnitaryModel', 'HEOMModel',
           'simulate_dynamics', 'simulate_with_fields', 'simulate_pump', 'Pulse', 'CustomPulse',
           'GaussianPulse', 'linear_response', 'absorption_spectra',
           'third_order_response', 'two_dimensional_spectra', 'impulsive_probe',
           'PUMP_PROBE_PATHWAYS',
           'THIRD_ORDER_PATHWAYS', 'fourier_transform', 'integrate', 'bound_signal',
           'IntegratorError', 'NativeUnavailable', 'NativeError', 'require_native']
