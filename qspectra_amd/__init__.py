"""qspectra_amd: MI355X-native density-matrix propagation for nonlinear
spectroscopy, a drop-in for the propagation path of qspectra.

The Python layer mirrors qspectra's DynamicalModel / simulate API; the
arithmetic runs in hand-written HIP kernels for gfx950 (libqsx.so) reached
through a C ABI (include/qsx.h).  There is no CPU fallback.
"""
from ._native import (NativeUnavailable, NativeError, require_native, load_library,
                      QSX_FP64, QSX_MIXED)

__version__ = '0.1.0'
