"""Classical pulses in the rotating frame (qspectra/pulse.py:7-114).

A pulse is a callable E(t, rw_freq) returning its complex field envelope at
the rotating-wave frequency rw_freq, with attributes ``t_init`` and
``t_final`` bounding the time it acts.  Field-driven propagation
(simulate_with_fields, simulate_pump) samples every pulse on the host at the
three stage times of every RK4 step and hands the samples to the driven
kernel, so any callable that accepts numpy arrays is a valid pulse.
"""
from abc import ABCMeta, abstractmethod

import numpy as np

from .constants import GAUSSIAN_SD_FWHM
from .utils import inspect_repr


class Pulse(metaclass=ABCMeta):
    """Pulse interface (qspectra/pulse.py:7-43): ``pulse(t, rw_freq)`` is the
    complex electric field at time(s) t; ``t_init`` / ``t_final`` bound it."""

    def __repr__(self):
        return inspect_repr(self)

    @abstractmethod
    def __call__(self, t, rw_freq):
        """Complex field at time(s) t in the frame rotating at rw_freq."""


class CustomPulse(Pulse):
    """Field given by a function call(t, rw_freq) (qspectra/pulse.py:51-71)."""

    def __init__(self, t_init, t_final, call):
        self.t_init = t_init
        self.t_final = t_final
        self.call = call

    def __call__(self, t, rw_freq):
        return self.call(t, rw_freq)


class GaussianPulse(Pulse):
    """Gaussian envelope at a carrier frequency (qspectra/pulse.py:74-114):

        E(t) = scale * exp(i c (carrier_freq - rw_freq) (t - t_peak)
                           - (t - t_peak)^2 / (2 sigma^2)),

    sigma = GAUSSIAN_SD_FWHM * fwhm, c = freq_convert (frequency to angular
    inverse time units, e.g. CM_FS); the pulse acts over t_peak -/+
    t_limits_multiple * sigma."""

    def __init__(self, carrier_freq, fwhm, t_peak=0, scale=1, freq_convert=1,
                 t_limits_multiple=3):
        sigma = GAUSSIAN_SD_FWHM * fwhm
        self.sigma = sigma
        self.two_sigma_squared = 2 * sigma ** 2
        self.t_init = t_peak - t_limits_multiple * sigma
        self.t_final = t_peak + t_limits_multiple * sigma
        self.t_peak = t_peak
        self.carrier_freq = carrier_freq
        self.scale = scale
        self.freq_convert = freq_convert

    def __call__(self, t, rw_freq):
        dt = np.asarray(t) - self.t_peak
        phase = 1j * self.freq_convert * (self.carrier_freq - rw_freq) * dt
        return self.scale * np.exp(phase - dt ** 2 / self.two_sigma_squared)


def sample_field(pulse, times, rw_freq):
    """pulse evaluated at every entry of ``times`` (a float64 array), as a
    complex128 array of the same shape; a pulse whose function does not
    accept arrays is evaluated one time at a time."""
    times = np.asarray(times, dtype=float)
    try:
        E = np.asarray(pulse(times, rw_freq), dtype=complex)
    except (TypeError, ValueError):
        E = None
    if E is None or E.shape != times.shape:
        E = np.array([complex(pulse(float(x), rw_freq)) for x in times.reshape(-1)],
                     dtype=complex).reshape(times.shape)
    return E


def field_rate(pulse, rw_freq, n=2049):
    """(max |E|, max |dE/dt| / max |E|) of a pulse over [t_init, t_final],
    from n samples: the field amplitude and the fastest relative rate of
    change (carrier detuning plus envelope), which bound the RK4 step."""
    t0, t1 = float(pulse.t_init), float(pulse.t_final)
    if not t1 > t0:
        E = sample_field(pulse, np.array([t0]), rw_freq)
        return float(np.abs(E).max()), 0.0
    ts = np.linspace(t0, t1, n)
    E = sample_field(pulse, ts, rw_freq)
    amp = float(np.abs(E).max())
    if amp == 0:
        return 0.0, 0.0
    rate = float(np.abs(np.diff(E)).max() / (ts[1] - ts[0])) / amp
    return amp, rate
