"""Bath descriptions (qspectra/bath.py).

The propagation path evaluates the Debye correlation function on the GPU
inside the Redfield builder (qsx_redfield_block); the host methods here give
the same function for inspection and tests of the host API.
"""
import numpy as np

from .utils import inspect_repr


class Bath(object):
    def corr_func_real(self, x):
        """Real part of the one-sided correlation function from detailed
        balance: (n(x) + 1) J(x) (T*J'(0) at x = 0)."""
        T = self.temperature
        if x == 0:
            return T * self.spectral_density_limit_at_zero
        J = self.spectral_density_func(x) if x >= 0 else -self.spectral_density_func(-x)
        return (1 / np.expm1(x / T) + 1) * J

    def __repr__(self):
        return inspect_repr(self)

    def spectral_density_func(self, x):
        raise NotImplementedError

    @property
    def spectral_density_limit_at_zero(self):
        raise NotImplementedError


class ArbitraryBath(Bath):
    """A bath given by a spectral density function J(x) (bath.py:50-57); only
    the real correlation function follows from it (detailed balance), so
    Redfield models use it with discard_imag_corr=True.  The host evaluates
    it at each member's eigen-energy differences for the GPU builder.

    The reference's constructor assigns spectral_density_limit_at_zero over
    the base class's read-only property (bath.py:42-47, :55) and so raises
    AttributeError; here the class attribute below shadows that property, so
    the value given is stored."""

    spectral_density_limit_at_zero = None

    def __init__(self, temperature, spectral_density_func, spectral_density_limit_at_zero):
        self.temperature = temperature
        self.spectral_density_func = spectral_density_func
        self.spectral_density_limit_at_zero = spectral_density_limit_at_zero


class UncoupledBath(Bath):
    """No system-bath coupling: C(w) = 0."""
    temperature = 0

    def corr_func_complex(self, _):
        return 0j

    def spectral_density_func(self, _):
        return 0

    @property
    def spectral_density_limit_at_zero(self):
        return 0


class DebyeBath(Bath):
    """Overdamped Brownian oscillator (Drude-Lorentz) bath,
    J(w) = 2 lam gam w / (gam^2 + w^2)."""

    def __init__(self, temperature, reorg_energy, cutoff_freq):
        self.temperature = float(temperature)
        self.reorg_energy = float(reorg_energy)
        self.cutoff_freq = float(cutoff_freq)

    def __eq__(self, other):
        return (isinstance(other, DebyeBath) and self.temperature == other.temperature
                and self.reorg_energy == other.reorg_energy
                and self.cutoff_freq == other.cutoff_freq)

    def __hash__(self):
        return hash((self.temperature, self.reorg_energy, self.cutoff_freq))

    def spectral_density_func(self, x):
        return 2 * self.reorg_energy * self.cutoff_freq * x / (self.cutoff_freq ** 2 + x ** 2)

    @property
    def spectral_density_limit_at_zero(self):
        return 2 * self.reorg_energy / self.cutoff_freq

    def corr_func_complex(self, x, matsubara_cutoff=1000):
        """C(w) = int_0^inf e^{iwt} C(t) dt with the Matsubara series cut
        after ``matsubara_cutoff`` terms (J. Chem. Phys. 112, 7953)."""
        T, lam, g = self.temperature, self.reorg_energy, self.cutoff_freq
        if x == 0:
            return lam * (2 * T / g - 1j)
        nu = 2 * np.pi * T * np.arange(1, matsubara_cutoff)
        tail = np.sum(nu / ((nu ** 2 - g ** 2) * (nu - 1j * x)))
        return lam * g * ((1 / np.tan(g / (2 * T)) - 1j) / (g - 1j * x) + 4 * T * tail)


def is_debye_form(bath):
    """True for the baths the GPU builder evaluates itself (Debye C(w) with
    its Matsubara sum, or no coupling)."""
    return isinstance(bath, (DebyeBath, UncoupledBath))


def debye_parameters(bath):
    """(T, lambda, gamma) for the GPU builder; an uncoupled bath is a Debye
    bath with lambda = 0."""
    if isinstance(bath, DebyeBath):
        return bath.temperature, bath.reorg_energy, bath.cutoff_freq
    if isinstance(bath, UncoupledBath):
        return 1.0, 0.0, 1.0
    raise NotImplementedError('the GPU Redfield builder supports DebyeBath and UncoupledBath; '
                              'got %s' % type(bath).__name__)


class PseudomodeBath(Bath):
    """Bath given by pseudomodes: Lorentzians of the bath correlation
    spectrum centred at Omega with widths gamma and weights huang
    (qspectra/bath.py:105-145); arrays of shape (numb_pm, n_sites) when the
    sites have separate baths.  It defines no temperature (thermal states
    are ground states), and is the bath of the ZOFE model."""

    def __init__(self, numb_pm, Omega, gamma, huang):
        self.numb_pm = numb_pm
        self.Omega = Omega
        self.gamma = gamma
        self.huang = huang
