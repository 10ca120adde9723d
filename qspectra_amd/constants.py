"""Unit conversions (same values as qspectra/constants.py:8-18).

Energies are in angular cm^-1 and times in fs: multiplying an energy by
CM_FS = 2*pi*c (c in cm/fs) gives rad/fs.
"""
import math

CM_FS_LINEAR = 2.99792458e-5               # speed of light in cm/fs
CM_FS = 2 * math.pi * CM_FS_LINEAR          # cm^-1 -> rad/fs
CM_K = 0.69503476                           # Boltzmann constant in cm^-1/K
GAUSSIAN_SD_FWHM = 1.0 / (2 * math.sqrt(2 * math.log(2)))  # sigma / FWHM
