"""Physical constants and temperature helpers (pibronic/constants.py:12-50).

k_B is built from the same two NIST values as the reference so that beta(T) is
bit-identical: beta(300 K) = 38.68174020133669.
"""
import numpy as np

# joules per eV
nist_j_per_ev = np.float64(1.6021766208e-19)
# eV per kelvin (pibronic/constants.py:24)
boltzman = np.float64(1.38064852e-23) / nist_j_per_ev
hbar = 1.0
# finite-difference step in beta for E and Cv (pibronic/constants.py:28)
delta_beta = 2.0e-4
wavenumber_per_eV = 8065.6


def beta(temperature):
    """1 / (k_B T) in 1/eV (pibronic/constants.py:41-44)."""
    return 1. / (temperature * boltzman)


def extract_T_from_beta(beta):
    """Temperature in kelvin for a given beta (pibronic/constants.py:47-50)."""
    return 1. / (beta * boltzman)
