"""Thermal conductance G(T) from a transmission spectrum on the GPU
(reference ``src/negfpy/workflows/ifc_bulk.py:408-438`` and
``src/negfpy/modeling/units.py:8-26``).

G(T) = (1/2pi) int k_B T(omega) x^2 e^-x / (1 - e^-x)^2 d omega,
x = hbar omega / (k_B T), with omega converted from QE-q2r units to rad/s;
the kernel integrand is 1 - x^2/12 below x = 1e-6 and 0 from x = 700, and
G = 0 for T <= 0.  One GPU block per temperature, FP64 trapezoid reduction.
"""

from __future__ import annotations

import numpy as np

from . import _engine

RYDBERG_J = 2.1798723611035e-18
BOHR_M = 5.29177210903e-11
ELECTRON_MASS_KG = 9.1093837015e-31
QE_RY_MASS_UNIT_KG = 2.0 * ELECTRON_MASS_KG


def qe_omega_to_rad_s(omega):
    """units.py:18-26: sqrt(Ry / (2 m_e a0^2)) per internal QE frequency unit."""
    return np.asarray(omega) * np.sqrt(RYDBERG_J / (QE_RY_MASS_UNIT_KG * BOHR_M * BOHR_M))


def thermal_conductance_from_spectrum(omega_qe, transmission_vals, temperatures_k) -> np.ndarray:
    """Thermal conductance in W/K per temperature (ifc_bulk.py:408-438)."""
    omega_qe = np.asarray(omega_qe, dtype=float)
    tvals = np.asarray(transmission_vals, dtype=float)
    temps = np.asarray(temperatures_k, dtype=float)
    if omega_qe.ndim != 1 or tvals.ndim != 1 or omega_qe.size != tvals.size:
        raise ValueError("omega_qe and transmission_vals must be 1D arrays of equal length.")
    if np.any(np.diff(omega_qe) <= 0.0):
        raise ValueError("omega_qe must be strictly increasing.")
    return thermal_conductance_rad_s(qe_omega_to_rad_s(omega_qe), tvals, temps)


def thermal_conductance_rad_s(omega_rad_s, transmission_vals, temperatures_k) -> np.ndarray:
    """Same integral with omega already in rad/s."""
    w = np.asarray(omega_rad_s, dtype=float)
    t = np.asarray(transmission_vals, dtype=float)
    temps = np.atleast_1d(np.asarray(temperatures_k, dtype=float))
    if w.size < 2:
        return np.zeros_like(temps)
    return _engine.thermal_conductance(w, t, temps)
