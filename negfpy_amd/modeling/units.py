"""QE internal frequency units (reference ``src/negfpy/modeling/units.py:9-54``):
omega in sqrt(Ry / (2 m_e bohr^2)) for IFCs in Ry/bohr^2 and masses in QE Ry mass
units (2 m_e)."""

from __future__ import annotations

import numpy as np

RYDBERG_J = 2.1798723611035e-18
BOHR_M = 5.29177210903e-11
ELECTRON_MASS_KG = 9.1093837015e-31
QE_RY_MASS_UNIT_KG = 2.0 * ELECTRON_MASS_KG
SPEED_OF_LIGHT_CM_S = 2.99792458e10
HBAR_J_S = 1.054571817e-34
EV_J = 1.602176634e-19


def qe_omega_to_rad_s(omega):
    return np.asarray(omega) * np.sqrt(RYDBERG_J / (QE_RY_MASS_UNIT_KG * BOHR_M * BOHR_M))


def qe_omega_to_thz(omega):
    return np.asarray(qe_omega_to_rad_s(omega), dtype=float) / (2.0 * np.pi * 1.0e12)


def qe_omega_to_cm1(omega):
    return np.asarray(qe_omega_to_rad_s(omega), dtype=float) / (2.0 * np.pi * SPEED_OF_LIGHT_CM_S)


def qe_omega_to_ev(omega):
    return (HBAR_J_S * np.asarray(qe_omega_to_rad_s(omega), dtype=float)) / EV_J


def qe_ev_to_omega(ev):
    return np.asarray(ev, dtype=float) / float(np.asarray(qe_omega_to_ev(1.0), dtype=float))


def qe_cm1_to_omega(freq_cm1: float) -> float:
    """Inverse of qe_omega_to_cm1 (reference workflows/ifc_bulk.py:113-116)."""
    if freq_cm1 <= 0.0:
        raise ValueError("cm^-1 frequency must be positive.")
    return float(freq_cm1 / float(np.asarray(qe_omega_to_cm1(1.0), dtype=float)))


def qe_thz_to_omega(freq_thz: float) -> float:
    """Inverse of qe_omega_to_thz (reference workflows/ifc_bulk.py:119-122)."""
    if freq_thz <= 0.0:
        raise ValueError("THz frequency must be positive.")
    return float(freq_thz / float(np.asarray(qe_omega_to_thz(1.0), dtype=float)))
