"""Full counting statistics of coherent phonon heat transport
(reference ``src/negfpy/core/fcs.py``).

The cumulant integrals run on the GPU (``k_fcs_cumulants``, C entry
``negf_fcs_cumulants``), one workgroup per temperature pair, so a whole bias or
temperature scan is one launch (``heat_current_cumulants_scan``).  The drop-in
functions keep the reference's names, argument checks, units and result types.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Iterable

import numpy as np
import torch

from negfpy_amd import _lib

from . import _engine

Array = np.ndarray

__all__ = [
    "HeatCurrentCumulants",
    "HeatCurrentUncertainty",
    "heat_current_cumulants_from_spectrum",
    "heat_current_cumulants_from_k_moments",
    "heat_current_cumulants_scan",
    "heat_current_uncertainty",
]


@dataclass(frozen=True)
class HeatCurrentCumulants:
    """First and second cumulants of heat current (per unit time), fcs.py:18-23."""

    c1_j_per_s: float
    c2_j2_per_s: float


@dataclass(frozen=True)
class HeatCurrentUncertainty:
    """Finite-time uncertainty derived from the second cumulant, fcs.py:26-33."""

    measurement_time_s: float
    std_current_j_per_s: float
    variance_energy_j2: float
    std_energy_j: float


def _grid(values: Iterable[float], name: str) -> Array:
    """fcs.py:36-44."""
    arr = np.asarray(list(values), dtype=float)
    if arr.ndim != 1:
        raise ValueError(f"{name} must be 1D.")
    if arr.size < 2:
        raise ValueError(f"{name} must contain at least two points.")
    if np.any(np.diff(arr) <= 0.0):
        raise ValueError(f"{name} must be strictly increasing.")
    return arr


def _run(omega: Array, t1: Array, t2: Array | None, temps_left, temps_right) -> Array:
    lib = _lib.load()
    tl = np.atleast_1d(np.asarray(temps_left, dtype=float))
    tr = np.atleast_1d(np.asarray(temps_right, dtype=float))
    if tl.shape != tr.shape or tl.ndim != 1:
        raise ValueError("temps_left and temps_right must be 1D arrays of equal length.")
    if np.any(tl <= 0.0) or np.any(tr <= 0.0):
        raise ValueError("Temperature must be positive in Kelvin.")
    w_d = _engine.to_dev_f64(omega)
    t1_d = _engine.to_dev_f64(t1)
    t2_d = None if t2 is None else _engine.to_dev_f64(t2)
    tl_d = _engine.to_dev_f64(tl)
    tr_d = _engine.to_dev_f64(tr)
    out = torch.empty((tl.size, 2), dtype=torch.float64, device=w_d.device)
    _lib.check(lib.negf_fcs_cumulants(
        _lib.stream(), int(w_d.numel()), w_d.data_ptr(), t1_d.data_ptr(),
        None if t2_d is None else t2_d.data_ptr(), int(tl.size), tl_d.data_ptr(),
        tr_d.data_ptr(), out.data_ptr()))
    return out.cpu().numpy()


def heat_current_cumulants_from_spectrum(omega_rad_s: Iterable[float],
                                         transmission_vals: Iterable[float], temp_left_k: float,
                                         temp_right_k: float) -> HeatCurrentCumulants:
    """C1 [J/s] and C2 [J^2/s] of one spectrum T(omega), fcs.py:89-122."""
    omega = _grid(omega_rad_s, "omega_rad_s")
    t = np.asarray(list(transmission_vals), dtype=float)
    if t.shape != omega.shape:
        raise ValueError("transmission_vals must have the same length as omega_rad_s.")
    if np.any(t < 0.0):
        raise ValueError("transmission_vals must be non-negative.")
    c = _run(omega, t, None, [float(temp_left_k)], [float(temp_right_k)])[0]
    return HeatCurrentCumulants(c1_j_per_s=float(c[0]), c2_j2_per_s=float(c[1]))


def heat_current_cumulants_from_k_moments(omega_rad_s: Iterable[float],
                                          t_mean_vs_omega: Iterable[float],
                                          t2_mean_vs_omega: Iterable[float], temp_left_k: float,
                                          temp_right_k: float) -> HeatCurrentCumulants:
    """Exact C1 and C2 from k-resolved moments <T>, <T^2>, fcs.py:125-156."""
    omega, t1, t2 = _moments(omega_rad_s, t_mean_vs_omega, t2_mean_vs_omega)
    c = _run(omega, t1, t2, [float(temp_left_k)], [float(temp_right_k)])[0]
    return HeatCurrentCumulants(c1_j_per_s=float(c[0]), c2_j2_per_s=float(c[1]))


def _moments(omega_rad_s, t_mean, t2_mean):
    omega = _grid(omega_rad_s, "omega_rad_s")
    t1 = np.asarray(list(t_mean), dtype=float)
    t2 = np.asarray(list(t2_mean), dtype=float)
    if t1.shape != omega.shape or t2.shape != omega.shape:
        raise ValueError("t_mean_vs_omega and t2_mean_vs_omega must match omega_rad_s shape.")
    if np.any(t1 < 0.0) or np.any(t2 < 0.0):
        raise ValueError("Transmission moments must be non-negative.")
    return omega, t1, t2


def heat_current_cumulants_scan(omega_rad_s: Iterable[float], t_mean_vs_omega: Iterable[float],
                                t2_mean_vs_omega: Iterable[float] | None, temps_left,
                                temps_right) -> Array:
    """(n_pairs, 2) array of (C1, C2) for every temperature pair of a bias or
    temperature scan, in one launch (the per-pair loop of the reference's FCS
    driver, workflows/ifc_bulk.py:1787-1936).  ``t2_mean_vs_omega=None`` uses the
    single-spectrum form."""
    if t2_mean_vs_omega is None:
        omega = _grid(omega_rad_s, "omega_rad_s")
        t1 = np.asarray(list(t_mean_vs_omega), dtype=float)
        if t1.shape != omega.shape:
            raise ValueError("transmission_vals must have the same length as omega_rad_s.")
        if np.any(t1 < 0.0):
            raise ValueError("transmission_vals must be non-negative.")
        return _run(omega, t1, None, temps_left, temps_right)
    omega, t1, t2 = _moments(omega_rad_s, t_mean_vs_omega, t2_mean_vs_omega)
    return _run(omega, t1, t2, temps_left, temps_right)


def heat_current_uncertainty(cumulants: HeatCurrentCumulants, *,
                             measurement_time_s: float) -> HeatCurrentUncertainty:
    """Finite-time uncertainty from C2, fcs.py:159-177 (scalar arithmetic)."""
    tau = float(measurement_time_s)
    if tau <= 0.0:
        raise ValueError("measurement_time_s must be positive.")
    var_e = float(cumulants.c2_j2_per_s * tau)
    std_e = float(np.sqrt(max(var_e, 0.0)))
    std_j = float(np.sqrt(max(cumulants.c2_j2_per_s / tau, 0.0)))
    return HeatCurrentUncertainty(measurement_time_s=tau, std_current_j_per_s=std_j,
                                  variance_energy_j2=var_e, std_energy_j=std_e)
