"""Landauer heat current from transmission spectra on the GPU
(reference ``src/negfpy/core/thermal_current.py:15-218``).

J = (1/2pi) int [k_B T_L phi(x_L) - k_B T_R phi(x_R)] T(omega) d omega with
phi(x) = x / (e^x - 1), omega in rad/s, trapezoid rule; the omega integral runs
in FP64 on the GPU (``negf_heat_current``).  The heat_current_* drivers solve
every (omega, k) transmission of the spectrum in one batch.
"""

from __future__ import annotations

from typing import Iterable

import numpy as np

from . import _engine
from ._status import transmission_value
from .negf import (DeviceLike, LeadLike, _resolve_many, kspace_sets, point_jobs, run_jobs,
                   solve_points)
from .surface_gf import _prec

Array = np.ndarray

HBAR = 1.054_571_817e-34  # J s
KB = 1.380_649e-23  # J / K


def _to_1d_array(values: Iterable[float], name: str) -> Array:
    arr = np.asarray(list(values), dtype=float)
    if arr.ndim != 1:
        raise ValueError(f"{name} must be 1D.")
    if arr.size < 2:
        raise ValueError(f"{name} must contain at least two points.")
    if np.any(np.diff(arr) <= 0.0):
        raise ValueError(f"{name} must be strictly increasing.")
    return arr


def heat_current_from_spectrum(omegas: Iterable[float], transmission_vals: Iterable[float],
                               temp_left: float, temp_right: float) -> float:
    """Heat current in W from T(omega), omega in rad/s (thermal_current.py:50-67)."""
    w = _to_1d_array(omegas, "omegas")
    t = np.asarray(list(transmission_vals), dtype=float)
    if t.shape != w.shape:
        raise ValueError("transmission_vals must have the same length as omegas.")
    if temp_left <= 0.0 or temp_right <= 0.0:
        raise ValueError("Temperatures must be positive in Kelvin.")
    return _engine.heat_current(w, t, temp_left, temp_right)


def transmission_spectrum(omegas, device, lead_left, lead_right, *, kpoints=None, eta=1e-8,
                          precision=None, device_to_lead_left=None,
                          device_to_lead_right=None) -> Array:
    """k-averaged T(omega) of a whole spectrum in one batch (kpoints=None: one point, kpar=None).
    ``device_to_lead_left`` / ``_right`` are the couplings of ``transmission_kavg``
    (negf.py:360-397).  Raises the reference's exception of the first failing
    (omega, k) point."""
    prec = _prec(precision)
    w = np.asarray(list(omegas), dtype=float)
    ks = [None] if kpoints is None else list(kpoints)
    if len(ks) == 0:
        raise ValueError("kpoints must contain at least one k-point.")
    nk = len(ks)
    om = np.repeat(w, nk)
    kk = np.tile(np.arange(nk), w.size)
    et = np.full(om.size, float(eta))
    sets = kspace_sets(device, lead_left, lead_right, ks, device_to_lead_left=device_to_lead_left,
                       device_to_lead_right=device_to_lead_right, precision=prec) \
        if kpoints is not None else None
    if sets is not None:
        res = _engine.TransmissionSolver(sets, precision=prec).run(kk.astype(np.int32), om, et, et)
        t, st = res.t.cpu().numpy(), res.status.cpu().numpy()
    else:
        kps = _resolve_many(device, lead_left, lead_right, ks, eta=eta, eta_device=None,
                            device_to_lead_left=device_to_lead_left,
                            device_to_lead_right=device_to_lead_right,
                            contact_left_indices=None, contact_right_indices=None,
                            omega_scale=None, precision=prec)
        t, st = solve_points(kps, kk, om, et, et, precision=prec)
    vals = np.array([transmission_value(int(c), float(v)) for c, v in zip(st, t)])
    return vals.reshape(w.size, nk).mean(axis=1)


def transmission_spectra(problems, *, eta=1e-8, precision=None) -> list[Array]:
    """k-averaged T(omega) of several systems in one call, one array per system.

    ``problems``: sequence of ``(omegas, device, lead_left, lead_right)`` or
    ``(omegas, device, lead_left, lead_right, kpoints)``, each as in
    :func:`transmission_spectrum`, with any block sizes.  In FP64 every (omega, k)
    point of every system goes through one persistent kernel launch (per eight
    block layouts), whose workgroups take the points of all systems from one work
    queue; mixed precision solves each layout with its own staged launches.
    Raises the reference's exception of the first failing point of the first
    system that has one.  No counterpart in the reference, which solves one point
    per call (negf.py:306-357).
    """
    prec = _prec(precision)
    jobs, where, shapes = [], [], []
    for pi, prob in enumerate(problems):
        if len(prob) not in (4, 5):
            raise ValueError("each problem is (omegas, device, lead_left, lead_right[, kpoints]).")
        omegas, device, lead_left, lead_right = prob[:4]
        kpoints = prob[4] if len(prob) == 5 else None
        w = np.asarray(list(omegas), dtype=float)
        ks = [None] if kpoints is None else list(kpoints)
        if len(ks) == 0:
            raise ValueError("kpoints must contain at least one k-point.")
        nk = len(ks)
        om = np.repeat(w, nk)
        kk = np.tile(np.arange(nk), w.size)
        et = np.full(om.size, float(eta))
        shapes.append((w.size, nk))
        sets = kspace_sets(device, lead_left, lead_right, ks, precision=prec) \
            if kpoints is not None else None
        if sets is not None:
            jobs.append((_engine.TransmissionSolver(sets, precision=prec), kk.astype(np.int32),
                         om, et, et))
            where.append((pi, np.arange(om.size)))
            continue
        kps = _resolve_many(device, lead_left, lead_right, ks, eta=eta, eta_device=None,
                            device_to_lead_left=None, device_to_lead_right=None,
                            contact_left_indices=None, contact_right_indices=None,
                            omega_scale=None, precision=prec)
        pj, sels = point_jobs(kps, kk, om, et, et, precision=prec)
        jobs.extend(pj)
        where.extend((pi, sel) for sel in sels)
    t_all = [np.full(a * b, np.nan) for a, b in shapes]
    st_all = [np.zeros(a * b, dtype=np.int32) for a, b in shapes]
    for (pi, sel), res in zip(where, run_jobs(jobs, prec)):
        t_all[pi][sel] = res.t.cpu().numpy()
        st_all[pi][sel] = res.status.cpu().numpy()
    out = []
    for (nw, nk), t, st in zip(shapes, t_all, st_all):
        vals = np.array([transmission_value(int(c), float(v)) for c, v in zip(st, t)])
        out.append(vals.reshape(nw, nk).mean(axis=1) if vals.size else np.zeros(nw))
    return out


def heat_current_1d(omegas: Iterable[float], device: DeviceLike, lead_left: LeadLike,
                    lead_right: LeadLike, temp_left: float, temp_right: float, eta: float = 1e-8,
                    *, precision: str | None = None) -> float:
    """thermal_current.py:70-100."""
    w = _to_1d_array(omegas, "omegas")
    t = transmission_spectrum(w, device, lead_left, lead_right, eta=eta, precision=precision)
    return heat_current_from_spectrum(w, t, temp_left, temp_right)


def heat_current_per_length_2d(omegas: Iterable[float], device: DeviceLike, lead_left: LeadLike,
                               lead_right: LeadLike, kpoints: list[tuple[float, ...]],
                               transverse_length: float, temp_left: float, temp_right: float,
                               eta: float = 1e-8, *, precision: str | None = None) -> float:
    """thermal_current.py:103-139 (W/m)."""
    if transverse_length <= 0.0:
        raise ValueError("transverse_length must be positive.")
    w = _to_1d_array(omegas, "omegas")
    if len(kpoints) == 0:
        raise ValueError("kpoints must contain at least one k-point.")
    t = transmission_spectrum(w, device, lead_left, lead_right, kpoints=kpoints, eta=eta,
                              precision=precision)
    return float(heat_current_from_spectrum(w, t, temp_left, temp_right) / transverse_length)


def heat_current_density_3d(omegas: Iterable[float], device: DeviceLike, lead_left: LeadLike,
                            lead_right: LeadLike, kpoints: list[tuple[float, ...]],
                            transverse_area: float, temp_left: float, temp_right: float,
                            eta: float = 1e-8, *, precision: str | None = None) -> float:
    """thermal_current.py:142-178 (W/m^2)."""
    if transverse_area <= 0.0:
        raise ValueError("transverse_area must be positive.")
    w = _to_1d_array(omegas, "omegas")
    if len(kpoints) == 0:
        raise ValueError("kpoints must contain at least one k-point.")
    t = transmission_spectrum(w, device, lead_left, lead_right, kpoints=kpoints, eta=eta,
                              precision=precision)
    return float(heat_current_from_spectrum(w, t, temp_left, temp_right) / transverse_area)


def transverse_length_from_vector(transverse_vec: Iterable[float],
                                  transport_dir: Iterable[float]) -> float:
    """thermal_current.py:181-199."""
    t = np.asarray(list(transverse_vec), dtype=float)
    e = np.asarray(list(transport_dir), dtype=float)
    if t.ndim != 1 or e.ndim != 1 or t.shape != e.shape:
        raise ValueError("transverse_vec and transport_dir must be 1D with the same size.")
    en = np.linalg.norm(e)
    if en == 0.0:
        raise ValueError("transport_dir must be non-zero.")
    eh = e / en
    length = float(np.linalg.norm(t - np.dot(t, eh) * eh))
    if length == 0.0:
        raise ValueError("transverse_vec has zero component orthogonal to transport_dir.")
    return length


def transverse_area_from_vectors(transverse_vec_1: Iterable[float],
                                 transverse_vec_2: Iterable[float],
                                 transport_dir: Iterable[float]) -> float:
    """thermal_current.py:202-218."""
    t1 = np.asarray(list(transverse_vec_1), dtype=float)
    t2 = np.asarray(list(transverse_vec_2), dtype=float)
    e = np.asarray(list(transport_dir), dtype=float)
    if t1.shape != (3,) or t2.shape != (3,) or e.shape != (3,):
        raise ValueError("All vectors must be 3D.")
    en = np.linalg.norm(e)
    if en == 0.0:
        raise ValueError("transport_dir must be non-zero.")
    return float(abs(np.dot(np.cross(t1, t2), e / en)))
