"""Lead surface phonon DOS on the GPU (reference ``src/negfpy/core/dos.py:18-251``).

DOS(omega, k) = (omega / pi) Re Tr[i (g - g^H)] of the retarded surface Green's
function; the trace is taken in FP64 inside the decimation kernel.  k-averages
and spectra batch every (omega, k) point of a call into one launch.
"""

from __future__ import annotations

from typing import Iterable

import numpy as np

from . import _engine
from ._status import raise_for_lead_status
from .negf import _resolve_lead_blocks
from .surface_gf import _prec, canonical_method, pack_lead, surface_route
from .types import KPar, LeadBlocks, LeadKSpace

Array = np.ndarray
LeadLike = LeadBlocks | LeadKSpace


def _lead_sets(lead: LeadLike, kpoints):
    blocks = [_resolve_lead_blocks(lead, k) for k in kpoints]
    return blocks, np.stack([pack_lead(b.d00, b.d01, b.d10) for b in blocks])


def _dos_points(lead, kpoints, omegas, etas, normalize_per_mode, method, precision):
    """DOS and status for every (omega_i, k_i, eta_i) triple (k given per point)."""
    route = surface_route(method)
    uniq = list(dict.fromkeys(range(len(kpoints))))
    blocks, sets = _lead_sets(lead, [kpoints[i] for i in uniq])
    res = _engine.surface_gf_batch(sets, np.arange(len(kpoints)), omegas, etas,
                                   precision=_prec(precision), route=route)
    dos = res.dos.cpu().numpy()
    st = res.status.cpu().numpy()
    if normalize_per_mode:
        dos = dos / np.array([b.d00.shape[0] for b in blocks], dtype=float)
    return dos, st


def lead_surface_dos(omega: float, lead: LeadLike, eta: float = 1e-8, kpar: KPar = None, *,
                     normalize_per_mode: bool = False, surface_gf_method: str = "sancho_rubio",
                     precision: str | None = None) -> float:
    """dos.py:31-61."""
    if omega < 0.0:
        raise ValueError("omega must be non-negative.")
    dos, st = _dos_points(lead, [kpar], [float(omega)], [float(eta)], normalize_per_mode,
                          surface_gf_method, precision)
    raise_for_lead_status(int(st[0]))
    return float(dos[0])


def lead_surface_dos_kavg(omega: float, lead: LeadLike, kpoints: list[tuple[float, ...]],
                          eta: float = 1e-8, *, normalize_per_mode: bool = False,
                          surface_gf_method: str = "sancho_rubio",
                          precision: str | None = None) -> float:
    """dos.py:64-88."""
    if len(kpoints) == 0:
        raise ValueError("kpoints must contain at least one k-point.")
    if omega < 0.0:
        raise ValueError("omega must be non-negative.")
    nk = len(kpoints)
    dos, st = _dos_points(lead, list(kpoints), np.full(nk, float(omega)), np.full(nk, float(eta)),
                          normalize_per_mode, surface_gf_method, precision)
    for c in st:
        raise_for_lead_status(int(c))
    return float(np.mean(dos))


def lead_surface_dos_kavg_adaptive(omega: float, lead: LeadLike, kpoints: list[tuple[float, ...]],
                                   eta_values: tuple[float, ...] = (1e-8, 1e-7, 1e-6, 1e-5),
                                   min_success_fraction: float = 0.0, *,
                                   normalize_per_mode: bool = False,
                                   surface_gf_method: str = "sancho_rubio",
                                   precision: str | None = None) -> tuple[float, dict[str, object]]:
    """dos.py:91-158: per-k eta ladder, each rung one batch over the pending k points."""
    if len(kpoints) == 0:
        raise ValueError("kpoints must contain at least one k-point.")
    if len(eta_values) == 0:
        raise ValueError("eta_values must contain at least one eta value.")
    if not (0.0 <= min_success_fraction <= 1.0):
        raise ValueError("min_success_fraction must be in [0, 1].")
    canonical_method(surface_gf_method)
    kpoints = list(kpoints)
    nk = len(kpoints)
    val = [None] * nk
    used = [None] * nk
    pending = list(range(nk))
    for eta in eta_values:
        if not pending:
            break
        if omega < 0.0:
            break  # every attempt raises ValueError in the reference
        dos, st = _dos_points(lead, [kpoints[k] for k in pending], np.full(len(pending), float(omega)),
                              np.full(len(pending), float(eta)), normalize_per_mode,
                              surface_gf_method, precision)
        still = []
        for j, k in enumerate(pending):
            if st[j] == 0 and np.isfinite(dos[j]):
                val[k] = float(dos[j])
                used[k] = float(eta)
            else:
                still.append(k)
        pending = still
    vals = [v for v in val if v is not None]
    n_success = len(vals)
    success_fraction = float(n_success / nk)
    if n_success == 0:
        raise RuntimeError("No k-point converged for adaptive k-averaged DOS.")
    if success_fraction < min_success_fraction:
        raise RuntimeError(
            "Adaptive k-averaged DOS failed success-fraction threshold "
            f"(got {success_fraction:.3f}, required >= {min_success_fraction:.3f}).")
    hist = {float(e): int(sum(1 for u in used if u == float(e))) for e in eta_values}
    info: dict[str, object] = {"n_total": int(nk), "n_success": int(n_success),
                               "n_failed": int(nk - n_success),
                               "success_fraction": success_fraction, "eta_histogram": hist}
    return float(np.mean(vals)), info


def lead_surface_dos_spectrum(omegas: Iterable[float], lead: LeadLike, *,
                              kpoints: list[tuple[float, ...]] | None = None, eta: float = 1e-8,
                              eta_values: tuple[float, ...] | None = None,
                              min_success_fraction: float = 0.0, normalize_per_mode: bool = False,
                              surface_gf_method: str = "sancho_rubio",
                              precision: str | None = None) -> Array:
    """dos.py:161-213.  Without an eta ladder all (omega, k) points run in one batch."""
    w = np.asarray(list(omegas), dtype=float)
    if w.ndim != 1 or w.size == 0:
        raise ValueError("omegas must be a non-empty 1D iterable.")
    if kpoints is None or eta_values is None:
        if np.any(w < 0.0):
            raise ValueError("omega must be non-negative.")
        ks = [None] if kpoints is None else list(kpoints)
        if len(ks) == 0:
            raise ValueError("kpoints must contain at least one k-point.")
        nk = len(ks)
        route = surface_route(surface_gf_method)
        blocks, sets = _lead_sets(lead, ks)
        om = np.repeat(w, nk)
        kk = np.tile(np.arange(nk), w.size)
        res = _engine.surface_gf_batch(sets, kk, om, np.full(om.size, float(eta)),
                                       precision=_prec(precision), route=route)
        st = res.status.cpu().numpy()
        dos = res.dos.cpu().numpy()
        for c in st:
            raise_for_lead_status(int(c))
        if normalize_per_mode:
            dos = dos / np.array([blocks[k].d00.shape[0] for k in kk], dtype=float)
        return dos.reshape(w.size, nk).mean(axis=1)
    out = np.zeros_like(w)
    for i, wi in enumerate(w):
        out[i], _ = lead_surface_dos_kavg_adaptive(
            float(wi), lead, kpoints, eta_values=eta_values,
            min_success_fraction=min_success_fraction, normalize_per_mode=normalize_per_mode,
            surface_gf_method=surface_gf_method, precision=precision)
    return out


def leads_surface_dos_spectrum(omegas: Iterable[float], lead_left: LeadLike, lead_right: LeadLike,
                               *, kpoints: list[tuple[float, ...]] | None = None,
                               eta: float = 1e-8, eta_values: tuple[float, ...] | None = None,
                               min_success_fraction: float = 0.0, normalize_per_mode: bool = False,
                               surface_gf_method: str = "sancho_rubio",
                               precision: str | None = None) -> dict[str, Array]:
    """dos.py:216-251."""
    omegas = list(omegas)
    kw = dict(kpoints=kpoints, eta=eta, eta_values=eta_values,
              min_success_fraction=min_success_fraction, normalize_per_mode=normalize_per_mode,
              surface_gf_method=surface_gf_method, precision=precision)
    return {"left": lead_surface_dos_spectrum(omegas, lead_left, **kw),
            "right": lead_surface_dos_spectrum(omegas, lead_right, **kw)}
