"""Bulk IFC transmission spectra from a study configuration
(reference ``src/negfpy/workflows/ifc_bulk.py``: ``_prepare_kmesh`` :154-216,
``_estimate_omega_max_from_ifc`` :229-257, ``_build_omega_grid`` :534-578,
``_resolve_eta_values_for_omega`` :581-595, ``_run_transmission`` :1059-1288, the
model set-up of ``run_ifc_bulk`` :2230-2290, ``_save_spectrum_data`` :1551-1581).

The reference solves every frequency in its own task (a Python loop, or a process
pool), each calling ``transmission_kavg_adaptive`` for its k points.  Here every
eta rung of the whole spectrum is one batched GPU solve over all pending
(omega, k) instances (``transmission_kavg_adaptive_spectrum`` engine), grouped by
k mesh (low/high) and eta ladder.
"""

from __future__ import annotations

import json
from pathlib import Path
from typing import Any

import numpy as np

from negfpy_amd.core.negf import _adaptive_batch, _KSolver, _t_max
from negfpy_amd.core.thermal_current import transverse_area_from_vectors, transverse_length_from_vector
from negfpy_amd.core.surface_gf import _prec, canonical_method, surface_route
from negfpy_amd.io import read_ifc
from negfpy_amd.modeling import (
    BuildConfig,
    apply_mass_mode,
    apply_transverse_stability_controls,
    interface_contacts,
    build_material_kspace_params,
    enforce_translational_asr_on_self_term,
    reorient_for_transport,
    transverse_cutoff,
)
from negfpy_amd.modeling.units import (
    qe_cm1_to_omega,
    qe_ev_to_omega,
    qe_omega_to_cm1,
    qe_omega_to_thz,
    qe_thz_to_omega,
)
from negfpy_amd.models import material_kspace_device

DEFAULT_ETA_VALUES = (1e-8, 1e-7, 1e-6, 1e-5)
ANGSTROM_M = 1.0e-10
BOHR_M = 5.29177210903e-11


# --------------------------------------------------------------------------
# k mesh and omega grid
# --------------------------------------------------------------------------

def _kline(nk: int, mode: str):
    if nk <= 0:
        raise ValueError("nk must be positive.")
    if nk == 1:
        return np.array([0.0]), "gamma"
    m = mode
    if m == "auto":
        m = "shifted" if nk % 2 == 0 else "centered"
    if m == "shifted":
        if nk % 2 != 0:
            raise ValueError("shifted k-mesh requires even nk to avoid sampling Gamma exactly.")
        return -np.pi + (np.arange(nk, dtype=float) + 0.5) * (2.0 * np.pi / nk), m
    if m == "centered":
        return (np.arange(nk, dtype=float) - (nk // 2)) * (2.0 * np.pi / nk), m
    if m in ("legacy_endpoint", "legacy"):
        return np.linspace(-np.pi, np.pi, nk, endpoint=True), "legacy_endpoint"
    raise ValueError(f"Unknown k-mesh mode '{mode}'.")


def build_kpoints(dim: int, nky: int, nkz: int, mode: str):
    if dim == 0:
        return [tuple()], {"dimension": 0, "mode_y": "gamma", "mode_z": "gamma", "nky": 1, "nkz": 1}
    if dim == 1:
        ky, my = _kline(nky, mode)
        return [(float(v),) for v in ky], {"dimension": 1, "mode_y": my, "mode_z": "none",
                                           "nky": nky, "nkz": 1}
    if dim == 2:
        ky, my = _kline(nky, mode)
        kz, mz = _kline(nkz, mode)
        return [(float(a), float(b)) for a in ky for b in kz], {
            "dimension": 2, "mode_y": my, "mode_z": mz, "nky": nky, "nkz": nkz}
    raise ValueError("kmesh.dimension must be 0, 1, or 2.")


def prepare_kmesh(cfg: dict) -> dict:
    k = dict(cfg.get("kmesh", {}))
    dim = int(k.get("dimension", 2))
    mode = str(k.get("mode", "auto"))
    nk = int(k.get("nk", 1))
    nky = int(k.get("nky", nk))
    nkz = int(k.get("nkz", nk))
    nk_low = int(k.get("nk_low", max(nky, nkz)))
    nky_low = int(k.get("nky_low", max(nky, nk_low)))
    nkz_low = int(k.get("nkz_low", max(nkz, nk_low)))
    hi, info_hi = build_kpoints(dim, nky, nkz, mode)
    lo, info_lo = build_kpoints(dim, nky_low, nkz_low, mode)
    return {"dim": dim, "mode": mode, "low_cm1": float(k.get("low_cm1", 0.0)),
            "high": {"kpoints": hi, **info_hi}, "low": {"kpoints": lo, **info_lo}}


def _kgrid(nk: int, centered: bool):
    if nk <= 0:
        raise ValueError("nk must be positive.")
    if centered:
        return (np.arange(nk, dtype=float) - (nk // 2)) * (2.0 * np.pi / nk)
    return -np.pi + (np.arange(nk, dtype=float) + 0.5) * (2.0 * np.pi / nk)


def estimate_omega_max(ifc, *, nkx: int, nky: int, nkz: int, centered_grid: bool,
                       safety_factor: float = 1.02) -> float:
    """Largest bulk phonon frequency on a k grid, times a safety factor."""
    nd = len(ifc.masses) * ifc.dof_per_atom
    minv = np.diag(1.0 / np.sqrt(np.repeat(np.asarray(ifc.masses, dtype=float), ifc.dof_per_atom)))
    terms = [(int(t.dx), int(t.dy), int(t.dz), np.asarray(t.block, dtype=np.complex128))
             for t in ifc.terms]
    wmax = 0.0
    for kx in _kgrid(nkx, centered_grid):
        for ky in _kgrid(nky, centered_grid):
            for kz in _kgrid(nkz, centered_grid):
                phi = np.zeros((nd, nd), dtype=np.complex128)
                for dx, dy, dz, b in terms:
                    phi += b * np.exp(1j * (kx * dx + ky * dy + kz * dz))
                d = minv @ phi @ minv
                vals = np.linalg.eigvalsh(0.5 * (d + d.conj().T))
                wmax = max(wmax, float(np.sqrt(max(float(np.max(vals.real)), 0.0))))
    return float(max(wmax * safety_factor, 1e-8))


def build_omega_grid(cfg: dict, ifc):
    o = dict(cfg.get("omega", {}))
    wmin = float(o.get("min", 0.0))
    npts = int(o.get("n_points", o.get("nw", 160)))
    if npts <= 1:
        raise ValueError("omega.n_points must be > 1.")
    if o.get("fmax_cm1", None) is not None:
        wmax, src = qe_cm1_to_omega(float(o["fmax_cm1"])), "fmax_cm1"
    elif o.get("fmax_thz", None) is not None:
        wmax, src = qe_thz_to_omega(float(o["fmax_thz"])), "fmax_thz"
    elif o.get("max", None) is not None:
        wmax, src = float(o["max"]), "omega.max"
    elif bool(o.get("auto_max", True)):
        nk = int(o.get("auto_sample_nk", 12))
        wmax = estimate_omega_max(ifc, nkx=nk, nky=nk, nkz=nk,
                                  centered_grid=bool(o.get("auto_centered_grid", False)),
                                  safety_factor=float(o.get("safety_factor", 1.02)))
        src = "auto_ifc_estimate"
    else:
        raise ValueError("No omega maximum specified. Set omega.max or omega.auto_max=true.")
    if wmax <= wmin:
        raise ValueError("Resolved omega max must be greater than omega min.")
    return np.linspace(wmin, wmax, npts), {
        "omega_min_qe": float(wmin), "omega_max_qe": float(wmax), "n_points": int(npts),
        "omega_max_source": src,
        "omega_max_thz": float(np.asarray(qe_omega_to_thz(wmax), dtype=float)),
        "omega_max_cm1": float(np.asarray(qe_omega_to_cm1(wmax), dtype=float))}


def eta_values_for_omega(omega: float, omega_cm1: float, *, eta_values, max_eta_over_omega,
                         eta_ratio_cm1_min):
    if max_eta_over_omega is None:
        return tuple(eta_values)
    if eta_ratio_cm1_min is not None and omega_cm1 < eta_ratio_cm1_min:
        return tuple(eta_values)
    kept = tuple(e for e in eta_values if (e / omega) <= max_eta_over_omega)
    return kept if kept else (min(eta_values),)


# --------------------------------------------------------------------------
# model
# --------------------------------------------------------------------------

def build_transport_model(cfg: dict, cfg_dir: Path | str = ".") -> dict[str, Any]:
    """IFC file -> conditioned force-constant model -> leads, couplings and device."""
    cfg_dir = Path(cfg_dir)
    ic = dict(cfg.get("ifc", {}))
    raw = ic.get("path", "si444.fc")
    path = cfg_dir / raw if not Path(raw).is_absolute() else Path(raw)
    for side in ("left_path", "right_path", "center_path"):
        other = ic.get(side, raw)
        if (cfg_dir / other).resolve() != path.resolve():
            raise NotImplementedError(
                "Different left/center/right IFC files are not implemented in this bulk runner yet. "
                "Set left_path=center_path=right_path or use ifc.path for identical leads/device.")
    ifc = read_ifc(path, reader=str(ic.get("reader", "qe")))
    asr_max = None
    if bool(ic.get("enforce_asr", False)):
        ifc, asr_max = enforce_translational_asr_on_self_term(ifc)
    mc = dict(cfg.get("model", {}))
    direction = int(mc.get("transport_direction", 1))
    ifc = reorient_for_transport(ifc, direction)
    dyc = ic.get("dy_cutoff", None)
    dzc = ic.get("dz_cutoff", None)
    ifc, n_removed = transverse_cutoff(ifc, dy_cutoff=None if dyc is None else int(dyc),
                                       dz_cutoff=None if dzc is None else int(dzc))
    bc = BuildConfig(
        onsite_pinning=float(mc.get("onsite_pinning", 0.0)),
        principal_layer_size=(None if mc.get("principal_layer_size", None) is None
                              else int(mc["principal_layer_size"])),
        auto_principal_layer_enlargement=bool(mc.get("auto_principal_layer_enlargement", True)),
        infer_fc01_from_negative_dx=bool(mc.get("infer_fc01_from_negative_dx", True)),
        enforce_hermitian_fc00=bool(mc.get("enforce_hermitian_fc00", True)),
        nyquist_split_half=bool(mc.get("nyquist_split_half", False)),
        dtype=str(mc.get("dtype", "complex128")))
    params = build_material_kspace_params(ifc, bc)
    params, mass_mode = apply_mass_mode(params, mc.get("mass_mode", "ifc"))
    params, nyq_info, pm_info = apply_transverse_stability_controls(
        params, ifc_metadata=dict(ifc.metadata),
        drop_nyquist=bool(mc.get("drop_nyquist_transverse", False)))
    lead_l, lead_r, v_l, v_r = interface_contacts(params)
    n_layers = int(mc.get("n_layers", 30))
    return {"ifc": ifc, "params": params, "lead_left": lead_l, "lead_right": lead_r,
            "device_to_lead_left": v_l, "device_to_lead_right": v_r,
            "device": material_kspace_device(n_layers=n_layers, params=params),
            "info": {"n_layers": n_layers, "ndof": int(params.ndof), "mass_mode": mass_mode,
                     "nyquist_filter": nyq_info, "pm_symmetry": pm_info,
                     "n_terms_removed_by_cutoff": int(n_removed), "asr_residual_max": asr_max}}


# --------------------------------------------------------------------------
# spectrum
# --------------------------------------------------------------------------

def transmission_spectrum_from_config(cfg: dict, model: dict, omegas, kmesh: dict, *,
                                      precision: str | None = None) -> dict[str, Any]:
    """The transmission spectrum of a study configuration (``_run_transmission``):
    per omega the low or high k mesh, its eta ladder and scheme, the reference's
    filters; every eta rung of all omegas sharing a mesh and ladder is one batch."""
    sc = dict(cfg.get("solver", {}))
    method = str(sc.get("surface_gf_method", "generalized_eigen_svd"))
    canonical_method(method)
    scheme = str(sc.get("eta_scheme", "adaptive")).lower().replace("-", "_")
    eta_fixed = float(sc.get("eta_fixed", 1e-4))
    base = tuple(float(v) for v in sc.get("eta_values", DEFAULT_ETA_VALUES))
    eta_device = None if sc.get("eta_device", None) is None else float(sc["eta_device"])
    min_frac = float(sc.get("min_success_fraction", 0.7))
    nonneg = float(sc.get("nonnegative_tolerance", 1e-8))
    mcf = sc.get("max_channel_factor", 1.5)
    mcf = None if mcf is None else float(mcf)
    meo = sc.get("max_eta_over_omega", None)
    meo = None if meo is None else float(meo)
    ratio_min = sc.get("eta_ratio_cm1_min", None)
    ratio_min = None if ratio_min is None else float(ratio_min)
    mode = str(sc.get("omega_scale_mode", "none")).lower()
    if mode == "ev":
        omega_scale = float(np.asarray(qe_ev_to_omega(1.0), dtype=float))
    elif mode == "none":
        omega_scale = None
    else:
        raise ValueError("solver.omega_scale_mode must be 'none' or 'ev'.")
    prec = _prec(precision)

    w = np.asarray(omegas, dtype=float)
    vals = np.zeros_like(w)
    records: list = [None] * w.size
    failures: list = [None] * w.size
    groups: dict = {}
    cm1 = {}
    for i, wi in enumerate(w):
        if wi <= 0.0:
            continue
        c = float(np.asarray(qe_omega_to_cm1(float(wi)), dtype=float))
        cm1[i] = c
        mesh = "low" if c <= float(kmesh["low_cm1"]) else "high"
        lad = eta_values_for_omega(float(wi), c, eta_values=base, max_eta_over_omega=meo,
                                   eta_ratio_cm1_min=ratio_min)
        if scheme == "fixed":
            lad = (eta_fixed,)
        groups.setdefault((mesh, lad), []).append(i)
    solvers: dict = {}
    for (mesh, lad), idx in groups.items():
        kpts = kmesh[mesh]["kpoints"]
        if mesh not in solvers:
            solvers[mesh] = (_KSolver(model["device"], model["lead_left"], model["lead_right"],
                                      kpts, eta_device=eta_device, precision=prec,
                                      route=surface_route(method),
                                      device_to_lead_left=model["device_to_lead_left"],
                                      device_to_lead_right=model["device_to_lead_right"],
                                      contact_left_indices=None, contact_right_indices=None,
                                      omega_scale=omega_scale),
                             _t_max(model["lead_left"], kpts, mcf))
        solver, t_max = solvers[mesh]
        res = _adaptive_batch(w[idx], solver, t_max, [lad] * len(idx), eta_device=eta_device,
                              global_eta=(scheme != "adaptive"), min_success_fraction=min_frac,
                              nonnegative_tolerance=nonneg, collect_rejected=False)
        for i, r in zip(idx, res):
            if r[0] != "ok":
                vals[i] = np.nan
                failures[i] = {"omega_qe": float(w[i]), "omega_cm1": cm1[i], "error": r[1]}
                continue
            vals[i] = float(r[1])
            info = r[2]
            hist = dict(info.get("eta_histogram", {}))
            used = [float(e) for e, c in hist.items() if int(c) > 0]
            emax = max(used) if used else np.nan
            records[i] = {"omega_qe": float(w[i]), "omega_cm1": cm1[i],
                          "n_failed_kpoints": int(info.get("n_failed", 0)),
                          "success_fraction": float(info.get("success_fraction", 0.0)),
                          "eta_histogram": hist, "eta_selected": info.get("eta_selected", None),
                          "max_eta_used": None if not np.isfinite(emax) else float(emax),
                          "max_eta_over_omega": None if not np.isfinite(emax)
                          else float(emax / w[i])}
    if np.all(~np.isfinite(vals)):
        raise RuntimeError("Transmission calculation failed for all omega points.")
    vals = np.where(np.isfinite(vals), vals, np.nan)
    vals = np.where(vals < 0.0, np.maximum(vals, -1e-12), vals)
    vals = np.where(np.isfinite(vals), np.clip(vals, 0.0, None), np.nan)
    return {"values": vals, "eta_records": [r for r in records if r is not None],
            "omega_failures": [f for f in failures if f is not None]}


def save_spectrum_tsv(path, omegas, values, *, value_name: str = "transmission",
                      metadata: dict | None = None, extra_columns: dict | None = None) -> None:
    """The reference's spectrum TSV: an optional ``# meta_json:`` line, a ``# `` header,
    then omega_qe (%.10e), THz and cm^-1 (%.8f), the value and extras (%.12e or nan)."""
    w = np.asarray(omegas, dtype=float)
    thz = np.asarray(qe_omega_to_thz(w), dtype=float)
    c = np.asarray(qe_omega_to_cm1(w), dtype=float)
    extras = {k: np.asarray(v, dtype=float) for k, v in (extra_columns or {}).items()}
    for k, v in extras.items():
        if v.shape != w.shape:
            raise ValueError(f"extra column '{k}' must have same shape as omegas.")
    fmt = lambda x: "nan" if not np.isfinite(x) else f"{float(x):.12e}"
    p = Path(path)
    p.parent.mkdir(parents=True, exist_ok=True)
    with p.open("w", encoding="utf-8") as fh:
        if metadata:
            fh.write(f"# meta_json: {json.dumps(metadata, sort_keys=True)}\n")
        fh.write("# " + "\t".join(["omega_qe", "freq_thz", "freq_cm^-1", value_name]
                                   + list(extras)) + "\n")
        for i in range(w.size):
            row = [f"{float(w[i]):.10e}", f"{float(thz[i]):.8f}", f"{float(c[i]):.8f}",
                   fmt(values[i])] + [fmt(extras[k][i]) for k in extras]
            fh.write("\t".join(row) + "\n")


def _lattice_vectors_m(cfg: dict, ifc):
    mc = dict(cfg.get("model", {}))
    lv = mc.get("lattice_vectors_angstrom", None)
    if lv is not None:
        lv = np.asarray(lv, dtype=float)
        if lv.shape != (3, 3):
            raise ValueError("model.lattice_vectors_angstrom must be a 3x3 array when provided.")
        return lv * ANGSTROM_M
    if ifc.lattice_vectors is None:
        return None
    src = str(ifc.metadata.get("source_format", "")).lower()
    if src == "phonopy_force_constants":
        scale = ANGSTROM_M
    elif src == "qe_q2r_fc":
        scale = BOHR_M
    else:
        raise ValueError("Unknown lattice-vector unit for area normalization. "
                         "Supported source_format: phonopy_force_constants, qe_q2r_fc.")
    return np.asarray(ifc.lattice_vectors, dtype=float) * scale


def effective_area_m2(lv_m, *, transport_direction: int, thickness_m: float | None = None,
                      thickness_direction: int | None = None) -> float:
    """Transverse cell area normal to transport, or in-plane length x thickness for a
    sheet (``_compute_effective_area_m2``, ifc_bulk.py:361-392)."""
    lv = np.asarray(lv_m, dtype=float)
    if lv.shape != (3, 3):
        raise ValueError("lattice_vectors must be shape (3,3) to compute transverse area.")
    t = int(transport_direction) - 1
    if t not in (0, 1, 2):
        raise ValueError("transport_direction must be one of 1, 2, 3.")
    axes = [a for a in (0, 1, 2) if a != t]
    if thickness_m is None:
        return float(transverse_area_from_vectors(lv[axes[0]], lv[axes[1]], lv[t]))
    if thickness_m <= 0.0:
        raise ValueError("model.thickness_angstrom must be positive when provided.")
    if thickness_direction is None:
        thick = axes[int(np.argmax([np.linalg.norm(lv[a]) for a in axes]))]
    else:
        thick = int(thickness_direction) - 1
        if thick not in axes:
            raise ValueError("model.thickness_direction must be one of the transverse lattice directions.")
    inplane = axes[0] if axes[1] == thick else axes[1]
    return float(transverse_length_from_vector(lv[inplane], lv[t]) * float(thickness_m))


def spectrum_metadata(cfg: dict, ifc):
    """The TSV header metadata of a transmission run and its transverse area (m^2, or
    None when the lattice is unknown) (ifc_bulk.py:2351-2383)."""
    mc = dict(cfg.get("model", {}))
    direction = int(mc.get("transport_direction", 1))
    th = mc.get("thickness_angstrom", None)
    th = None if th is None else float(th)
    td = mc.get("thickness_direction", None)
    td = None if td is None else int(td)
    lvo = mc.get("lattice_vectors_angstrom", None)
    meta = {"transport_direction": direction, "thickness_angstrom": th, "thickness_direction": td,
            "lattice_vectors_override_angstrom": None if lvo is None
            else np.asarray(lvo, dtype=float).tolist(),
            "atom_symbols": None if ifc.atom_symbols is None else list(ifc.atom_symbols),
            "lattice_vectors_raw": None if ifc.lattice_vectors is None
            else np.asarray(ifc.lattice_vectors, dtype=float).tolist(),
            "source_format": str(ifc.metadata.get("source_format", ""))}
    area = None
    lv = _lattice_vectors_m(cfg, ifc)
    if lv is not None:
        try:
            area = effective_area_m2(lv, transport_direction=direction,
                                     thickness_m=None if th is None else th * ANGSTROM_M,
                                     thickness_direction=td)
            meta["transverse_area_m2"] = float(area)
        except Exception as exc:
            meta["transverse_area_error"] = str(exc)
            area = None
    return meta, area


def run_transmission_config(config, *, base_dir=None, output_path=None,
                            precision: str | None = None) -> dict[str, Any]:
    """Compute the transmission spectrum of a study configuration (a JSON file, or
    its parsed dict with ``base_dir`` for relative IFC paths) and, with
    ``output_path``, write the reference's ``transmission.tsv`` there."""
    if isinstance(config, dict):
        cfg = config
        base = Path(".") if base_dir is None else Path(base_dir)
    else:
        path = Path(config)
        cfg = json.loads(path.read_text(encoding="utf-8"))
        base = path.parent if base_dir is None else Path(base_dir)
    calc = str(cfg.get("run", {}).get("calculation", "transmission")).lower()
    if calc != "transmission":
        raise NotImplementedError(f"calculation '{calc}' is not a transmission run.")
    model = build_transport_model(cfg, base)
    omegas, omega_info = build_omega_grid(cfg, model["ifc"])
    kmesh = prepare_kmesh(cfg)
    res = transmission_spectrum_from_config(cfg, model, omegas, kmesh, precision=precision)
    meta, area = spectrum_metadata(cfg, model["ifc"])
    if output_path is not None:
        extra = None
        if area is not None and area > 0.0:
            extra = {"transmission_per_m2": np.asarray(res["values"], dtype=float) / area}
        save_spectrum_tsv(output_path, omegas, res["values"], metadata=meta, extra_columns=extra)
    return {"metadata": meta, "omegas": omegas, "omega_info": omega_info, "kmesh": kmesh,
            "model": model["info"], **res}
