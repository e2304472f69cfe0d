"""Config-driven bulk runs with study bookkeeping (reference
``src/negfpy/workflows/ifc_bulk.py``: ``run_ifc_bulk`` :2061-2211 and :2340-2618,
``_run_thermal_conductance`` :1444-1503, ``_load_transmission_spectrum`` :441-482,
``_save_thermal_conductance_data`` :1596-1617, helpers :61-151).

Two calculations run here: ``transmission`` (the batched GPU spectrum of
``ifc_transmission``) and ``thermal_conductance`` (G(T) of a transmission TSV, the
integral on the GPU).  Output layout, file names, TSV formats, report and manifest
fields follow the reference, so a study tree written here reads like one of its own.
The reference's other calculations (surface DOS, dispersion, FCS) are not batched
transport runs and raise ``NotImplementedError``.
"""

from __future__ import annotations

import hashlib
import json
import re
import socket
import subprocess
import time
from datetime import datetime, timezone
from pathlib import Path
from typing import Any

import numpy as np

from negfpy_amd.core.thermal import thermal_conductance_from_spectrum
from negfpy_amd.modeling.units import qe_omega_to_cm1, qe_omega_to_thz

CALCULATIONS = ("transmission", "dos", "dispersion", "thermal_conductance", "fcs")


# --------------------------------------------------------------------------
# bookkeeping helpers (ifc_bulk.py:61-151)
# --------------------------------------------------------------------------

def utc_now_iso() -> str:
    return datetime.now(timezone.utc).isoformat()


def utc_stamp() -> str:
    return datetime.now(timezone.utc).strftime("%Y%m%d_%H%M%SZ")


def sanitize_token(value: str) -> str:
    """Runs of characters outside [A-Za-z0-9._-] become one underscore."""
    t = re.sub(r"[^A-Za-z0-9._-]+", "_", value.strip())
    return t or "unnamed"


def sha256_file(path: Path) -> str:
    h = hashlib.sha256()
    with Path(path).open("rb") as fh:
        for chunk in iter(lambda: fh.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def resolve_path(base_dir: Path, path_like) -> Path:
    p = Path(path_like)
    return p if p.is_absolute() else Path(base_dir) / p


def _git(cwd: Path, *args) -> str | None:
    try:
        return subprocess.check_output(["git", *args], cwd=str(cwd), text=True,
                                       stderr=subprocess.DEVNULL)
    except Exception:
        return None


def git_head_sha(cwd: Path) -> str | None:
    out = _git(cwd, "rev-parse", "HEAD")
    if out is None:
        return None
    return out.strip() or None


def git_dirty(cwd: Path) -> bool | None:
    out = _git(cwd, "status", "--porcelain")
    return None if out is None else bool(out.strip())


def to_builtin(value: Any) -> Any:
    """Paths, arrays and numpy scalars to JSON-ready Python values, recursively."""
    if isinstance(value, Path):
        return str(value)
    if isinstance(value, np.ndarray):
        return value.tolist()
    if isinstance(value, (np.floating, np.integer)):
        return value.item()
    if isinstance(value, dict):
        return {str(k): to_builtin(v) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        return [to_builtin(v) for v in value]
    return value


def load_json_config(path: Path) -> dict:
    with Path(path).open("r", encoding="utf-8-sig") as fh:
        cfg = json.load(fh)
    if not isinstance(cfg, dict):
        raise ValueError("Input config must be a JSON object.")
    return cfg


def save_json(path: Path, payload: dict) -> None:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with path.open("w", encoding="utf-8") as fh:
        json.dump(to_builtin(payload), fh, indent=2, sort_keys=True)
        fh.write("\n")


def append_manifest_jsonl(path: Path, payload: dict) -> None:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with path.open("a", encoding="utf-8") as fh:
        fh.write(json.dumps(to_builtin(payload), sort_keys=True) + "\n")


# --------------------------------------------------------------------------
# thermal conductance from a transmission TSV
# --------------------------------------------------------------------------

def load_transmission_spectrum(path) -> dict[str, Any]:
    """A transmission TSV as written by a transmission run: ``# meta_json:`` (optional,
    unreadable JSON gives {}), a ``# `` column header, whitespace-separated rows of at
    least four numbers; omega_qe is column 0, T column 3, T per m^2 column 4 (or the
    column the header names ``transmission_per_m2``)."""
    meta: dict = {}
    header = None
    rows = []
    with Path(path).open("r", encoding="utf-8") as fh:
        for raw in fh:
            line = raw.strip()
            if not line:
                continue
            if line.startswith("#"):
                txt = line[1:].strip()
                if txt.startswith("meta_json:"):
                    try:
                        meta = json.loads(txt.split("meta_json:", 1)[1].strip())
                    except Exception:
                        meta = {}
                elif txt and ("\t" in txt or "omega_qe" in txt):
                    header = [c.strip() for c in txt.split("\t")]
                continue
            rows.append([float(v) for v in line.split()])
    if not rows:
        raise ValueError(f"No numeric rows found in transmission file '{path}'.")
    arr = np.asarray(rows, dtype=float)
    if arr.ndim != 2 or arr.shape[1] < 4:
        raise ValueError("Transmission file must contain at least 4 numeric columns.")
    per_area = None
    if arr.shape[1] >= 5:
        per_area = arr[:, 4]
    elif header is not None and "transmission_per_m2" in header:
        i = header.index("transmission_per_m2")
        if i < arr.shape[1]:
            per_area = arr[:, i]
    return {"omega_qe": arr[:, 0].copy(), "transmission": arr[:, 3].copy(),
            "transmission_per_m2": None if per_area is None else per_area.copy(),
            "metadata": meta}


def resolve_transmission_source(tcfg: dict, cfg_dir: Path) -> Path:
    """The transmission TSV a ``thermal`` config section names: ``transmission_path``
    (relative to the config's directory), else ``outputs.data`` of the report at
    ``transmission_report_path``."""
    src = tcfg.get("transmission_path", None)
    rep = tcfg.get("transmission_report_path", None)
    if src is None and rep is None:
        raise ValueError("Provide thermal.transmission_path or thermal.transmission_report_path "
                         "for thermal_conductance calculation.")
    if src is not None:
        return resolve_path(cfg_dir, src).resolve()
    rep_path = resolve_path(cfg_dir, rep).resolve()
    data = dict(load_json_config(rep_path).get("outputs", {})).get("data", None)
    if data is None:
        raise ValueError("transmission report does not contain outputs.data.")
    p = Path(data)
    if p.is_absolute():
        return p
    # a relative outputs.data is taken from the working directory when it exists there,
    # else from the report's own directory
    cwd = p.resolve()
    return cwd if cwd.exists() else (rep_path.parent / p).resolve()


def thermal_conductance_run(cfg: dict, *, cfg_dir: Path) -> dict[str, Any]:
    """G(T) on ``thermal.n_temperature_points`` (default 161) temperatures from
    ``temperature_min_k`` (0) to ``temperature_max_k`` (400), from the transmission
    TSV the ``thermal`` section names; also G per m^2 when the TSV has that column."""
    tcfg = dict(cfg.get("thermal", {}))
    path = resolve_transmission_source(tcfg, Path(cfg_dir))
    spec = load_transmission_spectrum(path)
    t_min = float(tcfg.get("temperature_min_k", 0.0))
    t_max = float(tcfg.get("temperature_max_k", 400.0))
    n_t = int(tcfg.get("n_temperature_points", 161))
    if n_t <= 1:
        raise ValueError("thermal.n_temperature_points must be > 1.")
    if t_max < t_min:
        raise ValueError("thermal.temperature_max_k must be >= thermal.temperature_min_k.")
    temps = np.linspace(t_min, t_max, n_t, dtype=float)
    g = thermal_conductance_from_spectrum(spec["omega_qe"], spec["transmission"], temps)
    g_area = None
    if spec["transmission_per_m2"] is not None:
        g_area = thermal_conductance_from_spectrum(spec["omega_qe"], spec["transmission_per_m2"],
                                                   temps)
    return {"temperatures_k": temps, "conductance_w_per_k": np.asarray(g, dtype=float),
            "conductance_per_area_w_m2_k": None if g_area is None else np.asarray(g_area, float),
            "summary": {"transmission_path": str(path), "temperature_min_k": t_min,
                        "temperature_max_k": t_max, "n_temperature_points": n_t,
                        "has_transmission_per_m2": spec["transmission_per_m2"] is not None,
                        "transmission_metadata": dict(spec["metadata"])}}


def save_thermal_conductance_tsv(path, temperatures_k, conductance_w_per_k, *,
                                 conductance_per_area_w_m2_k=None, metadata=None) -> None:
    """``# meta_json:`` (when metadata is non-empty), the column header, then the
    temperature (%.8f) and conductances (%.12e), tab-separated."""
    p = Path(path)
    p.parent.mkdir(parents=True, exist_ok=True)
    with p.open("w", encoding="utf-8") as fh:
        if metadata:
            fh.write(f"# meta_json: {json.dumps(to_builtin(metadata), sort_keys=True)}\n")
        if conductance_per_area_w_m2_k is None:
            fh.write("# temperature_K\tconductance_W_per_K\n")
            for t, gv in zip(temperatures_k, conductance_w_per_k):
                fh.write(f"{float(t):.8f}\t{float(gv):.12e}\n")
        else:
            fh.write("# temperature_K\tconductance_W_per_K\tconductance_per_area_W_per_m2K\n")
            for t, gv, ga in zip(temperatures_k, conductance_w_per_k, conductance_per_area_w_m2_k):
                fh.write(f"{float(t):.8f}\t{float(gv):.12e}\t{float(ga):.12e}\n")


# --------------------------------------------------------------------------
# figures (matplotlib, when it is installed)
# --------------------------------------------------------------------------

def _plot(path: Path, x, y, *, xlabel: str, ylabel: str, title: str, color: str, lw: float) -> bool:
    try:
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
    except Exception:
        return False
    path.parent.mkdir(parents=True, exist_ok=True)
    fig, ax = plt.subplots(figsize=(7.2, 4.6))
    ax.plot(x, y, color=color, lw=lw)
    ax.set_xlabel(xlabel)
    ax.set_ylabel(ylabel)
    ax.set_title(title)
    ax.grid(alpha=0.3)
    fig.tight_layout()
    fig.savefig(path, dpi=220)
    plt.close(fig)
    return True


# --------------------------------------------------------------------------
# the run driver
# --------------------------------------------------------------------------

class _RunLayout:
    """Output directory, study root and manifest of a run (flat or study mode)."""

    def __init__(self, run_cfg: dict, calc: str, stem: str):
        self.name = str(run_cfg.get("name", f"{calc}_{stem}"))
        safe = sanitize_token(self.name)
        self.mode = str(run_cfg.get("output_mode", "flat")).lower()
        self.study_root = self.manifest = self.run_dir_name = None
        self.study = sanitize_token(str(run_cfg.get("study", "bulk_study")))
        self.material = sanitize_token(str(run_cfg.get("material", "material")))
        if self.mode == "flat":
            self.out = Path(run_cfg.get("output_dir", "outputs/bulk_ifc_runs"))
        elif self.mode == "study":
            stamped = bool(run_cfg.get("timestamped_run_dir", True))
            self.run_dir_name = f"{utc_stamp()}_{safe}" if stamped else safe
            self.study_root = (Path(run_cfg.get("studies_root", "studies")) / self.study
                               / self.material)
            self.out = self.study_root / "runs" / self.run_dir_name
            self.manifest = self.study_root / str(run_cfg.get("manifest_jsonl", "manifest.jsonl"))
        else:
            raise ValueError("run.output_mode must be one of: flat, study.")

    def make_dirs(self) -> None:
        if self.study_root is not None:
            for sub in (("inputs", "configs"), ("inputs", "ifc"), ("analysis", "figures"),
                        ("analysis", "tables")):
                (self.study_root.joinpath(*sub)).mkdir(parents=True, exist_ok=True)
        self.out.mkdir(parents=True, exist_ok=True)

    def file(self, run_cfg: dict, key: str, study_name: str) -> Path:
        """``run.<key>`` if set, else ``study_name`` in study mode or
        ``<run name>_<study_name>`` in flat mode."""
        default = study_name if self.mode == "study" else f"{self.name}_{study_name}"
        return self.out / run_cfg.get(key, default)

    def study_block(self) -> dict:
        return {"root": str(self.study_root.resolve()), "study": self.study,
                "material": self.material, "run_dir_name": self.run_dir_name,
                "manifest_jsonl": str(self.manifest.resolve())}


def run_ifc_bulk(config_path, *, precision: str | None = None) -> dict[str, Any]:
    """Run the calculation a study configuration (JSON file) names and write its
    outputs; returns the run report.  ``precision`` ("fp64" or "mixed") overrides the
    transport precision of a transmission run."""
    cfg_path = Path(config_path)
    cfg_dir = cfg_path.parent if cfg_path.parent != Path("") else Path(".")
    cfg = load_json_config(cfg_path)
    run_cfg = dict(cfg.get("run", {}))
    calc = str(run_cfg.get("calculation", "transmission")).lower()
    if calc not in CALCULATIONS:
        raise ValueError("run.calculation must be one of: transmission, dos, dispersion, "
                         "thermal_conductance, fcs.")
    layout = _RunLayout(run_cfg, calc, cfg_path.stem)
    write_plot = bool(run_cfg.get("write_plot", True))
    write_plot_area = bool(run_cfg.get("write_plot_per_area", True))
    write_data = bool(run_cfg.get("write_data", True))
    write_report = bool(run_cfg.get("write_report", True))
    x_unit = str(run_cfg.get("x_unit", "cm-1")).lower()
    if x_unit not in ("cm-1", "thz", "qe"):
        raise ValueError("run.x_unit must be one of: cm-1, thz, qe.")
    if calc not in ("transmission", "thermal_conductance"):
        raise NotImplementedError(f"run.calculation '{calc}' is not a batched transport run; "
                                  "this driver runs transmission and thermal_conductance.")
    if layout.mode == "study":
        layout.make_dirs()

    t0 = time.perf_counter()
    started = utc_now_iso()
    cfg_abs = cfg_path.resolve()
    workspace = Path.cwd()
    prov = {"config_sha256": sha256_file(cfg_abs), "git_head": git_head_sha(workspace),
            "git_dirty": git_dirty(workspace), "hostname": socket.gethostname(),
            "workspace": str(workspace.resolve())}
    outputs: dict[str, str] = {}
    extra: dict[str, Any] = {}
    manifest_extra: dict[str, Any] = {}

    if calc == "thermal_conductance":
        res = thermal_conductance_run(cfg, cfg_dir=cfg_dir)
        layout.out.mkdir(parents=True, exist_ok=True)
        if bool(run_cfg.get("write_input_snapshot", True)):
            save_json(layout.out / "input.json", cfg)
        temps, g, ga = (res["temperatures_k"], res["conductance_w_per_k"],
                        res["conductance_per_area_w_m2_k"])
        if write_data:
            p = layout.file(run_cfg, "data_filename", "thermal_conductance.tsv")
            save_thermal_conductance_tsv(p, temps, g, conductance_per_area_w_m2_k=ga,
                                         metadata={"summary": dict(res["summary"])})
            outputs["data"] = str(p)
        if write_plot:
            p = layout.file(run_cfg, "plot_filename", "thermal_conductance.png")
            if _plot(p, temps, g, xlabel="Temperature (K)", ylabel="Thermal Conductance (W/K)",
                     title=f"Thermal Conductance ({layout.name})", color="tab:red", lw=1.8):
                outputs["plot"] = str(p)
            if ga is not None:
                p = layout.file(run_cfg, "plot_area_filename", "thermal_conductance_per_area.png")
                if _plot(p, temps, ga, xlabel="Temperature (K)",
                         ylabel=r"Thermal Conductance per Area (W/m$^2$/K)",
                         title=f"Thermal Conductance per Area ({layout.name})", color="tab:red",
                         lw=1.8):
                    outputs["plot_per_area"] = str(p)
        extra["thermal"] = dict(res["summary"])
        manifest_extra = {"config_path": str(cfg_abs), "config_sha256": prov["config_sha256"]}
    else:
        from .ifc_transmission import (build_omega_grid, build_transport_model, prepare_kmesh,
                                       spectrum_metadata, transmission_spectrum_from_config)

        ic = dict(cfg.get("ifc", {}))
        ifc_path = resolve_path(cfg_dir, ic.get("center_path", ic.get("path", "si444.fc"))).resolve()
        ifc_sha = sha256_file(ifc_path)
        model = build_transport_model(cfg, cfg_dir)
        omegas, omega_info = build_omega_grid(cfg, model["ifc"])
        kmesh = prepare_kmesh(cfg)
        res = transmission_spectrum_from_config(cfg, model, omegas, kmesh, precision=precision)
        layout.out.mkdir(parents=True, exist_ok=True)
        if bool(run_cfg.get("write_input_snapshot", True)):
            save_json(layout.out / "input.json", cfg)
        vals = np.asarray(res["values"], dtype=float)
        meta, area = spectrum_metadata(cfg, model["ifc"])
        per_m2 = vals / area if area is not None and area > 0.0 else None
        if write_data:
            from .ifc_transmission import save_spectrum_tsv

            p = layout.file(run_cfg, "data_filename", "transmission.tsv")
            save_spectrum_tsv(p, omegas, vals, metadata=meta,
                              extra_columns=None if per_m2 is None
                              else {"transmission_per_m2": per_m2})
            outputs["data"] = str(p)
        if write_plot:
            if x_unit == "thz":
                x, xl = np.asarray(qe_omega_to_thz(omegas), float), "Frequency (THz)"
            elif x_unit == "qe":
                x, xl = np.asarray(omegas, float), "Frequency (QE internal omega unit)"
            else:
                x, xl = np.asarray(qe_omega_to_cm1(omegas), float), r"Frequency (cm$^{-1}$)"
            p = layout.file(run_cfg, "plot_filename", "transmission.png")
            if _plot(p, x, vals, xlabel=xl, ylabel=r"$T(\omega)$",
                     title=f"Transmission ({layout.name})", color="tab:blue", lw=1.6):
                outputs["plot"] = str(p)
            if write_plot_area and per_m2 is not None:
                p = layout.file(run_cfg, "plot_per_area_filename", "transmission_per_area.png")
                if _plot(p, x, per_m2, xlabel=xl, ylabel=r"$T(\omega)$/Area (m$^{-2}$)",
                         title=f"Transmission per Area ({layout.name})", color="tab:blue",
                         lw=1.6):
                    outputs["plot_per_area"] = str(p)
        p = layout.file(run_cfg, "eta_diag_filename", "transmission_eta_diag.json")
        save_json(p, {"eta_records": res.get("eta_records", []),
                      "omega_failures": res.get("omega_failures", []),
                      "n_rejected_total": res.get("n_rejected_total", 0)})
        outputs["eta_diagnostics"] = str(p)
        info = dict(model["info"])
        extra["ifc"] = {"path": str(ifc_path), "sha256": ifc_sha,
                        "reader": str(cfg.get("ifc", {}).get("reader", "qe")),
                        "n_atoms": int(len(model["ifc"].masses)),
                        "atom_symbols": None if model["ifc"].atom_symbols is None
                        else list(model["ifc"].atom_symbols),
                        "asr_residual_max": info.pop("asr_residual_max", None),
                        "n_terms_removed_by_cutoff": info.pop("n_terms_removed_by_cutoff", 0)}
        extra["model"] = info
        extra["kmesh"] = {"dimension": int(kmesh["dim"]), "mode": str(kmesh["mode"]),
                          "low_cm1": float(kmesh["low_cm1"]),
                          "high": {k: v for k, v in kmesh["high"].items() if k != "kpoints"},
                          "low": {k: v for k, v in kmesh["low"].items() if k != "kpoints"}}
        extra["omega"] = omega_info
        extra["solver_summary"] = res.get("summary", {})
        manifest_extra = {"ifc_path": str(ifc_path), "ifc_sha256": ifc_sha}

    finished = utc_now_iso()
    report = {"run": {"name": layout.name, "calculation": calc, "output_mode": layout.mode,
                      "input_config": str(cfg_abs), "started_utc": started,
                      "finished_utc": finished, "runtime_seconds": time.perf_counter() - t0,
                      "run_dir": str(layout.out.resolve())},
              **extra, "outputs": outputs, "provenance": prov}
    if layout.mode == "study":
        report["study"] = layout.study_block()
    if write_report:
        p = layout.file(run_cfg, "report_filename", f"{calc}_report.json")
        save_json(p, report)
        outputs["report"] = str(p)
    if layout.manifest is not None:
        row = {"run_name": layout.name, "calculation": calc, "run_dir": str(layout.out.resolve()),
               "git_head": prov["git_head"], "git_dirty": prov["git_dirty"]}
        # thermal rows carry "created_utc" and the config, transmission rows
        # "timestamp_utc", the report and the IFC file (ifc_bulk.py:2199-2211, :2605-2617)
        if calc == "thermal_conductance":
            row["created_utc"] = finished
        else:
            row["timestamp_utc"] = finished
            row["report"] = outputs.get("report", None)
        row.update(manifest_extra)
        append_manifest_jsonl(layout.manifest, row)
    return report
