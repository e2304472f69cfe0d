"""Host side of the config-driven bulk runs (``negfpy_amd.workflows.run_bulk``): the
TSV reader and writer against the reference's own study files, transmission-source
resolution and the argument errors raised before any GPU work."""

import json
from pathlib import Path

import numpy as np
import pytest

from negfpy_amd.workflows import (load_transmission_spectrum, resolve_transmission_source,
                                  run_ifc_bulk, save_thermal_conductance_tsv,
                                  thermal_conductance_run)
from negfpy_amd.workflows.run_bulk import sanitize_token, to_builtin
from study_runs_util import RUNS, RUNS_DIR, THERMAL_RUNS


def _thermal_ref(name):
    text = (RUNS_DIR / name / "thermal_conductance.tsv").read_text()
    meta = json.loads(text.splitlines()[0].split("meta_json:", 1)[1])
    return text, meta, np.loadtxt(RUNS_DIR / name / "thermal_conductance.tsv")


def test_fixture_inventory():
    assert len(THERMAL_RUNS) == 16 and len(RUNS) == 17
    for name in THERMAL_RUNS:
        src = json.loads((RUNS_DIR / name / "source.json").read_text())["transmission_run"]
        assert src in RUNS


@pytest.mark.parametrize("name", THERMAL_RUNS)
def test_thermal_writer_reproduces_reference_file(tmp_path, name):
    """Written from the reference file's own numbers and metadata, the TSV is the
    reference's file byte for byte (meta line, header, %.8f / %.12e columns)."""
    text, meta, rows = _thermal_ref(name)
    out = tmp_path / "t.tsv"
    per_area = rows[:, 2] if rows.shape[1] > 2 else None
    save_thermal_conductance_tsv(out, rows[:, 0], rows[:, 1], conductance_per_area_w_m2_k=per_area,
                                 metadata=meta)
    assert out.read_text() == text


def test_thermal_writer_without_area_or_metadata(tmp_path):
    out = tmp_path / "t.tsv"
    save_thermal_conductance_tsv(out, np.array([0.0, 2.5]), np.array([0.0, 1.25e-13]))
    assert out.read_text() == ("# temperature_K\tconductance_W_per_K\n"
                               "0.00000000\t0.000000000000e+00\n"
                               "2.50000000\t1.250000000000e-13\n")


@pytest.mark.parametrize("name", RUNS)
def test_transmission_reader_on_reference_files(name):
    path = RUNS_DIR / name / "transmission.tsv"
    spec = load_transmission_spectrum(path)
    rows = np.loadtxt(path)
    np.testing.assert_array_equal(spec["omega_qe"], rows[:, 0])
    np.testing.assert_array_equal(spec["transmission"], rows[:, 3])
    first = path.read_text().splitlines()[0]
    assert spec["metadata"] == json.loads(first.split("meta_json:", 1)[1])
    if rows.shape[1] > 4:
        np.testing.assert_array_equal(spec["transmission_per_m2"], rows[:, 4])
    else:
        assert spec["transmission_per_m2"] is None


def test_transmission_reader_edge_cases(tmp_path):
    p = tmp_path / "a.tsv"
    # broken meta line -> {}, four columns -> no per-area values, blank lines skipped
    p.write_text("# meta_json: {not json\n# omega_qe\tfreq_thz\tfreq_cm^-1\ttransmission\n\n"
                 "1.0 2.0 3.0 0.5\n2.0 2.0 3.0 0.75\n")
    s = load_transmission_spectrum(p)
    assert s["metadata"] == {} and s["transmission_per_m2"] is None
    np.testing.assert_array_equal(s["transmission"], [0.5, 0.75])
    p.write_text("# only comments\n")
    with pytest.raises(ValueError, match="No numeric rows"):
        load_transmission_spectrum(p)
    p.write_text("1.0 2.0 3.0\n")
    with pytest.raises(ValueError, match="at least 4"):
        load_transmission_spectrum(p)


def test_transmission_source_resolution(tmp_path, monkeypatch):
    cfg_dir = tmp_path / "inputs" / "configs"
    run = tmp_path / "runs" / "r1"
    cfg_dir.mkdir(parents=True)
    run.mkdir(parents=True)
    (run / "transmission.tsv").write_text("1 2 3 4\n")
    # transmission_path relative to the config directory
    got = resolve_transmission_source({"transmission_path": "../../runs/r1/transmission.tsv"},
                                      cfg_dir)
    assert got == (run / "transmission.tsv").resolve()
    # report with a relative outputs.data: taken from the report's directory when the
    # working directory has no such file ...
    (run / "transmission_report.json").write_text(json.dumps({"outputs": {"data": "transmission.tsv"}}))
    monkeypatch.chdir(tmp_path)
    got = resolve_transmission_source({"transmission_report_path": "../../runs/r1/transmission_report.json"},
                                      cfg_dir)
    assert got == (run / "transmission.tsv").resolve()
    # ... and from the working directory when it has one
    (tmp_path / "transmission.tsv").write_text("1 2 3 4\n")
    got = resolve_transmission_source({"transmission_report_path": str(run / "transmission_report.json")},
                                      cfg_dir)
    assert got == (tmp_path / "transmission.tsv").resolve()
    (run / "transmission_report.json").write_text(json.dumps({"outputs": {}}))
    with pytest.raises(ValueError, match="outputs.data"):
        resolve_transmission_source({"transmission_report_path": str(run / "transmission_report.json")},
                                    cfg_dir)
    with pytest.raises(ValueError, match="transmission_path"):
        resolve_transmission_source({}, cfg_dir)


def test_thermal_config_errors_before_gpu_work(tmp_path):
    src = RUNS_DIR / RUNS[0] / "transmission.tsv"
    for bad, msg in (({"n_temperature_points": 1}, "n_temperature_points"),
                     ({"temperature_min_k": 10.0, "temperature_max_k": 5.0}, "temperature_max_k")):
        with pytest.raises(ValueError, match=msg):
            thermal_conductance_run({"thermal": {"transmission_path": str(src), **bad}},
                                    cfg_dir=tmp_path)


def _write_cfg(tmp_path, run, **sections):
    p = tmp_path / "cfg.json"
    p.write_text(json.dumps({"run": run, **sections}))
    return p


def test_run_argument_errors(tmp_path):
    th = {"thermal": {"transmission_path": str(RUNS_DIR / RUNS[0] / "transmission.tsv")}}
    with pytest.raises(ValueError, match="run.calculation"):
        run_ifc_bulk(_write_cfg(tmp_path, {"calculation": "bands"}, **th))
    with pytest.raises(ValueError, match="output_mode"):
        run_ifc_bulk(_write_cfg(tmp_path, {"calculation": "thermal_conductance",
                                           "output_mode": "tree"}, **th))
    with pytest.raises(ValueError, match="x_unit"):
        run_ifc_bulk(_write_cfg(tmp_path, {"calculation": "thermal_conductance",
                                           "x_unit": "ev"}, **th))
    for calc in ("dos", "dispersion", "fcs"):
        with pytest.raises(NotImplementedError, match=calc):
            run_ifc_bulk(_write_cfg(tmp_path, {"calculation": calc}))
    (tmp_path / "list.json").write_text("[1, 2]")
    with pytest.raises(ValueError, match="JSON object"):
        run_ifc_bulk(tmp_path / "list.json")


def test_helpers():
    assert sanitize_token("  Si bulk / x (area) ") == "Si_bulk_x_area_"
    assert sanitize_token("   ") == "unnamed"
    assert to_builtin({Path("a"): (np.float64(1.5), np.arange(2), [np.int32(3)])}) == \
        {"a": [1.5, [0, 1], [3]]}
