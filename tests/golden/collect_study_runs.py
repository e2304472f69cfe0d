"""Copy the reference's own transmission study runs into ``study_runs/`` as fixtures
(run in the development container only; the reference tree is its argument, default
``/root/reference``).

For every run under ``studies/*/*/runs/`` whose calculation is "transmission" with
QE or phonopy force constants, and whose TSV carries the metadata line of the
current ``run_ifc_bulk`` (older runs were written by earlier versions of the
driver), this keeps ``input.json``, ``transmission.tsv`` and
``transmission_eta_diag.json`` under ``study_runs/<run>/`` and the force-constant
input gzip-compressed under ``study_runs/ifc/`` (QE files by name; a phonopy
FORCE_CONSTANTS with its POSCAR under ``ifc/<study>/``).  Data only: no reference
code is copied.
"""

from __future__ import annotations

import gzip
import json
import shutil
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "study_runs"
KEEP = ("input.json", "transmission.tsv", "transmission_eta_diag.json")


def main(ref: str) -> None:
    root = Path(ref) / "studies"
    (OUT / "ifc").mkdir(parents=True, exist_ok=True)
    for inp in sorted(root.glob("*/*/runs/*/input.json")):
        cfg = json.loads(inp.read_text())
        reader = cfg.get("ifc", {}).get("reader")
        if cfg.get("run", {}).get("calculation") != "transmission" or reader not in ("qe", "phonopy"):
            continue
        first = (inp.parent / "transmission.tsv").read_text().splitlines()[0]
        if not first.startswith("# meta_json:") or "lattice_vectors_override_angstrom" not in first:
            continue
        if reader == "qe":
            name = Path(cfg["ifc"]["path"]).name
            pairs = [(next(root.glob(f"*/inputs/ifc/{name}")), OUT / "ifc" / (name + ".gz"))]
        else:
            # the path is relative to the study's input configs (studies/<study>/inputs/configs)
            fc = (inp.parents[3] / "inputs" / "configs" / cfg["ifc"]["path"]).resolve()
            study = fc.parent.name
            (OUT / "ifc" / study).mkdir(exist_ok=True)
            pairs = [(fc.parent / f, OUT / "ifc" / study / (f + ".gz"))
                     for f in (fc.name, "POSCAR")]
        for src, dst in pairs:
            if not dst.exists():
                with open(src, "rb") as f, gzip.GzipFile(dst, "wb", mtime=0) as g:
                    shutil.copyfileobj(f, g)
        run = OUT / inp.parent.name
        run.mkdir(exist_ok=True)
        for k in KEEP:
            shutil.copyfile(inp.parent / k, run / k)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
