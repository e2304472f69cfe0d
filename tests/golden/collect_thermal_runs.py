"""Copy the reference's thermal-conductance study runs into ``study_runs/`` as fixtures
(run in the development container only; the reference tree is its argument, default
``/root/reference``).

For every run under ``studies/*/*/runs/`` whose calculation is "thermal_conductance"
and whose source transmission run is itself kept under ``study_runs/`` (by
``collect_study_runs.py``), this keeps the run's ``input.json`` and
``thermal_conductance.tsv`` under ``study_runs/<run>/`` and names the source run in
``source.json``.  Data only: no reference code is copied.
"""

from __future__ import annotations

import json
import re
import shutil
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "study_runs"


def main(ref: str) -> None:
    root = Path(ref) / "studies"
    for inp in sorted(root.glob("*/*/runs/*/input.json")):
        cfg = json.loads(inp.read_text())
        if cfg.get("run", {}).get("calculation") != "thermal_conductance":
            continue
        rep = json.loads((inp.parent / "thermal_conductance_report.json").read_text())
        # the report records the transmission TSV it read (a Windows path)
        src = re.split(r"[\\/]", rep["thermal"]["transmission_path"])[-2]
        if not (OUT / src / "transmission.tsv").exists():
            continue
        run = OUT / inp.parent.name
        run.mkdir(exist_ok=True)
        for k in ("input.json", "thermal_conductance.tsv"):
            shutil.copyfile(inp.parent / k, run / k)
        (run / "source.json").write_text(json.dumps({"transmission_run": src}, indent=1) + "\n")
        print(run.name, "<-", src)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
