"""Record the reference's standard-atomic-weight table as a fixture (development
container only): ``atomic_masses.json`` next to this file, symbol -> amu.  The
phonopy reader weights its force constants with these masses, so the table of
``negfpy_amd.modeling.atomic_masses`` must agree with it entry for entry."""

from __future__ import annotations

import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def main(ref_src: str) -> None:
    sys.path.insert(0, ref_src)
    from negfpy.modeling.atomic_masses import ATOMIC_MASS_AMU

    table = {str(k): float(v) for k, v in ATOMIC_MASS_AMU.items()}
    (HERE / "atomic_masses.json").write_text(json.dumps(table, indent=0, sort_keys=True) + "\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src")
