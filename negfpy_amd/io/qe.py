"""Quantum ESPRESSO q2r.x real-space force-constant files (``*.fc``)
(reference ``src/negfpy/io/readers/qe.py:47-197``).

Layout of the text format: a header (ntyp nat ibrav + cell parameters, with
explicit lattice vectors right after it for ibrav = 0), ntyp species lines
``id 'label' mass``, nat atom lines ``index species x y z``, an ``T``/``F`` flag
followed (for ``T``) by the dielectric tensor and the Born effective charges, the
grid ``nr1 nr2 nr3``, then for every (a, b, atom_i, atom_j) a header line and
nr1*nr2*nr3 lines ``m1 m2 m3 value``.  Grid indices m are 1-based FFT indices;
they map to signed cell shifts with the upper half (including the Nyquist index of
even grids) on the negative side.
"""

from __future__ import annotations

import re
from pathlib import Path

import numpy as np

from negfpy_amd.modeling.schema import IFCData, IFCTerm, validate_ifc_data

_SPECIES = re.compile(r"^\s*(\d+)\s+'([^']+)'\s+([Ee0-9+\-.]+)\s*$")


def _shift(m: int, nr: int) -> int:
    s = int(m) - 1
    return s - nr if (nr > 1 and s >= nr // 2) else s


class _Lines:
    """Cursor over the non-empty lines of the file."""

    def __init__(self, path: Path):
        self.lines = [ln.strip() for ln in path.read_text(encoding="utf-8").splitlines()]
        self.i = 0

    def peek(self) -> str:
        j = self.i
        while j < len(self.lines) and not self.lines[j]:
            j += 1
        if j >= len(self.lines):
            raise ValueError("Unexpected end of file while parsing QE IFC file.")
        return self.lines[j]

    def next(self) -> str:
        while self.i < len(self.lines) and not self.lines[self.i]:
            self.i += 1
        if self.i >= len(self.lines):
            raise ValueError("Unexpected end of file while parsing QE IFC file.")
        line = self.lines[self.i]
        self.i += 1
        return line


def _three_floats(line: str):
    tok = line.split()
    if len(tok) < 3:
        return None
    try:
        return [float(tok[0]), float(tok[1]), float(tok[2])]
    except ValueError:
        return None


def read_qe_q2r_ifc(source) -> IFCData:
    path = Path(source)
    cur = _Lines(path)
    head = cur.next().split()
    if len(head) < 3:
        raise ValueError("Invalid QE IFC header line.")
    ntyp, nat, ibrav = int(head[0]), int(head[1]), int(head[2])

    lattice = None
    if ibrav == 0 and _SPECIES.match(cur.peek()) is None and _three_floats(cur.peek()) is not None:
        rows = [_three_floats(cur.next())]
        for _ in range(2):
            r = _three_floats(cur.next())
            if r is None:
                raise ValueError("Invalid explicit lattice-vector block in QE IFC file.")
            rows.append(r)
        lattice = np.asarray(rows, dtype=float)

    labels, masses_by_species = {}, {}
    for _ in range(ntyp):
        line = cur.next()
        m = _SPECIES.match(line)
        if m is None:
            raise ValueError(f"Invalid species line in QE IFC file: '{line}'")
        labels[int(m.group(1))] = m.group(2).strip()
        masses_by_species[int(m.group(1))] = float(m.group(3))

    species, positions = [], []
    for _ in range(nat):
        line = cur.next()
        tok = line.split()
        if len(tok) < 5:
            raise ValueError(f"Invalid atomic-position line in QE IFC file: '{line}'")
        species.append(int(tok[1]))
        positions.append([float(tok[2]), float(tok[3]), float(tok[4])])

    born = None
    flag = cur.peek()
    if flag in ("T", "F"):
        cur.next()
        if flag == "T":
            eps = [[float(x) for x in cur.next().split()[:3]] for _ in range(3)]
            born = np.zeros((nat, 3, 3))
            for ia in range(nat):
                cur.next()  # atom index
                for a in range(3):
                    born[ia, a, :] = [float(x) for x in cur.next().split()[:3]]
            lattice = np.asarray(eps, dtype=float)

    grid = cur.next().split()
    if len(grid) < 3:
        raise ValueError("Missing real-space IFC grid dimensions in QE IFC file.")
    nr = (int(grid[0]), int(grid[1]), int(grid[2]))
    npts = nr[0] * nr[1] * nr[2]

    ndof = 3 * nat
    blocks: dict = {}
    for _ in range(nat * nat * 9):
        line = cur.next()
        h = line.split()
        if len(h) < 4:
            raise ValueError(f"Invalid IFC block header in QE IFC file: '{line}'")
        a, b, ia, ib = int(h[0]) - 1, int(h[1]) - 1, int(h[2]) - 1, int(h[3]) - 1
        if not (0 <= ia < nat and 0 <= ib < nat and 0 <= a < 3 and 0 <= b < 3):
            raise ValueError(f"Out-of-range IFC block indices in line: '{line}'")
        for _ in range(npts):
            row = cur.next()
            tok = row.split()
            if len(tok) < 4:
                raise ValueError(f"Invalid IFC value line in QE IFC file: '{row}'")
            key = (_shift(int(tok[0]), nr[0]), _shift(int(tok[1]), nr[1]),
                   _shift(int(tok[2]), nr[2]))
            blk = blocks.get(key)
            if blk is None:
                blk = blocks[key] = np.zeros((ndof, ndof), dtype=np.complex128)
            blk[3 * ia + a, 3 * ib + b] = float(tok[3].replace("D", "E"))

    ifc = IFCData(
        masses=np.asarray([masses_by_species[s] for s in species], dtype=float),
        dof_per_atom=3,
        terms=tuple(IFCTerm(dx=k[0], dy=k[1], dz=k[2], block=v) for k, v in sorted(blocks.items())),
        units="sqrt(Ry/(me*bohr^2))",
        metadata={"source_format": "qe_q2r_fc", "path": str(path), "ntyp": ntyp, "nat": nat,
                  "ibrav": ibrav, "nr": nr, "force_constant_unit": "Ry/bohr^2",
                  "mass_unit": "2*electron_mass (QE Ry mass unit)",
                  "omega_internal_unit": "sqrt(Ry/(2*me*bohr^2))", "born_charges": born},
        lattice_vectors=lattice,
        atom_positions=np.asarray(positions, dtype=float),
        atom_symbols=tuple(labels[s] for s in species),
        index_convention="qe_q2r_fft",
    )
    validate_ifc_data(ifc)
    return ifc
