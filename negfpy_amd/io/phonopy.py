"""phonopy force constants: a supercell ``FORCE_CONSTANTS`` text file plus its
``SPOSCAR``/``POSCAR``, or a normalized IFC payload (dict or JSON file)
(reference ``src/negfpy/io/readers/phonopy.py``).

The supercell is folded onto its primitive cell: the repeat along each lattice
vector is the largest n dividing the atom count for which shifting every atom by
1/n maps each species onto itself; atoms of the home cell, ordered by species and
fractional position, form the basis; every other atom is matched to the nearest
basis atom of its species.  The 3x3 block of each supercell pair (i, j) goes to
the term of the cell shift t_j - t_i (modulo the repeats, as a signed FFT index)
at the (basis_i, basis_j) position, and blocks that several pairs reach are
averaged.  Units: eV/Angstrom^2 -> Ry/bohr^2 and amu -> the QE mass unit
(2 electron masses), so the model shares the QE reader's frequency units.
"""

from __future__ import annotations

import json
from pathlib import Path
from typing import Any

import numpy as np

from negfpy_amd.modeling.atomic_masses import get_atomic_mass_amu
from negfpy_amd.modeling.schema import IFCData, IFCTerm, validate_ifc_data
from negfpy_amd.modeling.units import BOHR_M, EV_J, QE_RY_MASS_UNIT_KG, RYDBERG_J

AMU_KG = 1.66053906660e-27


def fc_scale_to_qe() -> float:
    """eV/Angstrom^2 -> Ry/bohr^2."""
    a2b = 1.0e-10 / BOHR_M
    return float(EV_J / RYDBERG_J) / (a2b * a2b)


def mass_scale_to_qe() -> float:
    """amu -> QE Ry mass unit."""
    return float(AMU_KG / QE_RY_MASS_UNIT_KG)


def parse_poscar(path) -> dict[str, Any]:
    lines = Path(path).read_text(encoding="utf-8").splitlines()
    if len(lines) < 8:
        raise ValueError(f"POSCAR seems too short: '{path}'")
    scale = float(lines[1].split()[0])
    lv = np.array([[float(x) for x in lines[2 + r].split()[:3]] for r in range(3)]) * scale
    tok = lines[5].split()
    if all(t.lstrip("+-").isdigit() for t in tok):
        counts = [int(x) for x in tok]
        species = [f"X{j + 1}" for j in range(len(counts))]
        i = 6
    else:
        species = tok
        counts = [int(x) for x in lines[6].split()]
        i = 7
    if i < len(lines) and lines[i].strip().lower().startswith("s"):
        i += 1   # selective dynamics
    if i >= len(lines):
        raise ValueError(f"POSCAR missing coordinate mode line: '{path}'")
    mode = lines[i].strip().lower()
    i += 1
    nat = sum(counts)
    if i + nat > len(lines):
        raise ValueError(f"POSCAR atom coordinate section is incomplete: '{path}'")
    xyz = []
    for j in range(nat):
        t = lines[i + j].split()
        if len(t) < 3:
            raise ValueError(f"Invalid POSCAR coordinate line: '{lines[i + j]}'")
        xyz.append([float(t[0]), float(t[1]), float(t[2])])
    pos = np.array(xyz, dtype=float)
    if mode.startswith("c") or mode.startswith("k"):
        pos = pos @ np.linalg.inv(lv)
    elif not mode.startswith("d"):
        raise ValueError(f"Unsupported POSCAR coordinate mode '{lines[i - 1].strip()}'.")
    symbols = tuple(s for s, c in zip(species, counts) for _ in range(int(c)))
    return {"lattice_vectors": lv, "atom_positions": np.mod(pos, 1.0), "atom_symbols": symbols,
            "species": tuple(species), "counts": tuple(int(c) for c in counts)}


def parse_force_constants(path) -> np.ndarray:
    """(nat, nat, 3, 3) from a phonopy FORCE_CONSTANTS text file."""
    rows = [ln.split() for ln in Path(path).read_text(encoding="utf-8").splitlines() if ln.strip()]
    if not rows:
        raise ValueError("Unexpected end of file while parsing phonopy FORCE_CONSTANTS.")
    head = rows[0]
    nat = int(head[0])
    if len(head) >= 2 and int(head[1]) != nat:
        raise ValueError("FORCE_CONSTANTS first line must have matching nat values.")
    if len(rows) < 1 + 4 * nat * nat:
        raise ValueError("Unexpected end of file while parsing phonopy FORCE_CONSTANTS.")
    fc = np.zeros((nat, nat, 3, 3), dtype=float)
    r = 1
    for _ in range(nat * nat):
        pair = rows[r]
        if len(pair) < 2:
            raise ValueError(f"Invalid FORCE_CONSTANTS pair line: '{' '.join(pair)}'")
        ia, ja = int(pair[0]) - 1, int(pair[1]) - 1
        if not (0 <= ia < nat and 0 <= ja < nat):
            raise ValueError(f"FORCE_CONSTANTS pair indices out of range: '{' '.join(pair)}'")
        for a in range(3):
            v = rows[r + 1 + a]
            if len(v) < 3:
                raise ValueError(f"Invalid FORCE_CONSTANTS matrix row: '{' '.join(v)}'")
            fc[ia, ja, a, :] = (float(v[0]), float(v[1]), float(v[2]))
        r += 4
    return fc


def _frac_key(p, tol=1e-6):
    return tuple(int(np.round(float(x) / tol)) for x in np.mod(p, 1.0))


def _repeat_along(frac, symbols, axis, n):
    if n <= 1:
        return True
    d = np.zeros(3)
    d[axis] = 1.0 / n
    sym = np.array(symbols)
    for s in set(symbols):
        pts = frac[sym == s]
        if {_frac_key(p) for p in pts} != {_frac_key(p + d) for p in pts}:
            return False
    return True


def infer_supercell_repeats(frac, symbols) -> tuple[int, int, int]:
    nat = len(symbols)
    rep = []
    for axis in range(3):
        best = 1
        for n in range(2, max(1, nat) + 1):
            if nat % n == 0 and _repeat_along(frac, symbols, axis, n):
                best = n
        rep.append(best)
    if nat % (rep[0] * rep[1] * rep[2]) != 0:
        raise ValueError("Failed to infer a valid supercell repeat vector from POSCAR.")
    return rep[0], rep[1], rep[2]


def _signed(raw, n):
    return np.where((n > 1) & (raw >= n // 2), raw - n, raw)


def read_phonopy_force_constants(fc_path, poscar_path=None, *, supercell=None) -> IFCData:
    fc_path = Path(fc_path).resolve()
    if poscar_path is None:
        cands = [fc_path.parent / "SPOSCAR", fc_path.parent / "POSCAR"]
        found = [c for c in cands if c.exists()]
        if not found:
            raise FileNotFoundError(f"Could not find POSCAR/SPOSCAR next to '{fc_path}'. "
                                    "Provide a POSCAR in the same folder as FORCE_CONSTANTS.")
        poscar_path = found[0]
    poscar_path = Path(poscar_path).resolve()
    pos = parse_poscar(poscar_path)
    frac = pos["atom_positions"]
    symbols = pos["atom_symbols"]
    nat = len(symbols)
    fc = parse_force_constants(fc_path)
    if fc.shape[0] != nat:
        raise ValueError(f"Mismatch between POSCAR atoms ({nat}) and FORCE_CONSTANTS nat "
                         f"({fc.shape[0]}).")
    if supercell is None:
        reps = infer_supercell_repeats(frac, symbols)
    else:
        reps = tuple(int(v) for v in supercell)
        if min(reps) <= 0:
            raise ValueError("supercell repeats must be positive integers.")
    nmult = reps[0] * reps[1] * reps[2]
    if nat % nmult != 0:
        raise ValueError(f"nat_super={nat} is not divisible by inferred/provided supercell size "
                         f"{reps[0]}x{reps[1]}x{reps[2]}.")
    nprim = nat // nmult
    nvec = np.array(reps, dtype=float)
    q = frac * nvec[None, :]
    cell = np.floor(q + 1e-8).astype(int)
    rem = np.mod(q - cell, 1.0)
    cell %= np.array(reps)[None, :]
    home = np.flatnonzero(np.all(cell == 0, axis=1))
    if home.size != nprim:
        raise ValueError("Could not identify primitive home cell atoms from supercell mapping. "
                         f"Expected {nprim}, found {int(home.size)}.")
    order = sorted(home.tolist(), key=lambda a: (str(symbols[a]), float(rem[a, 0]),
                                                 float(rem[a, 1]), float(rem[a, 2])))
    bfrac = rem[order]
    bsym = [str(symbols[a]) for a in order]
    basis = np.empty(nat, dtype=int)
    for a in range(nat):
        best, bd = -1, None
        for b in range(nprim):
            if bsym[b] != symbols[a]:
                continue
            d = float(np.linalg.norm(np.mod(rem[a] - bfrac[b] + 0.5, 1.0) - 0.5))
            if bd is None or d < bd:
                best, bd = b, d
        if best < 0 or bd > 5e-5:
            raise ValueError("Failed to map supercell atom to primitive basis with geometric tolerance.")
        basis[a] = best
    ndof = 3 * nprim
    # cell shift of every ordered pair (i outer, j inner), as signed FFT indices
    rp = np.array(reps)
    raw = np.mod(cell[None, :, :] - cell[:, None, :], rp[None, None, :])
    key = _signed(raw, rp[None, None, :]).reshape(nat * nat, 3)
    uniq, kid = np.unique(key, axis=0, return_inverse=True)
    kid = kid.reshape(-1)
    bi = np.repeat(basis, nat)
    bj = np.tile(basis, nat)
    acc = np.zeros((len(uniq), nprim, nprim, 3, 3), dtype=np.complex128)
    cnt = np.zeros((len(uniq), nprim, nprim), dtype=float)
    np.add.at(acc, (kid, bi, bj), fc.reshape(nat * nat, 3, 3).astype(np.complex128))
    np.add.at(cnt, (kid, bi, bj), 1.0)
    s_fc = fc_scale_to_qe()
    terms = []
    for u in range(len(uniq)):        # np.unique sorts the keys lexicographically
        blk = acc[u].transpose(0, 2, 1, 3).reshape(ndof, ndof)
        c = np.repeat(np.repeat(cnt[u], 3, axis=0), 3, axis=1)
        val = np.zeros_like(blk)
        np.divide(blk, c, out=val, where=c > 0.0)
        terms.append(IFCTerm(dx=int(uniq[u, 0]), dy=int(uniq[u, 1]), dz=int(uniq[u, 2]),
                             block=val * s_fc))
    s_m = mass_scale_to_qe()
    masses = np.array([get_atomic_mass_amu(s) for s in bsym], dtype=float) * s_m
    ifc = IFCData(
        masses=masses, dof_per_atom=3, terms=tuple(terms), units="sqrt(Ry/(me*bohr^2))",
        metadata={"source_format": "phonopy_force_constants", "force_constants_path": str(fc_path),
                  "poscar_path": str(poscar_path), "nat_supercell": nat,
                  "nat_primitive_inferred": nprim, "nr": tuple(reps), "supercell_repeats": tuple(reps),
                  "force_constant_unit_source": "eV/Angstrom^2 (phonopy default)",
                  "mass_unit_source": "amu", "force_constant_unit_converted": "Ry/bohr^2",
                  "mass_unit_converted": "2*electron_mass (QE Ry mass unit)",
                  "force_constant_scale_to_qe": s_fc, "mass_scale_to_qe": s_m,
                  "omega_internal_unit": "sqrt(Ry/(2*me*bohr^2))"},
        lattice_vectors=pos["lattice_vectors"] / nvec[:, None], atom_positions=bfrac,
        atom_symbols=tuple(bsym), index_convention="phonopy_fft_signed")
    validate_ifc_data(ifc)
    return ifc


def _from_payload(p: dict) -> IFCData:
    terms = []
    for it in p["terms"]:
        dx, dy, dz = it["translation"] if "translation" in it else (it["dx"], it["dy"], it["dz"])
        terms.append(IFCTerm(dx=int(dx), dy=int(dy), dz=int(dz),
                             block=np.asarray(it["block"], dtype=np.complex128)))
    lv = p.get("lattice_vectors")
    ap = p.get("atom_positions")
    sy = p.get("atom_symbols")
    ifc = IFCData(masses=np.asarray(p["masses"], dtype=float), dof_per_atom=int(p.get("dof_per_atom", 3)),
                  terms=tuple(terms), units=str(p.get("units", "unknown")),
                  metadata=dict(p.get("metadata", {})),
                  lattice_vectors=None if lv is None else np.asarray(lv, dtype=float),
                  atom_positions=None if ap is None else np.asarray(ap, dtype=float),
                  atom_symbols=None if sy is None else tuple(sy),
                  index_convention=str(p.get("index_convention", "layer-major")))
    validate_ifc_data(ifc)
    return ifc


def _supercell_arg(raw):
    if raw is None:
        return None
    if not isinstance(raw, (list, tuple)) or len(raw) != 3:
        raise ValueError("source['supercell'] must be a 3-element sequence when provided.")
    return tuple(int(v) for v in raw)


def read_phonopy_ifc(source: Any) -> IFCData:
    """A FORCE_CONSTANTS path (POSCAR/SPOSCAR beside it), a JSON file or dict holding
    a normalized IFC payload (``masses``, ``terms``, ...) or the keys
    ``force_constants_path``/``path`` [, ``poscar_path``, ``supercell``]."""
    if isinstance(source, (str, Path)):
        p = Path(source)
        if p.is_file() and p.suffix.lower() == ".json":
            payload = json.loads(p.read_text(encoding="utf-8"))
            if "terms" in payload and "masses" in payload:
                return _from_payload(payload)
            raw = payload.get("force_constants_path", payload.get("path"))
            if raw is None:
                raise ValueError("Unsupported phonopy JSON payload. Expected normalized IFC keys "
                                 "('masses','terms') or force_constants_path/path.")
            fcp = Path(raw) if Path(raw).is_absolute() else (p.parent / raw).resolve()
            pp = payload.get("poscar_path")
            if pp is not None and not Path(pp).is_absolute():
                pp = (p.parent / pp).resolve()
            return read_phonopy_force_constants(fcp, pp, supercell=_supercell_arg(payload.get("supercell")))
        return read_phonopy_force_constants(p)
    if isinstance(source, dict):
        if "terms" in source and "masses" in source:
            return _from_payload(source)
        if "force_constants_path" in source or "path" in source:
            return read_phonopy_force_constants(
                source.get("force_constants_path", source.get("path")), source.get("poscar_path"),
                supercell=_supercell_arg(source.get("supercell")))
        raise ValueError("Unsupported phonopy source dict. Expected normalized IFC payload or keys "
                         "'force_constants_path'/'path'.")
    raise TypeError("Unsupported source type for phonopy reader.")
