"""Model conditioning between the IFC file and the transport blocks
(reference ``src/negfpy/workflows/ifc_bulk.py:260-300, :484-531, :745-905``)."""

from __future__ import annotations

from dataclasses import replace

import numpy as np

from .schema import IFCData

_PERM = {1: (0, 1, 2), 2: (1, 0, 2), 3: (2, 0, 1)}


def _with_terms(ifc: IFCData, terms, metadata) -> IFCData:
    return IFCData(masses=np.asarray(ifc.masses, dtype=float), dof_per_atom=int(ifc.dof_per_atom),
                   terms=tuple(terms), units=ifc.units, metadata=metadata,
                   lattice_vectors=ifc.lattice_vectors, atom_positions=ifc.atom_positions,
                   atom_symbols=ifc.atom_symbols, index_convention=ifc.index_convention)


def reorient_for_transport(ifc: IFCData, direction: int) -> IFCData:
    """Relabel lattice directions so transport along a1/a2/a3 runs along +x."""
    if int(direction) not in _PERM:
        raise ValueError("model.transport_direction must be one of: 1, 2, 3.")
    perm = _PERM[int(direction)]
    if perm == (0, 1, 2):
        return ifc
    terms = []
    for t in ifc.terms:
        v = (int(t.dx), int(t.dy), int(t.dz))
        terms.append(replace(t, dx=int(v[perm[0]]), dy=int(v[perm[1]]), dz=int(v[perm[2]])))
    meta = dict(ifc.metadata)
    nr = meta.get("nr")
    if isinstance(nr, (list, tuple)) and len(nr) == 3:
        meta["nr"] = (int(nr[perm[0]]), int(nr[perm[1]]), int(nr[perm[2]]))
    meta["transport_direction"] = int(direction)
    meta["transport_permutation"] = perm
    return _with_terms(ifc, terms, meta)


def transverse_cutoff(ifc: IFCData, *, dy_cutoff=None, dz_cutoff=None):
    """Drop terms with |dy| > dy_cutoff or |dz| > dz_cutoff; (ifc, number removed)."""
    if dy_cutoff is None and dz_cutoff is None:
        return ifc, 0
    keep = [t for t in ifc.terms
            if not ((dy_cutoff is not None and abs(int(t.dy)) > dy_cutoff)
                    or (dz_cutoff is not None and abs(int(t.dz)) > dz_cutoff))]
    removed = len(ifc.terms) - len(keep)
    if not keep:
        raise ValueError("All IFC terms were removed by dy/dz cutoffs.")
    if not any(t.dx == 0 and t.dy == 0 and t.dz == 0 for t in keep):
        raise ValueError("dy/dz cutoff removed IFC self-term (0,0,0), invalid model.")
    meta = dict(ifc.metadata)
    meta.update(dy_cutoff=dy_cutoff, dz_cutoff=dz_cutoff, n_terms_removed_by_cutoff=int(removed))
    return _with_terms(ifc, keep, meta), removed


def apply_mass_mode(params, mode):
    m = str(mode).lower().replace("-", "_")
    if m == "ifc":
        return params, m
    if m in ("unit", "ones", "legacy_unit"):
        return replace(params, masses=np.ones_like(params.masses, dtype=float)), "unit"
    raise ValueError("model.mass_mode must be one of: ifc, unit.")


def _even_half(n):
    return (n // 2) if (n > 1 and n % 2 == 0) else None


def _transverse_nr(meta):
    nr = meta.get("nr") if isinstance(meta, dict) else None
    if nr is None or len(nr) != 3:
        return None
    try:
        return int(nr[1]), int(nr[2])
    except (TypeError, ValueError):
        return None


def drop_nyquist_transverse_terms(params, ifc_metadata):
    """Remove the terms at the negative transverse Nyquist shifts of even grids."""
    none = {"dropped_dy": None, "dropped_dz": None, "n_dropped_fc00": 0, "n_dropped_fc01": 0,
            "n_dropped_fc10": 0}
    nr = _transverse_nr(ifc_metadata)
    if nr is None:
        return params, none
    ddy = -nr[0] // 2 if _even_half(nr[0]) is not None else None
    ddz = -nr[1] // 2 if _even_half(nr[1]) is not None else None
    if ddy is None and ddz is None:
        return params, none

    def keep(k):
        return not ((ddy is not None and int(k[0]) == ddy) or (ddz is not None and int(k[1]) == ddz))

    f00 = {k: v for k, v in params.fc00_terms.items() if keep(k)}
    f01 = {k: v for k, v in params.fc01_terms.items() if keep(k)}
    f10 = None if params.fc10_terms is None else {k: v for k, v in params.fc10_terms.items()
                                                   if keep(k)}
    info = {"dropped_dy": ddy, "dropped_dz": ddz,
            "n_dropped_fc00": int(len(params.fc00_terms) - len(f00)),
            "n_dropped_fc01": int(len(params.fc01_terms) - len(f01)),
            "n_dropped_fc10": int(0 if f10 is None else len(params.fc10_terms) - len(f10))}
    return replace(params, fc00_terms=f00, fc01_terms=f01, fc10_terms=f10), info


def enforce_transverse_pm_symmetry(params, ifc_metadata=None):
    """Complete the +/- partner of every term at a transverse Nyquist shift of an
    even grid: Phi00(-s) = Phi00(s)^H (Hermitised when s = -s), Phi10(-s) = Phi01(s)^H
    and Phi01(-s) = Phi10(s)^H.  Other terms are untouched."""
    nr = _transverse_nr(ifc_metadata)
    if nr is None:
        return params, {"enabled": False, "reason": "missing_nr_metadata"}
    ny, nz = _even_half(nr[0]), _even_half(nr[1])
    if ny is None and nz is None:
        return params, {"enabled": False, "reason": "no_even_transverse_grid"}

    def nyq(k):
        return (ny is not None and abs(int(k[0])) == ny) or (nz is not None and abs(int(k[1])) == nz)

    def flip(k):
        return (-int(k[0]), -int(k[1]))

    f00 = {k: np.asarray(v, dtype=np.complex128).copy() for k, v in params.fc00_terms.items()}
    added00 = herm00 = 0
    for k in list(f00.keys()):
        if not nyq(k):
            continue
        kp = flip(k)
        if k == kp:
            f00[k] = 0.5 * (f00[k] + f00[k].conj().T)
            herm00 += 1
        elif kp not in f00:
            f00[kp] = f00[k].conj().T
            added00 += 1
    info00 = {"n_terms_in": int(len(params.fc00_terms)), "n_terms_out": int(len(f00)),
              "n_pairs_added": int(added00), "n_self_hermitianized": int(herm00)}

    f01 = {k: np.asarray(v, dtype=np.complex128).copy() for k, v in params.fc01_terms.items()}
    if params.fc10_terms is None:
        f10 = None
        info01 = {"n_fc01_in": int(len(params.fc01_terms)), "n_fc01_out": int(len(f01)),
                  "n_fc10_in": 0, "n_fc10_out": 0, "n_pairs_added_to_fc01": 0,
                  "n_pairs_added_to_fc10": 0}
    else:
        f10 = {k: np.asarray(v, dtype=np.complex128).copy() for k, v in params.fc10_terms.items()}
        add10 = add01 = 0
        for k, b in list(f01.items()):
            if nyq(k) and flip(k) not in f10:
                f10[flip(k)] = b.conj().T
                add10 += 1
        for k, b in list(f10.items()):
            if nyq(k) and flip(k) not in f01:
                f01[flip(k)] = b.conj().T
                add01 += 1
        info01 = {"n_fc01_in": int(len(params.fc01_terms)), "n_fc01_out": int(len(f01)),
                  "n_fc10_in": int(len(params.fc10_terms)), "n_fc10_out": int(len(f10)),
                  "n_pairs_added_to_fc01": int(add01), "n_pairs_added_to_fc10": int(add10)}
    info = {"enabled": True, "nyquist_abs_dy": ny, "nyquist_abs_dz": nz, "fc00": info00,
            "fc01_fc10": info01}
    return replace(params, fc00_terms=f00, fc01_terms=f01, fc10_terms=f10), info


def apply_transverse_stability_controls(params, *, ifc_metadata, drop_nyquist: bool):
    """Optional Nyquist drop, then the (always applied) Nyquist +/- completion."""
    nyq_info = {"dropped_dy": None, "dropped_dz": None, "n_dropped_fc00": 0,
                "n_dropped_fc01": 0, "n_dropped_fc10": 0}
    out = params
    if drop_nyquist:
        out, nyq_info = drop_nyquist_transverse_terms(out, ifc_metadata)
    out, pm_info = enforce_transverse_pm_symmetry(out, ifc_metadata)
    return out, nyq_info, pm_info
