"""Write reference-made vectors for the IFC study pipeline (development container only).

Imports the reference package ``negfpy`` from its source tree (first argument,
default ``/root/reference/src``) and records, for the two study configurations
held in ``study_transmission/`` (graphene, 1D k mesh; silicon along x, 2D k mesh),
the intermediate stages of ``negfpy.workflows.ifc_bulk.run_ifc_bulk`` whose
conventions the end-to-end spectra alone do not isolate:

* the conditioned force-constant terms after reading, reorienting, building the
  principal layer, the mass mode and the transverse controls (ifc_bulk.py:2244-2287);
* the interface-contact leads and device-to-lead couplings at a few k points
  (``_build_interface_contacts``), and the device blocks there;
* the k meshes (``_prepare_kmesh``) and the omega grid, including the eigenvalue
  estimate of the maximum frequency (``_estimate_omega_max_from_ifc``);
* per-k transmissions through the explicit contacts at fixed eta;
* the phonopy reader's primitive-cell terms, masses and cell for the TaN study.

Output (data only): ``ifc_vectors.npz`` next to this file.
"""

from __future__ import annotations

import json
import os
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
STUDY = HERE / "study_transmission"
CASES = (("graphene", "graphene.fc"), ("silicon_x", "si444.fc"))


def _terms(prefix, name, terms, out):
    keys = sorted(terms)
    out[f"{prefix}_{name}_keys"] = np.array(keys, dtype=np.int64).reshape(len(keys), -1)
    out[f"{prefix}_{name}_blocks"] = np.stack([np.asarray(terms[k]) for k in keys])


def main(ref_src: str) -> None:
    sys.path.insert(0, ref_src)
    import negfpy.core as rc
    import negfpy.workflows.ifc_bulk as wb
    from negfpy.io import read_ifc
    from negfpy.modeling import BuildConfig
    from negfpy.modeling.builders import build_material_kspace_params
    from negfpy.models import material_kspace_device

    out = {}
    for tag, fc in CASES:
        cfg = json.loads((STUDY / f"{tag}_input.json").read_text())
        mc = cfg["model"]
        ifc = read_ifc(STUDY / fc, reader="qe")
        ifc = wb._reorient_ifc_for_transport_direction(ifc, direction=int(mc.get("transport_direction", 1)))
        ifc, _ = wb._apply_transverse_cutoff(ifc, dy_cutoff=None, dz_cutoff=None)
        bc = BuildConfig(onsite_pinning=float(mc.get("onsite_pinning", 0.0)),
                         principal_layer_size=int(mc["principal_layer_size"]),
                         auto_principal_layer_enlargement=True, infer_fc01_from_negative_dx=True,
                         enforce_hermitian_fc00=True, nyquist_split_half=False, dtype="complex128")
        params = build_material_kspace_params(ifc=ifc, config=bc)
        params, _ = wb._apply_mass_mode(params, "ifc")
        params, _, _ = wb._apply_transverse_stability_controls(
            params, ifc_metadata=dict(ifc.metadata), drop_nyquist=False)
        out[f"{tag}_masses"] = np.asarray(params.masses, dtype=float)
        _terms(tag, "fc00", params.fc00_terms, out)
        _terms(tag, "fc01", params.fc01_terms, out)
        if params.fc10_terms is not None:
            _terms(tag, "fc10", params.fc10_terms, out)
        lead_l, lead_r, v_l, v_r = wb._build_interface_contacts(params)
        n_layers = 6
        dev = material_kspace_device(n_layers=n_layers, params=params)
        km = wb._prepare_kmesh(cfg)
        out[f"{tag}_kmesh_high"] = np.array(km["high"]["kpoints"], dtype=float)
        out[f"{tag}_kmesh_low"] = np.array(km["low"]["kpoints"], dtype=float)
        hi = km["high"]["kpoints"]
        sel = [hi[0], hi[len(hi) // 3], hi[len(hi) // 2 + 1]]
        out[f"{tag}_k"] = np.array(sel, dtype=float)
        out[f"{tag}_lead_l"] = np.stack([np.stack(lead_l.blocks_builder(k)) for k in sel])
        out[f"{tag}_lead_r"] = np.stack([np.stack(lead_r.blocks_builder(k)) for k in sel])
        out[f"{tag}_v_l"] = np.stack([v_l(k) for k in sel])
        out[f"{tag}_v_r"] = np.stack([v_r(k) for k in sel])
        out[f"{tag}_onsite"] = np.stack([np.stack(dev.device_builder(k).onsite_blocks) for k in sel])
        out[f"{tag}_coupling"] = np.stack([np.stack(dev.device_builder(k).coupling_blocks)
                                           for k in sel])
        wmax = float(np.asarray(wb._build_omega_grid(cfg, ifc)[0]).max())
        ws = np.array([0.13, 0.41, 0.77]) * wmax
        out[f"{tag}_t_omegas"] = ws
        out[f"{tag}_t_layers"] = np.array(n_layers)
        out[f"{tag}_t"] = np.array([[rc.transmission(w, dev, lead_l, lead_r, kpar=k, eta=1e-6,
                                                     device_to_lead_left=v_l,
                                                     device_to_lead_right=v_r)
                                     for w in ws] for k in sel])
        grid, info = wb._build_omega_grid(cfg, ifc)
        out[f"{tag}_omega_grid"] = np.asarray(grid, dtype=float)
    # the phonopy reader on the TaN study's supercell force constants
    tan = read_ifc(Path(ref_src).parent / "studies" / "TaN" / "FORCE_CONSTANTS", reader="phonopy")
    out["tan_masses"] = np.asarray(tan.masses, dtype=float)
    out["tan_keys"] = np.array([[t.dx, t.dy, t.dz] for t in tan.terms], dtype=np.int64)
    out["tan_blocks"] = np.stack([np.asarray(t.block) for t in tan.terms])
    out["tan_lattice"] = np.asarray(tan.lattice_vectors, dtype=float)
    out["tan_positions"] = np.asarray(tan.atom_positions, dtype=float)
    # the eigenvalue estimate of the maximum frequency on a small k grid
    ifc = read_ifc(STUDY / "si444.fc", reader="qe")
    for centered in (False, True):
        out[f"si_auto_max_{int(centered)}"] = np.array(wb._estimate_omega_max_from_ifc(
            ifc, nkx=4, nky=3, nkz=2, centered_grid=centered, safety_factor=1.02))
    np.savez_compressed(HERE / "ifc_vectors.npz", **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src")
