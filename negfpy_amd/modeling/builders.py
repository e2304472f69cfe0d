"""From real-space force constants to principal-layer k-space terms
(reference ``src/negfpy/modeling/builders/ifc_to_terms.py``,
``material_kspace_builder.py``, ``modeling/ifc_tools.py``).

A principal layer groups ``pl`` consecutive cells along x.  A term (dx, dy, dz)
acting from cell r of a layer lands in cell r + dx: inside the layer it adds to
the same-layer term Phi00(dy, dz), in the next layer to Phi01(dy, dz), in the
previous layer to Phi10(dy, dz) (as a (cell r, cell r + dx mod pl) sub-block).
"""

from __future__ import annotations

import numpy as np

from negfpy_amd.models import MaterialKspaceParams

from .schema import BuildConfig, IFCData, IFCTerm, validate_ifc_data, validate_transport_connectivity


def enforce_translational_asr_on_self_term(ifc: IFCData):
    """Correct the R = 0 block so every row sums to zero per Cartesian direction,
    then Hermitise it (reference modeling/ifc_tools.py:10-59).  Returns
    (corrected ifc, largest residual before the correction)."""
    dof = ifc.dof_per_atom
    ndof = len(ifc.masses) * dof
    terms = list(ifc.terms)
    total = np.zeros((ndof, ndof), dtype=np.complex128)
    i0 = None
    for idx, t in enumerate(terms):
        total += np.asarray(t.block, dtype=np.complex128)
        if t.dx == 0 and t.dy == 0 and t.dz == 0:
            i0 = idx
    if i0 is None:
        raise ValueError("Missing IFC term at translation (0,0,0); cannot enforce ASR.")
    res = np.zeros((ndof, dof), dtype=np.complex128)
    for beta in range(dof):
        res[:, beta] = np.sum(total[:, np.arange(beta, ndof, dof)], axis=1)
    fixed = np.asarray(terms[i0].block, dtype=np.complex128).copy()
    for row in range(ndof):
        base = (row // dof) * dof
        for beta in range(dof):
            fixed[row, base + beta] -= res[row, beta]
    fixed = 0.5 * (fixed + fixed.conj().T)
    terms[i0] = IFCTerm(dx=0, dy=0, dz=0, block=fixed)
    meta = dict(ifc.metadata)
    meta["asr_enforced"] = True
    return (IFCData(masses=np.asarray(ifc.masses, dtype=float), dof_per_atom=dof,
                    terms=tuple(terms), units=ifc.units, metadata=meta,
                    lattice_vectors=ifc.lattice_vectors, atom_positions=ifc.atom_positions,
                    atom_symbols=ifc.atom_symbols, index_convention=ifc.index_convention),
            float(np.max(np.abs(res))))


def _nr1(ifc: IFCData):
    nr = ifc.metadata.get("nr") if isinstance(ifc.metadata, dict) else None
    try:
        return int(nr[0]) if nr is not None and len(nr) == 3 else None
    except (TypeError, ValueError):
        return None


def full_realspace_grid(ifc: IFCData) -> bool:
    """True when the terms cover every shift of the nr1 x nr2 x nr3 grid (as q2r writes)."""
    nr = ifc.metadata.get("nr") if isinstance(ifc.metadata, dict) else None
    if nr is None or len(nr) != 3:
        return False
    try:
        n1, n2, n3 = int(nr[0]), int(nr[1]), int(nr[2])
    except (TypeError, ValueError):
        return False
    if min(n1, n2, n3) <= 0:
        return False
    return len({(int(t.dx), int(t.dy), int(t.dz)) for t in ifc.terms}) == n1 * n2 * n3


def principal_layer_size(ifc: IFCData, config: BuildConfig | None = None) -> int:
    """Cells per principal layer (ifc_to_terms.py:18-52): the configured size (refused
    when it would split the Nyquist x coupling of an even full grid, unless
    nyquist_split_half), else max|dx| with auto enlargement, else 1."""
    config = config or BuildConfig()
    max_dx = max(abs(int(t.dx)) for t in ifc.terms)
    if config.principal_layer_size is not None:
        if config.principal_layer_size <= 0:
            raise ValueError("principal_layer_size must be positive when provided.")
        pl = int(config.principal_layer_size)
        n1 = _nr1(ifc)
        nyquist_x = n1 is not None and n1 > 1 and n1 % 2 == 0 and max_dx == n1 // 2
        if full_realspace_grid(ifc) and nyquist_x and pl > max_dx and not config.nyquist_split_half:
            raise ValueError(
                "principal_layer_size is too large for this full-grid IFC: "
                f"got {pl}, but max |dx| is {max_dx}. "
                "For even FFT grids, use principal_layer_size <= max|dx| "
                "(recommended: omit it and use auto inference). "
                "Alternatively set BuildConfig.nyquist_split_half=true "
                "to enable symmetric half-half Nyquist splitting.")
        return pl
    if not config.auto_principal_layer_enlargement:
        return 1
    return max(1, max_dx)


def build_fc_terms(ifc: IFCData, config: BuildConfig | None = None):
    """(Phi00, Phi01, Phi10 or None) term dictionaries keyed by (dy, dz)
    (ifc_to_terms.py:70-185)."""
    config = config or BuildConfig()
    validate_ifc_data(ifc)
    validate_transport_connectivity(ifc)
    dtype = np.dtype(config.dtype)
    cell = np.asarray(ifc.masses).size * ifc.dof_per_atom
    pl = principal_layer_size(ifc, config)
    nd = pl * cell
    full = full_realspace_grid(ifc)
    infer01 = config.infer_fc01_from_negative_dx and not full
    terms = tuple(ifc.terms)
    present = {(int(t.dx), int(t.dy), int(t.dz)) for t in terms}
    n1 = _nr1(ifc)
    max_dx = max(abs(int(t.dx)) for t in terms)
    if (config.nyquist_split_half and full and n1 is not None and n1 > 1 and n1 % 2 == 0
            and max_dx == n1 // 2 and pl > max_dx):
        nyq = n1 // 2
        if any(int(t.dx) == -nyq for t in terms) and not any(int(t.dx) == nyq for t in terms):
            # an even grid stores the Nyquist x coupling once, at -nyq: split it
            # Hermitian-symmetrically between -nyq and +nyq
            split = []
            for t in terms:
                if int(t.dx) == -nyq:
                    b = np.asarray(t.block, dtype=dtype)
                    half = 0.5 * (0.5 * (b + b.conj().T))
                    split.append(IFCTerm(dx=-nyq, dy=int(t.dy), dz=int(t.dz), block=half))
                    split.append(IFCTerm(dx=+nyq, dy=int(t.dy), dz=int(t.dz), block=half))
                else:
                    split.append(t)
            terms = tuple(split)
            present = {(int(t.dx), int(t.dy), int(t.dz)) for t in terms}

    acc00: dict = {}
    acc01: dict = {}
    acc10: dict = {}

    def add(acc, key, r, c, blk):
        m = acc.get(key)
        if m is None:
            m = acc[key] = np.zeros((nd, nd), dtype=dtype)
        m[r * cell:(r + 1) * cell, c * cell:(c + 1) * cell] += blk

    for t in terms:
        blk = np.asarray(t.block, dtype=dtype)
        for r in range(pl):
            c = r + int(t.dx)
            if 0 <= c < pl:
                add(acc00, (t.dy, t.dz), r, c, blk)
            elif pl <= c < 2 * pl:
                add(acc01, (t.dy, t.dz), r, c - pl, blk)
            elif -pl <= c < 0:
                add(acc10, (t.dy, t.dz), r, c + pl, blk)
                if infer01 and (-int(t.dx), -int(t.dy), -int(t.dz)) not in present:
                    add(acc01, (-t.dy, -t.dz), c + pl, r, blk.conj().T)
    if not acc00:
        raise ValueError("No dx=0 IFC terms were mapped to fc00_terms.")
    if not acc01:
        raise ValueError("No x-coupling IFC terms were mapped to fc01_terms.")
    fc00 = {k: np.asarray(v, dtype=dtype) for k, v in acc00.items()}
    fc01 = {k: np.asarray(v, dtype=dtype) for k, v in acc01.items()}
    fc10 = {k: np.asarray(v, dtype=dtype) for k, v in acc10.items()}
    if config.enforce_hermitian_fc00 and not full:
        fc00 = {k: 0.5 * (b + b.conj().T) for k, b in fc00.items()}
    return fc00, fc01, (None if (not full and not fc10) else fc10)


def build_material_kspace_params(ifc: IFCData, config: BuildConfig | None = None
                                 ) -> MaterialKspaceParams:
    """MaterialKspaceParams of the principal layer (material_kspace_builder.py:15-31)."""
    config = config or BuildConfig()
    validate_ifc_data(ifc)
    pl = principal_layer_size(ifc, config)
    fc00, fc01, fc10 = build_fc_terms(ifc, config)
    return MaterialKspaceParams(masses=np.tile(np.asarray(ifc.masses, dtype=float), pl),
                                fc00_terms=fc00, fc01_terms=fc01, fc10_terms=fc10,
                                dof_per_atom=ifc.dof_per_atom,
                                onsite_pinning=float(config.onsite_pinning))


infer_principal_layer_size = principal_layer_size


def build_transport_components(ifc: IFCData, n_layers: int, config: BuildConfig | None = None,
                               layer_masses=None) -> dict:
    """Force-constant params with their lead and device (reference
    ``modeling/builders/material_kspace_builder.py:32``)."""
    from negfpy_amd.models import material_kspace_device, material_kspace_lead

    params = build_material_kspace_params(ifc, config)
    return {"params": params, "lead": material_kspace_lead(params),
            "device": material_kspace_device(n_layers=n_layers, params=params,
                                             layer_masses=layer_masses)}
