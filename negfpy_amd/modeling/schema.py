"""Intermediate force-constant schema (reference ``src/negfpy/modeling/schema/*``,
``modeling/validators/ifc_validator.py``)."""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import numpy as np

Array = np.ndarray


@dataclass(frozen=True)
class IFCTerm:
    """One real-space force-constant block between a cell and the cell shifted by
    (dx, dy, dz)."""

    dx: int
    dy: int
    dz: int
    block: Array


@dataclass(frozen=True)
class IFCData:
    masses: Array
    dof_per_atom: int
    terms: tuple
    units: str = "unknown"
    metadata: dict[str, Any] = field(default_factory=dict)
    lattice_vectors: Array | None = None
    atom_positions: Array | None = None
    atom_symbols: tuple | None = None
    index_convention: str = "layer-major"


@dataclass(frozen=True)
class InterfaceData:
    """Coupling matrices and index maps of an interface (reference
    ``modeling/schema/interface_data.py``)."""

    left_contact: Array | None = None
    right_contact: Array | None = None
    index_map_left: tuple | None = None
    index_map_right: tuple | None = None
    convention: str = "direct"
    metadata: dict[str, Any] = field(default_factory=dict)


@dataclass(frozen=True)
class BuildConfig:
    onsite_pinning: float = 0.0
    principal_layer_size: int | None = None
    auto_principal_layer_enlargement: bool = True
    infer_fc01_from_negative_dx: bool = True
    enforce_hermitian_fc00: bool = True
    nyquist_split_half: bool = False
    dtype: str = "complex128"


def validate_ifc_data(ifc: IFCData) -> None:
    masses = np.asarray(ifc.masses, dtype=float)
    if masses.ndim != 1 or masses.size == 0:
        raise ValueError("IFCData.masses must be a non-empty 1D array.")
    if np.any(masses <= 0.0):
        raise ValueError("All IFCData masses must be positive.")
    if ifc.dof_per_atom <= 0:
        raise ValueError("IFCData.dof_per_atom must be positive.")
    if len(ifc.terms) == 0:
        raise ValueError("IFCData.terms must contain at least one term.")
    ndof = masses.size * ifc.dof_per_atom
    for term in ifc.terms:
        if np.asarray(term.block).shape != (ndof, ndof):
            raise ValueError("All IFC terms must have shape (n_atoms*dof, n_atoms*dof).")


def validate_transport_connectivity(ifc: IFCData) -> None:
    if not any(t.dx == 0 for t in ifc.terms):
        raise ValueError("IFCData must include at least one dx=0 term for onsite blocks.")
    if not any(t.dx != 0 for t in ifc.terms):
        raise ValueError("IFCData must include at least one nonzero dx term for layer coupling.")


def validate_material_kspace_params(params) -> None:
    """The (0, 0) same-layer term exists and has the model's shape (reference
    ``modeling/validators/model_validator.py:10``)."""
    ndof = params.ndof
    d00 = params.fc00_terms.get((0, 0))
    if d00 is None:
        raise ValueError("fc00_terms must contain the (0, 0) term.")
    if np.asarray(d00).shape != (ndof, ndof):
        raise ValueError("fc00_terms[(0, 0)] has invalid shape.")
