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
