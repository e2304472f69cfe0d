"""Host-side lattice builders with the API of the reference ``negfpy.models``
(``src/negfpy/models/__init__.py:1-31``).  They produce the block containers of
``negfpy_amd.core.types`` that the GPU transport path consumes."""

from .chain import ChainParams, analytic_band_max, device_mass_defect, device_perfect_chain, lead_blocks
from .lattice_kspace import (
    CubicLatticeParams,
    SquareLatticeParams,
    cubic_lattice_device,
    cubic_lattice_lead,
    square_lattice_device,
    square_lattice_lead,
)
from .material_kspace import MaterialKspaceParams, material_kspace_device, material_kspace_lead
from .multidof import MultiDofCellParams, multidof_device, multidof_lead_blocks

__all__ = [
    "ChainParams",
    "lead_blocks",
    "device_perfect_chain",
    "device_mass_defect",
    "analytic_band_max",
    "SquareLatticeParams",
    "CubicLatticeParams",
    "square_lattice_lead",
    "square_lattice_device",
    "cubic_lattice_lead",
    "cubic_lattice_device",
    "MaterialKspaceParams",
    "material_kspace_lead",
    "material_kspace_device",
    "MultiDofCellParams",
    "multidof_lead_blocks",
    "multidof_device",
]
