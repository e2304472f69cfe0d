"""General force-constant k_par builder
(reference ``src/negfpy/models/material_kspace.py:16-195``).

Phi_ab(k) = sum_{(dy,dz)} Phi_ab(dy,dz) exp(i (ky dy + kz dz)) for the same-layer
(ab = 00), next-layer (01) and optional previous-layer (10) terms; without
explicit 10 terms Phi10(k) = Phi01(k)^H.  Blocks are mass-weighted as
D = M^-1/2 Phi M^-1/2 with the left/right layer masses.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from negfpy_amd.core.types import Device1D, DeviceKSpace, LeadKSpace

Array = np.ndarray
Shift2D = tuple[int, int]


@dataclass(frozen=True)
class MaterialKspaceParams:
    masses: Array
    fc00_terms: dict
    fc01_terms: dict
    fc10_terms: dict | None = None
    dof_per_atom: int = 3
    onsite_pinning: float = 0.0

    def __post_init__(self) -> None:
        m = np.asarray(self.masses, dtype=float)
        if m.ndim != 1 or m.size == 0:
            raise ValueError("masses must be a non-empty 1D array.")
        if np.any(m <= 0.0):
            raise ValueError("All masses must be positive.")
        if self.dof_per_atom <= 0:
            raise ValueError("dof_per_atom must be positive.")
        if len(self.fc00_terms) == 0:
            raise ValueError("fc00_terms must contain at least one term.")
        if len(self.fc01_terms) == 0:
            raise ValueError("fc01_terms must contain at least one term.")
        nd = m.size * self.dof_per_atom
        for name, terms in (("fc00_terms", self.fc00_terms), ("fc01_terms", self.fc01_terms)):
            for key, blk in terms.items():
                if len(key) != 2:
                    raise ValueError(f"{name} keys must be (dy, dz).")
                if blk.shape != (nd, nd):
                    raise ValueError(f"{name} blocks must have shape (n_atoms*dof, n_atoms*dof).")
        if self.fc10_terms is not None:
            if len(self.fc10_terms) == 0:
                raise ValueError("fc10_terms must be non-empty when provided.")
            for key, blk in self.fc10_terms.items():
                if len(key) != 2:
                    raise ValueError("fc10_terms keys must be (dy, dz).")
                if blk.shape != (nd, nd):
                    raise ValueError("fc10_terms blocks must have shape (n_atoms*dof, n_atoms*dof).")

    @property
    def n_atoms(self) -> int:
        return int(np.asarray(self.masses).size)

    @property
    def ndof(self) -> int:
        return self.n_atoms * self.dof_per_atom


def _kyz(kpar) -> tuple[float, float]:
    if kpar is None or len(kpar) == 0:
        return 0.0, 0.0
    if len(kpar) > 2:
        raise ValueError("kpar must contain at most two transverse components: (ky,) or (ky, kz).")
    if len(kpar) == 1:
        return float(kpar[0]), 0.0
    return float(kpar[0]), float(kpar[1])


def _fourier(terms: dict, ky: float, kz: float) -> Array:
    nd = next(iter(terms.values())).shape[0]
    out = np.zeros((nd, nd), dtype=np.complex128)
    for (dy, dz), blk in terms.items():
        out += np.asarray(blk, dtype=np.complex128) * np.exp(1j * (ky * dy + kz * dz))
    return out


def _dyn_blocks(params: MaterialKspaceParams, ky, kz, m_left, m_right):
    phi00 = _fourier(params.fc00_terms, ky, kz)
    if params.onsite_pinning != 0.0:
        phi00 = phi00 + params.onsite_pinning * np.eye(params.ndof, dtype=np.complex128)
    phi01 = _fourier(params.fc01_terms, ky, kz)
    phi10 = np.conj(phi01).T if params.fc10_terms is None else _fourier(params.fc10_terms, ky, kz)
    sl = 1.0 / np.sqrt(np.repeat(np.asarray(m_left, dtype=float), params.dof_per_atom))
    sr = 1.0 / np.sqrt(np.repeat(np.asarray(m_right, dtype=float), params.dof_per_atom))
    d00 = sl[:, None] * phi00 * sl[None, :]
    d01 = sl[:, None] * phi01 * sr[None, :]
    d10 = sr[:, None] * phi10 * sl[None, :]
    return d00, d01, d10


def material_kspace_lead(params: MaterialKspaceParams) -> LeadKSpace:
    def build(kpar):
        ky, kz = _kyz(kpar)
        return _dyn_blocks(params, ky, kz, params.masses, params.masses)

    # lets batched solvers build every k point's blocks on the GPU in one launch
    build.material_kspace = params
    return LeadKSpace(blocks_builder=build)


def material_kspace_device(n_layers: int, params: MaterialKspaceParams,
                           layer_masses: Array | None = None) -> DeviceKSpace:
    if n_layers <= 0:
        raise ValueError("n_layers must be positive.")
    if layer_masses is None:
        masses = np.repeat(np.asarray(params.masses, dtype=float)[None, :], n_layers, axis=0)
    else:
        masses = np.asarray(layer_masses, dtype=float)
        if masses.shape != (n_layers, params.n_atoms):
            raise ValueError("layer_masses shape must be (n_layers, n_atoms).")
        if np.any(masses <= 0.0):
            raise ValueError("All layer_masses entries must be positive.")

    def build(kpar):
        ky, kz = _kyz(kpar)
        onsite = [_dyn_blocks(params, ky, kz, masses[i], masses[i])[0] for i in range(n_layers)]
        coupling = [_dyn_blocks(params, ky, kz, masses[i], masses[i + 1])[1]
                    for i in range(n_layers - 1)]
        return Device1D(onsite_blocks=onsite, coupling_blocks=coupling)

    build.material_kspace = (params, int(n_layers), masses)
    return DeviceKSpace(device_builder=build)
