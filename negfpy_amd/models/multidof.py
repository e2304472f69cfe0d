"""Multi-atom, multi-DOF principal layers from force constants
(reference ``src/negfpy/models/multidof.py:16-142``).

Dynamical blocks are D = M^-1/2 Phi M^-1/2: onsite D00 = M_i^-1/2 (Phi00 + p I) M_i^-1/2,
coupling D01 = M_i^-1/2 Phi01 M_{i+1}^-1/2 (p = onsite pinning, onsite only).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from negfpy_amd.core.types import Device1D, LeadBlocks

Array = np.ndarray


@dataclass(frozen=True)
class MultiDofCellParams:
    masses: Array
    fc00: Array
    fc01: Array
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
        nd = m.size * self.dof_per_atom
        if self.fc00.shape != (nd, nd):
            raise ValueError("fc00 shape must be (n_atoms*dof_per_atom, n_atoms*dof_per_atom).")
        if self.fc01.shape != (nd, nd):
            raise ValueError("fc01 shape must match fc00 shape.")

    @property
    def n_atoms(self) -> int:
        return int(np.asarray(self.masses).size)

    @property
    def ndof(self) -> int:
        return self.n_atoms * self.dof_per_atom


def _mass_scale(masses: Array, dof_per_atom: int) -> Array:
    """Vector of m^-1/2 per DOF (diagonal of M^-1/2)."""
    return 1.0 / np.sqrt(np.repeat(np.asarray(masses, dtype=float), dof_per_atom))


def _blocks(m_left, m_right, fc00, fc01, dof, pinning=0.0):
    sl = _mass_scale(m_left, dof)
    sr = _mass_scale(m_right, dof)
    phi00 = np.asarray(fc00, dtype=np.complex128)
    if pinning != 0.0:
        phi00 = phi00 + pinning * np.eye(phi00.shape[0])
    d00 = sl[:, None] * phi00 * sl[None, :]
    d01 = sl[:, None] * np.asarray(fc01, dtype=np.complex128) * sr[None, :]
    return d00, d01


def multidof_lead_blocks(params: MultiDofCellParams) -> LeadBlocks:
    d00, d01 = _blocks(params.masses, params.masses, params.fc00, params.fc01,
                       params.dof_per_atom, params.onsite_pinning)
    return LeadBlocks(d00=d00, d01=d01)


def multidof_device(n_layers: int, params: MultiDofCellParams,
                    layer_masses: Array | None = None) -> Device1D:
    """Device of ``n_layers`` cells; ``layer_masses`` (n_layers, n_atoms) sets per-layer masses."""
    if n_layers <= 0:
        raise ValueError("n_layers must be positive.")
    if layer_masses is None:
        masses = np.repeat(np.asarray(params.masses, dtype=float)[None, :], n_layers, axis=0)
    else:
        masses = np.asarray(layer_masses, dtype=float)
        if masses.shape != (n_layers, params.n_atoms):
            raise ValueError("layer_masses shape must be (n_layers, n_atoms).")
        if np.any(masses <= 0.0):
            raise ValueError("layer_masses entries must be positive.")
    dof = params.dof_per_atom
    onsite = [_blocks(masses[i], masses[i], params.fc00, params.fc01, dof,
                      params.onsite_pinning)[0] for i in range(n_layers)]
    coupling = [_blocks(masses[i], masses[i + 1], params.fc00, params.fc01, dof)[1]
                for i in range(n_layers - 1)]
    return Device1D(onsite_blocks=onsite, coupling_blocks=coupling)
