"""1D monatomic chain builders (reference ``src/negfpy/models/chain.py:16-57``).

Mass-weighted dynamical blocks of a chain of masses m joined by springs K:
onsite 2K/m, nearest-neighbour coupling -K/m; band edge w_max = 2 sqrt(K/m).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from negfpy_amd.core.types import Device1D, LeadBlocks


@dataclass(frozen=True)
class ChainParams:
    mass: float
    spring: float


def lead_blocks(params: ChainParams) -> LeadBlocks:
    k, m = params.spring, params.mass
    return LeadBlocks(d00=np.array([[2.0 * k / m]]), d01=np.array([[-k / m]]))


def device_perfect_chain(n_layers: int, params: ChainParams) -> Device1D:
    k, m = params.spring, params.mass
    onsite = [np.array([[2.0 * k / m]]) for _ in range(n_layers)]
    coupling = [np.array([[-k / m]]) for _ in range(n_layers - 1)]
    return Device1D(onsite_blocks=onsite, coupling_blocks=coupling)


def device_mass_defect(n_layers: int, params: ChainParams, defect_index: int,
                       defect_mass: float) -> Device1D:
    """Perfect chain with one site of mass ``defect_mass``; the couplings keep -K/m
    (reference chain.py:38-53 keeps the host mass in the coupling blocks)."""
    if defect_index < 0 or defect_index >= n_layers:
        raise ValueError("defect_index out of range")
    k, m = params.spring, params.mass
    onsite = [np.array([[2.0 * k / (defect_mass if i == defect_index else m)]])
              for i in range(n_layers)]
    coupling = [np.array([[-k / m]]) for _ in range(n_layers - 1)]
    return Device1D(onsite_blocks=onsite, coupling_blocks=coupling)


def analytic_band_max(params: ChainParams) -> float:
    return float(2.0 * np.sqrt(params.spring / params.mass))
