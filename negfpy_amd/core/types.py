"""Block containers for the phonon-NEGF transport path.

Same names, fields and validation as the reference ``src/negfpy/core/types.py``
(``LeadBlocks`` :16-35, ``LeadKSpace`` :38-44, ``Device1D`` :48-76,
``DeviceKSpace`` :80-86), so user code that builds these objects works
unchanged.  The GPU path packs them into device tensors once per k-point
(``negfpy_amd.core._pack``).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable

import numpy as np

Array = np.ndarray
KPar = tuple[float, ...] | None


@dataclass(frozen=True)
class LeadBlocks:
    """Periodic lead: onsite block d00 and coupling d01 to the next cell (d10 optional)."""

    d00: Array
    d01: Array
    d10: Array | None = None

    def __post_init__(self) -> None:
        if self.d00.ndim != 2 or self.d00.shape[0] != self.d00.shape[1]:
            raise ValueError("d00 must be a square 2D array.")
        if self.d01.ndim != 2:
            raise ValueError("d01 must be a 2D array.")
        if self.d01.shape != self.d00.shape:
            raise ValueError("d01 must have the same shape as d00.")
        if self.d10 is not None:
            if self.d10.ndim != 2:
                raise ValueError("d10 must be a 2D array when provided.")
            if self.d10.shape != self.d00.shape:
                raise ValueError("d10 must have the same shape as d00.")


@dataclass(frozen=True)
class LeadKSpace:
    """Lead whose blocks depend on the transverse wave vector k_par."""

    blocks_builder: Callable[[KPar], tuple]

    def blocks(self, kpar: KPar = None) -> tuple:
        return self.blocks_builder(kpar)


@dataclass(frozen=True)
class Device1D:
    """Finite block-tridiagonal device: L onsite blocks and L-1 couplings (upper diagonal)."""

    onsite_blocks: list[Array]
    coupling_blocks: list[Array]

    def __post_init__(self) -> None:
        nl = len(self.onsite_blocks)
        if nl == 0:
            raise ValueError("Device must contain at least one principal layer.")
        if len(self.coupling_blocks) != nl - 1:
            raise ValueError("coupling_blocks length must be n_layers - 1.")
        shape = self.onsite_blocks[0].shape
        if len(shape) != 2 or shape[0] != shape[1]:
            raise ValueError("Each onsite block must be a square 2D array.")
        for blk in self.onsite_blocks:
            if blk.shape != shape:
                raise ValueError("All onsite blocks must share the same shape.")
        for blk in self.coupling_blocks:
            if blk.shape != shape:
                raise ValueError("All coupling blocks must match onsite block shape.")

    @property
    def n_layers(self) -> int:
        return len(self.onsite_blocks)

    @property
    def dof_per_layer(self) -> int:
        return int(self.onsite_blocks[0].shape[0])


@dataclass(frozen=True)
class DeviceKSpace:
    """Device whose blocks depend on k_par."""

    device_builder: Callable[[KPar], Device1D]

    def device(self, kpar: KPar = None) -> Device1D:
        return self.device_builder(kpar)
