"""Square / cubic lattices with periodic transverse directions
(reference ``src/negfpy/models/lattice_kspace.py:13-157``).

Transport along x; one DOF per site.  At transverse wave vector k the lead and
device blocks are 1x1:  d00(k) = [2Kx + sum_a 2Ka (1 - cos k_a) + p] / m,
d01 = -Kx / m  (the coupling keeps the host mass also next to a mass defect).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from negfpy_amd.core.types import Device1D, DeviceKSpace, LeadKSpace


@dataclass(frozen=True)
class SquareLatticeParams:
    mass: float
    spring_x: float
    spring_y: float


@dataclass(frozen=True)
class CubicLatticeParams:
    mass: float
    spring_x: float
    spring_y: float
    spring_z: float
    onsite_pinning: float = 0.0


def _ky(kpar) -> float:
    if kpar is None or len(kpar) == 0:
        return 0.0
    if len(kpar) > 1:
        raise ValueError("Square-lattice kpar must contain at most one transverse component: (ky,).")
    return float(kpar[0])


def _kyz(kpar) -> tuple[float, float]:
    if kpar is None:
        return 0.0, 0.0
    if len(kpar) > 2:
        raise ValueError(
            "Cubic-lattice kpar must contain at most two transverse components: (ky,) or (ky, kz).")
    if len(kpar) == 1:
        return float(kpar[0]), 0.0
    if len(kpar) == 0:
        raise IndexError("tuple index out of range")
    return float(kpar[0]), float(kpar[1])


def _square_onsite_k(p: SquareLatticeParams, ky: float) -> float:
    return 2.0 * p.spring_x + 2.0 * p.spring_y - 2.0 * p.spring_y * np.cos(ky)


def _cubic_onsite_k(p: CubicLatticeParams, ky: float, kz: float) -> float:
    return (2.0 * p.spring_x + 2.0 * p.spring_y + 2.0 * p.spring_z + p.onsite_pinning
            - 2.0 * p.spring_y * np.cos(ky) - 2.0 * p.spring_z * np.cos(kz))


def _chain_device(n_layers, onsite_k, host_mass, spring_x, defect_index, defect_mass):
    onsite = []
    for i in range(n_layers):
        m = host_mass
        if defect_index is not None and defect_mass is not None and i == defect_index:
            m = defect_mass
        onsite.append(np.array([[onsite_k / m]], dtype=np.complex128))
    c = np.array([[-spring_x / host_mass]], dtype=np.complex128)
    return Device1D(onsite_blocks=onsite, coupling_blocks=[c.copy() for _ in range(n_layers - 1)])


def square_lattice_lead(params: SquareLatticeParams) -> LeadKSpace:
    def build(kpar):
        d00 = np.array([[_square_onsite_k(params, _ky(kpar)) / params.mass]], dtype=np.complex128)
        d01 = np.array([[-params.spring_x / params.mass]], dtype=np.complex128)
        return d00, d01

    return LeadKSpace(blocks_builder=build)


def square_lattice_device(n_layers: int, params: SquareLatticeParams,
                          defect_index: int | None = None,
                          defect_mass: float | None = None) -> DeviceKSpace:
    if defect_index is not None and (defect_index < 0 or defect_index >= n_layers):
        raise ValueError("defect_index out of range")

    def build(kpar):
        return _chain_device(n_layers, _square_onsite_k(params, _ky(kpar)), params.mass,
                             params.spring_x, defect_index, defect_mass)

    return DeviceKSpace(device_builder=build)


def cubic_lattice_lead(params: CubicLatticeParams) -> LeadKSpace:
    def build(kpar):
        ky, kz = _kyz(kpar)
        d00 = np.array([[_cubic_onsite_k(params, ky, kz) / params.mass]], dtype=np.complex128)
        d01 = np.array([[-params.spring_x / params.mass]], dtype=np.complex128)
        return d00, d01

    return LeadKSpace(blocks_builder=build)


def cubic_lattice_device(n_layers: int, params: CubicLatticeParams,
                         defect_index: int | None = None,
                         defect_mass: float | None = None) -> DeviceKSpace:
    if defect_index is not None and (defect_index < 0 or defect_index >= n_layers):
        raise ValueError("defect_index out of range")

    def build(kpar):
        ky, kz = _kyz(kpar)
        return _chain_device(n_layers, _cubic_onsite_k(params, ky, kz), params.mass,
                             params.spring_x, defect_index, defect_mass)

    return DeviceKSpace(device_builder=build)
