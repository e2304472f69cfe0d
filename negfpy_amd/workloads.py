"""Seeded synthetic transport problems used by the benchmark and the tests.

Force constants are sums of pair spring tensors: each interacting atom pair
(a, b) gets a random positive semi-definite 3x3 tensor K_ab, Phi_ab = -K_ab off
the diagonal and the acoustic sum rule Phi_aa = sum_b K_ab on it.  That keeps
D(k) = M^-1/2 Phi(k) M^-1/2 Hermitian and positive semi-definite with an
acoustic branch at k = 0, as a physical crystal has.  Units: force constants in
eV/A^2, masses in amu, omega in sqrt(eV/A^2/amu) (OMEGA_UNIT_RAD_S rad/s).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from negfpy_amd.core.types import Device1D, LeadBlocks
from negfpy_amd.models.multidof import MultiDofCellParams, multidof_device, multidof_lead_blocks

EV_J = 1.602176634e-19
AMU_KG = 1.66053906660e-27
ANGSTROM_M = 1e-10
OMEGA_UNIT_RAD_S = float(np.sqrt(EV_J / (ANGSTROM_M ** 2 * AMU_KG)))


def _spring(rng, fc_max):
    a = rng.uniform(-1.0, 1.0, (3, 3))
    k = a @ a.T
    return fc_max * k / np.abs(k).max()


def ribbon_params(seed: int = 0, n_atoms: int = 8, fc_max: float = 5.0,
                  mass_range=(10.0, 30.0)) -> MultiDofCellParams:
    """Quasi-1D ribbon cell: every atom pair inside the cell and to the next cell is
    coupled by a random spring tensor with entries in [-fc_max, fc_max]."""
    rng = np.random.default_rng(seed)
    masses = rng.uniform(mass_range[0], mass_range[1], n_atoms)
    nd = 3 * n_atoms
    fc00 = np.zeros((nd, nd))
    fc01 = np.zeros((nd, nd))
    for a in range(n_atoms):
        for b in range(a + 1, n_atoms):
            k = _spring(rng, fc_max)
            fc00[3 * a:3 * a + 3, 3 * b:3 * b + 3] = -k
            fc00[3 * b:3 * b + 3, 3 * a:3 * a + 3] = -k
            fc00[3 * a:3 * a + 3, 3 * a:3 * a + 3] += k
            fc00[3 * b:3 * b + 3, 3 * b:3 * b + 3] += k
    for a in range(n_atoms):
        for b in range(n_atoms):
            k = _spring(rng, fc_max)
            fc01[3 * a:3 * a + 3, 3 * b:3 * b + 3] = -k
            fc00[3 * a:3 * a + 3, 3 * a:3 * a + 3] += k
            fc00[3 * b:3 * b + 3, 3 * b:3 * b + 3] += k
    return MultiDofCellParams(masses=masses, fc00=fc00, fc01=fc01, dof_per_atom=3)


def omega_max(lead: LeadBlocks, nk: int = 64) -> float:
    """Top of the lead's bulk band: max over kx of sqrt(eig D(kx))."""
    d00 = np.asarray(lead.d00, dtype=np.complex128)
    d01 = np.asarray(lead.d01, dtype=np.complex128)
    d10 = d01.conj().T if lead.d10 is None else np.asarray(lead.d10, dtype=np.complex128)
    top = 0.0
    for kx in np.linspace(0.0, np.pi, nk):
        dk = d00 + d01 * np.exp(1j * kx) + d10 * np.exp(-1j * kx)
        top = max(top, float(np.linalg.eigvalsh(0.5 * (dk + dk.conj().T)).max()))
    return float(np.sqrt(max(top, 0.0)))


@dataclass
class RibbonProblem:
    """Quasi-1D multi-DOF ribbon: n = 3 * n_atoms DOF per layer, mass-disordered device."""

    lead: LeadBlocks
    device: Device1D
    omegas: np.ndarray
    eta: float
    params: MultiDofCellParams
    wmax: float


def ribbon_problem(seed: int = 0, n_atoms: int = 8, n_layers: int = 32, n_omega: int = 4096,
                   eta: float = 1e-6, span: float = 1.1) -> RibbonProblem:
    """The 24-DOF, 32-layer, 4096-omega ribbon workload (for n_atoms = 8).

    Lead cells use the seeded masses; every device layer draws its own masses
    from the same range, so the device scatters.  omega_i = span * w_max (i + 1/2) / n_omega.
    """
    params = ribbon_params(seed, n_atoms)
    rng = np.random.default_rng(seed + 1)
    layer_masses = rng.uniform(10.0, 30.0, (n_layers, n_atoms))
    lead = multidof_lead_blocks(params)
    device = multidof_device(n_layers, params, layer_masses=layer_masses)
    wmax = omega_max(lead)
    omegas = span * wmax * (np.arange(n_omega) + 0.5) / n_omega
    return RibbonProblem(lead=lead, device=device, omegas=omegas, eta=eta, params=params,
                         wmax=wmax)


def random_lead(seed: int, n: int, scale: float = 1.0):
    """Small random Hermitian lead (d00, d01) with a real band, for parity tests."""
    n_atoms = max(1, n // 3)
    if 3 * n_atoms == n:
        p = ribbon_params(seed, n_atoms, fc_max=scale)
        lb = multidof_lead_blocks(p)
        return np.asarray(lb.d00), np.asarray(lb.d01)
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n))
    d01 = -scale * (a @ a.conj().T) / n
    b = rng.uniform(-1, 1, (n, n))
    d00 = scale * (b @ b.T) / n
    d00 = d00 - d01 - d01.conj().T
    d00 = 0.5 * (d00 + d00.conj().T)
    return d00, d01


# (atoms per layer -> 24, 48, 96 DOF, omega points) of the uneven batch
UNEVEN_PARTS = ((8, 4096), (16, 1024), (32, 256))


def uneven_problems(seed: int = 0, n_layers: int = 8, edge_fraction: float = 0.25,
                    eta: float = 1e-6) -> list[RibbonProblem]:
    """An uneven batch: ribbons of 24-, 48- and 96-DOF layers (4096, 1024 and 256
    omega points, n_layers device layers each).  A quarter of each spectrum is
    packed within 1 % of the band top, where the decimation needs the most steps, so
    step counts vary strongly between neighbouring instances."""
    out = []
    for i, (n_atoms, n_omega) in enumerate(UNEVEN_PARTS):
        p = ribbon_problem(seed + 10 * i, n_atoms=n_atoms, n_layers=n_layers, n_omega=n_omega,
                           eta=eta)
        m = int(round(edge_fraction * n_omega))
        w = p.omegas.copy()
        w[::max(n_omega // m, 1)][:m] = p.wmax * (1.0 + np.linspace(-0.01, 0.01, m))
        p.omegas = np.sort(w)
        out.append(p)
    return out


@dataclass
class InterfaceProblem:
    """Mass-mismatch interface between two different leads."""

    lead_left: LeadBlocks
    lead_right: LeadBlocks
    device: Device1D
    omegas: np.ndarray
    eta: float


def interface_problem(seed: int = 0, n_atoms: int = 16, n_layers: int = 64, n_omega: int = 512,
                      m_left: float = 28.0, m_right: float = 73.0, perturb: float = 0.2,
                      n_perturbed: int = 4, eta: float = 1e-6) -> InterfaceProblem:
    """The adjoint workload: a ribbon cell's force constants (3 n_atoms DOF per layer)
    with every atom of mass m_left in the left lead and the left half of the device
    and m_right in the right half and the right lead; the onsite force constants of
    the last n_perturbed device layers get random symmetric relative perturbations
    of up to +-perturb.  omega_i = 1.1 w_max (i + 1/2) / n_omega with w_max the band
    top of the lighter (left) lead."""
    base = ribbon_params(seed, n_atoms)
    left = MultiDofCellParams(masses=np.full(n_atoms, m_left), fc00=base.fc00, fc01=base.fc01)
    right = MultiDofCellParams(masses=np.full(n_atoms, m_right), fc00=base.fc00,
                               fc01=base.fc01)
    masses = np.where(np.arange(n_layers)[:, None] < n_layers // 2, m_left, m_right)
    dev = multidof_device(n_layers, left, layer_masses=masses * np.ones((1, n_atoms)))
    rng = np.random.default_rng(seed + 3)
    onsite = [np.array(b) for b in dev.onsite_blocks]
    nd = 3 * n_atoms
    for i in range(n_layers - n_perturbed, n_layers):
        r = rng.uniform(-1.0, 1.0, (nd, nd))
        onsite[i] = onsite[i] * (1.0 + perturb * 0.5 * (r + r.T))
    device = Device1D(onsite_blocks=onsite, coupling_blocks=dev.coupling_blocks)
    lead_l = multidof_lead_blocks(left)
    wmax = omega_max(lead_l)
    omegas = 1.1 * wmax * (np.arange(n_omega) + 0.5) / n_omega
    return InterfaceProblem(lead_left=lead_l, lead_right=multidof_lead_blocks(right),
                            device=device, omegas=omegas, eta=eta)


# bond length (A) and (radial, in-plane tangential, out-of-plane) springs in eV/A^2 of
# the first, second and third neighbour shells of the graphene-like sheet
SHEET_BOND = 1.42
SHEET_SPRINGS = ((22.8, 15.3, 6.1), (5.5, 2.0, 0.5), (1.9, 1.0, 0.2))


def sheet_params(seed: int = 0, jitter: float = 0.05, mass: float = 12.0):
    """Graphene-like honeycomb sheet in its 4-atom orthorhombic cell (12 DOF):
    transport along x (period 3 d), one transverse direction y (period sqrt(3) d),
    d = SHEET_BOND.  Every bond of the first three neighbour shells (d, sqrt(3) d,
    2 d) is a spring K = kr e e^T + kt t t^T + kz z z^T (e the bond direction, t its
    in-plane normal) with the shell's springs times a seeded factor in
    [1 - jitter, 1 + jitter] per bond; Phi_ab = -K off the diagonal, acoustic sum
    rule on it.  Returns ``MaterialKspaceParams`` with (dy, 0) terms."""
    from negfpy_amd.models.material_kspace import MaterialKspaceParams

    d = SHEET_BOND
    pos = np.array([[0.0, 0.0], [d, 0.0], [1.5 * d, 0.5 * np.sqrt(3) * d],
                    [2.5 * d, 0.5 * np.sqrt(3) * d]])
    a1, a2 = np.array([3.0 * d, 0.0]), np.array([0.0, np.sqrt(3) * d])
    shells = (d, np.sqrt(3) * d, 2.0 * d)
    rng = np.random.default_rng(seed)
    factor: dict = {}
    phi: dict = {}
    onsite = np.zeros((12, 12))
    for a in range(4):
        for b in range(4):
            for sx in (-1, 0, 1):
                for sy in (-1, 0, 1):
                    r = pos[b] + sx * a1 + sy * a2 - pos[a]
                    dist = float(np.hypot(*r))
                    shell = [i for i, s in enumerate(shells) if abs(dist - s) < 1e-6]
                    if not shell:
                        continue
                    # one factor per bond: (a, b, s) and (b, a, -s) are the same bond
                    key = (a, b, sx, sy)
                    twin = (b, a, -sx, -sy)
                    if twin in factor:
                        factor[key] = factor[twin]
                    elif key not in factor:
                        factor[key] = 1.0 + jitter * rng.uniform(-1.0, 1.0)
                    kr, kt, kz = (factor[key] * k for k in SHEET_SPRINGS[shell[0]])
                    e = np.array([r[0], r[1], 0.0]) / dist
                    t = np.array([-e[1], e[0], 0.0])
                    z = np.array([0.0, 0.0, 1.0])
                    k = kr * np.outer(e, e) + kt * np.outer(t, t) + kz * np.outer(z, z)
                    blk = phi.setdefault((sx, sy), np.zeros((12, 12)))
                    blk[3 * a:3 * a + 3, 3 * b:3 * b + 3] -= k
                    onsite[3 * a:3 * a + 3, 3 * a:3 * a + 3] += k
    phi.setdefault((0, 0), np.zeros((12, 12)))
    phi[(0, 0)] += onsite
    fc00 = {(sy, 0): phi[(0, sy)] for sy in (-1, 0, 1) if (0, sy) in phi}
    fc01 = {(sy, 0): phi[(1, sy)] for sy in (-1, 0, 1) if (1, sy) in phi}
    return MaterialKspaceParams(masses=np.full(4, float(mass)), fc00_terms=fc00,
                                fc01_terms=fc01)


def sheet_omega_max(params, nk: int = 24) -> float:
    """Top of the sheet's bulk band: max over a (kx, ky) grid of sqrt(eig D)."""
    from negfpy_amd.models.material_kspace import material_kspace_lead

    build = material_kspace_lead(params).blocks_builder
    top = 0.0
    for ky in np.linspace(-np.pi, np.pi, nk, endpoint=False):
        d00, d01, d10 = build((ky,))
        for kx in np.linspace(0.0, np.pi, nk):
            dk = d00 + d01 * np.exp(1j * kx) + d10 * np.exp(-1j * kx)
            top = max(top, float(np.linalg.eigvalsh(0.5 * (dk + dk.conj().T)).max()))
    return float(np.sqrt(top))


@dataclass
class SheetProblem:
    """k-averaged transport through a pristine graphene-like sheet."""

    params: object
    lead_left: object
    lead_right: object
    v_left: object
    v_right: object
    device: object
    kpoints: list
    omegas: np.ndarray
    eta: float
    wmax: float


def sheet_problem(seed: int = 0, n_layers: int = 16, n_k: int = 128, n_omega: int = 2048,
                  eta: float = 1e-6, span: float = 1.1) -> SheetProblem:
    """The 2-D sheet workload: an n_layers device of ``sheet_params`` cells
    (``material_kspace_device``) between leads and couplings from
    ``interface_contacts`` (the left lead's outward coupling is Phi_10, so its
    surface Green's function faces left; the sheet's cell has no x mirror), all
    built per k point; ky_j = -pi + 2 pi (j + 1/2) / n_k and
    omega_i = span * w_max (i + 1/2) / n_omega: n_k * n_omega instances."""
    from negfpy_amd.modeling.contacts import interface_contacts
    from negfpy_amd.models.material_kspace import material_kspace_device

    params = sheet_params(seed)
    wmax = sheet_omega_max(params)
    kpoints = [(-np.pi + 2.0 * np.pi * (j + 0.5) / n_k,) for j in range(n_k)]
    omegas = span * wmax * (np.arange(n_omega) + 0.5) / n_omega
    lead_l, lead_r, v_l, v_r = interface_contacts(params)
    return SheetProblem(params=params, lead_left=lead_l, lead_right=lead_r, v_left=v_l,
                        v_right=v_r, device=material_kspace_device(n_layers, params),
                        kpoints=kpoints, omegas=omegas, eta=eta, wmax=wmax)
