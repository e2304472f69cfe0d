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


# (radial, tangential) springs in eV/A^2 of the first and second neighbour shells of
# the simple-cubic cell of the 3-D interface workload
CUBIC_SPRINGS = ((10.0, 2.0), (3.0, 0.5))


def _pair_phi(pos, size, shells, springs, seed: int | None, jitter: float):
    """Force constants Phi(dx, dy, dz) (3 N x 3 N) between a cell of N atoms at ``pos``
    (box ``size``, transport along x) and its images shifted by whole cells: for every
    pair at one of the ``shells`` distances, K = kr e e^T + kt (I - e e^T) with that
    shell's (kr, kt) times a seeded factor in [1 - jitter, 1 + jitter] (the same for
    the bond seen from either atom), Phi_ab = -K off the diagonal and the acoustic sum
    rule on it.  Bonds must be shorter than the cell so only adjacent cells couple."""
    pos = np.asarray(pos, float)
    size = np.asarray(size, float)
    na = len(pos)
    rng = np.random.default_rng(seed)
    factor: dict = {}
    phi: dict = {}
    onsite = np.zeros((3 * na, 3 * na))
    for a in range(na):
        for b in range(na):
            for s in np.ndindex(3, 3, 3):
                sh = np.array(s) - 1
                r = pos[b] + sh * size - pos[a]
                d = float(np.linalg.norm(r))
                shell = next((q for q, ds in enumerate(shells) if abs(d - ds) < 1e-9), -1)
                if shell < 0:
                    continue
                key, twin = (a, b) + tuple(sh), (b, a) + tuple(-sh)
                if key not in factor:
                    factor[key] = factor.get(twin, 1.0 + jitter * rng.uniform(-1.0, 1.0))
                kr, kt = (factor[key] * x for x in springs[shell])
                e = r / d
                k = kr * np.outer(e, e) + kt * (np.eye(3) - np.outer(e, e))
                blk = phi.setdefault(tuple(sh), np.zeros((3 * na, 3 * na)))
                blk[3 * a:3 * a + 3, 3 * b:3 * b + 3] -= k
                onsite[3 * a:3 * a + 3, 3 * a:3 * a + 3] += k
    phi.setdefault((0, 0, 0), np.zeros((3 * na, 3 * na)))
    phi[(0, 0, 0)] += onsite
    return phi


def _cubic_phi(cell=(4, 2, 2), seed: int | None = None, jitter: float = 0.0):
    """``_pair_phi`` of a simple-cubic cell of nx x ny x nz atoms (unit spacing) with
    first and second neighbour springs (``CUBIC_SPRINGS``)."""
    nx, ny, nz = cell
    pos = [(i, j, k) for i in range(nx) for j in range(ny) for k in range(nz)]
    return _pair_phi(pos, cell, (1.0, np.sqrt(2.0)), CUBIC_SPRINGS, seed, jitter)


# (radial, tangential) springs in eV/A^2 of the first (a sqrt(3) / 4, along <111>) and
# second (a / sqrt(2), along <110>) neighbour shells of the silicon-like diamond cells
DIAMOND_SPRINGS = ((9.0, 1.5), (1.5, 0.25))
_FCC = ((0.0, 0.0, 0.0), (0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0))


def _diamond_phi(cells=(1, 1, 1), seed: int | None = None, jitter: float = 0.0):
    """``_pair_phi`` of cx x cy x cz conventional diamond cubes (cube edge 1, 8 atoms
    each: fcc sites and the (1/4, 1/4, 1/4) basis) with first and second neighbour
    springs (``DIAMOND_SPRINGS``)."""
    cx, cy, cz = cells
    pos = [(i + f[0] + o, j + f[1] + o, k + f[2] + o)
           for i in range(cx) for j in range(cy) for k in range(cz)
           for o in (0.0, 0.25) for f in _FCC]
    return _pair_phi(pos, cells, (np.sqrt(3.0) / 4.0, np.sqrt(2.0) / 2.0), DIAMOND_SPRINGS,
                     seed, jitter)


@dataclass
class Interface3DProblem:
    """Mass-mismatch interface in a 3-D crystal, k-resolved, with the device force
    constants as per-layer terms (gradients of kappa flow to them)."""

    terms: list          # transverse shifts (dy, dz) of the force-constant terms
    fc00: np.ndarray     # (L, n_terms, n, n) same-layer force constants of every layer
    fc01: np.ndarray     # (L - 1, n_terms, n, n) layer l -> l + 1
    masses: np.ndarray   # (L, n) per-DOF masses of every device layer
    lead_left: np.ndarray    # (S, 3, n, n) left lead, facing left: (d00, d10, d01)
    lead_right: np.ndarray   # (S, 3, n, n)
    v_left: np.ndarray   # (S, n, n) device layer 0 -> left lead
    v_right: np.ndarray  # (S, n, n) last device layer -> right lead
    phases: np.ndarray   # (S, n_terms) e^{i (ky dy + kz dz)}
    kpoints: list
    omegas: np.ndarray
    eta: float


def interface3d_problem(seed: int = 0, cell=(4, 2, 2), n_layers: int = 64, nk: int = 16,
                        n_omega: int = 1024, m_left: float = 28.0, m_right: float = 73.0,
                        perturb: float = 0.2, n_perturbed: int = 4,
                        eta: float = 1e-6) -> Interface3DProblem:
    """The 3-D gradient workload: simple-cubic cells of prod(cell) atoms (48 DOF for
    4 x 2 x 2) with first and second neighbour springs, n_layers device layers, mass
    m_left in the left lead and the left half of the device, m_right in the right half
    and the right lead; the same-layer force constants of the last n_perturbed layers
    get random symmetric relative perturbations of up to +-perturb (Phi00(-d) =
    Phi00(d)^T kept).  An nk x nk (ky, kz) mesh, ky_j = -pi + 2 pi (j + 1/2) / nk, and
    omega_i = 1.1 w_max (i + 1/2) / n_omega, w_max the band top of the light side.
    Leads as the interface contacts of a bulk model (the left lead's outward coupling
    is Phi_10)."""
    phi = _cubic_phi(cell)
    n = next(iter(phi.values())).shape[0]
    terms = sorted({(dy, dz) for (_, dy, dz) in phi})
    zero = np.zeros((n, n))
    f00 = np.stack([phi.get((0,) + t, zero) for t in terms])
    f01 = np.stack([phi.get((1,) + t, zero) for t in terms])
    rng = np.random.default_rng(seed)
    fc00 = np.repeat(f00[None], n_layers, axis=0)
    index = {t: i for i, t in enumerate(terms)}
    for layer in range(n_layers - n_perturbed, n_layers):
        done = set()
        for t in terms:
            if t in done:
                continue
            mt = (-t[0], -t[1])
            r = rng.uniform(-1.0, 1.0, (n, n))
            if mt == t:
                r = 0.5 * (r + r.T)
            fc00[layer, index[t]] = f00[index[t]] * (1.0 + perturb * r)
            fc00[layer, index[mt]] = fc00[layer, index[t]].T if mt != t else fc00[layer, index[t]]
            done.update({t, mt})
    fc01 = np.repeat(f01[None], n_layers - 1, axis=0)
    dof_mass = np.where(np.arange(n_layers)[:, None] < n_layers // 2, m_left, m_right)
    masses = dof_mass * np.ones((1, n))
    kpoints = [(-np.pi + 2 * np.pi * (a + 0.5) / nk, -np.pi + 2 * np.pi * (b + 0.5) / nk)
               for a in range(nk) for b in range(nk)]
    ph = np.array([[np.exp(1j * (ky * dy + kz * dz)) for (dy, dz) in terms]
                   for (ky, kz) in kpoints])
    p00 = np.einsum("st,tij->sij", ph, f00)
    p01 = np.einsum("st,tij->sij", ph, f01)
    sl, sr = 1.0 / np.sqrt(m_left), 1.0 / np.sqrt(m_right)
    d01_l, d01_r = sl * sl * p01, sr * sr * p01
    d10_l = np.conj(np.swapaxes(d01_l, 1, 2))
    d10_r = np.conj(np.swapaxes(d01_r, 1, 2))
    lead_left = np.stack([sl * sl * p00, d10_l, d01_l], axis=1)
    lead_right = np.stack([sr * sr * p00, d01_r, d10_r], axis=1)
    # V_L = D(layer 0, lead layer -1) = M_0^-1/2 Phi_10 M_L^-1/2; V_R = M_{L-1}^-1/2 Phi_01 M_R^-1/2
    v_left = np.conj(np.swapaxes(p01, 1, 2)) * (sl / np.sqrt(masses[0, 0]))
    v_right = p01 * (sr / np.sqrt(masses[-1, 0]))
    # band top of the light side, over a coarse (kx, k) sample
    top = 0.0
    for s in range(0, len(kpoints), max(1, len(kpoints) // 16)):
        for kx in np.linspace(0.0, np.pi, 9):
            dk = lead_left[s, 0] + d01_l[s] * np.exp(1j * kx) + d10_l[s] * np.exp(-1j * kx)
            top = max(top, float(np.linalg.eigvalsh(0.5 * (dk + dk.conj().T)).max()))
    omegas = 1.1 * np.sqrt(top) * (np.arange(n_omega) + 0.5) / n_omega
    return Interface3DProblem(terms=terms, fc00=fc00, fc01=fc01, masses=masses,
                              lead_left=lead_left, lead_right=lead_right, v_left=v_left,
                              v_right=v_right, phases=ph, kpoints=kpoints, omegas=omegas,
                              eta=eta)


def interface3d_blocks(p: Interface3DProblem, fc00, fc01):
    """Device blocks of every k point from the layer force constants ``fc00`` (L, T, n, n)
    and ``fc01`` (L - 1, T, n, n) as a differentiable torch expression:
    D_l(k) = M_l^-1/2 sum_t Phi_l(t) e^{i k.t} M_l^-1/2 and the same between layers
    l and l + 1.  Returns (onsite (S, L, n, n), coupling (S, L - 1, n, n))."""
    import torch

    dev = fc00.device
    c128 = torch.complex128
    ph = torch.as_tensor(p.phases, dtype=c128, device=dev)
    inv = torch.as_tensor(1.0 / np.sqrt(p.masses), dtype=torch.float64, device=dev)
    d00 = torch.einsum("st,ltij->slij", ph, fc00.to(c128))
    d01 = torch.einsum("st,ltij->slij", ph, fc01.to(c128))
    ons = inv[None, :, :, None] * d00 * inv[None, :, None, :]
    cpl = inv[None, :-1, :, None] * d01 * inv[None, 1:, None, :]
    return ons, cpl


def interface3d_lead_blocks(p: Interface3DProblem, f00, f01):
    """Lead blocks and device-to-lead couplings of every k point from the bulk force
    constants ``f00`` / ``f01`` (n_terms, n, n) as a differentiable torch expression,
    the same construction as ``interface3d_problem``'s (interface contacts: the left
    lead faces left, V_L = M_0^-1/2 Phi_10 M_L^-1/2, V_R = M_{L-1}^-1/2 Phi_01 M_R^-1/2).
    Returns (lead_left, lead_right (S, 3, n, n), v_left, v_right (S, n, n))."""
    import torch

    dev = f00.device
    c128 = torch.complex128
    ph = torch.as_tensor(p.phases, dtype=c128, device=dev)
    m_l, m_r = float(p.masses[0, 0]), float(p.masses[-1, 0])
    p00 = torch.einsum("st,tij->sij", ph, f00.to(c128))
    p01 = torch.einsum("st,tij->sij", ph, f01.to(c128))
    p10 = p01.conj().transpose(1, 2)
    lead_left = torch.stack([p00 / m_l, p10 / m_l, p01 / m_l], dim=1)
    lead_right = torch.stack([p00 / m_r, p01 / m_r, p10 / m_r], dim=1)
    v_left = p10 / np.sqrt(m_l * m_l)
    v_right = p01 / np.sqrt(m_r * m_r)
    return lead_left, lead_right, v_left, v_right


# (cell, transverse k mesh) of the 3-D uneven batch: 24-, 48- and 96-DOF layers
# (conventional diamond cubes along x, y, z; k mesh per transverse direction)
UNEVEN3D_PARTS = (((1, 1, 1), 32), ((2, 1, 1), 8), ((2, 2, 1), 4))


@dataclass
class KspaceProblem:
    """A bulk force-constant model solved over a transverse k mesh."""

    params: object
    lead: object
    device: object
    kpoints: list
    omegas: np.ndarray
    eta: float
    wmax: float


def _kspace_params(phi, mass: float):
    from negfpy_amd.models.material_kspace import MaterialKspaceParams

    na = phi[(0, 0, 0)].shape[0] // 3
    f00 = {(dy, dz): v for (dx, dy, dz), v in phi.items() if dx == 0}
    f01 = {(dy, dz): v for (dx, dy, dz), v in phi.items() if dx == 1}
    return MaterialKspaceParams(masses=np.full(na, float(mass)), fc00_terms=f00,
                                fc01_terms=f01)


def cubic_params(cell, seed: int = 0, jitter: float = 0.05, mass: float = 28.0):
    """``MaterialKspaceParams`` of a simple-cubic cell (``_cubic_phi``, seeded bond
    jitter) with (dy, dz) terms for the same layer and the next one."""
    return _kspace_params(_cubic_phi(cell, seed, jitter), mass)


def diamond_params(cells, seed: int = 0, jitter: float = 0.05, mass: float = 28.0):
    """``MaterialKspaceParams`` of cx x cy x cz conventional silicon-like diamond cubes
    (``_diamond_phi``, seeded bond jitter), mass 28 amu."""
    return _kspace_params(_diamond_phi(cells, seed, jitter), mass)


def uneven3d_problems(seed: int = 0, n_layers: int = 8, n_omega: int = 1024,
                      edge_fraction: float = 0.25, eta: float = 1e-6) -> list[KspaceProblem]:
    """The 3-D uneven batch: silicon-like diamond cells of 8, 16 and 32 atoms (one, two
    and four conventional cubes: 24-, 48- and 96-DOF layers, mass 28 amu, first and
    second neighbour springs with 5 % seeded bond jitter) on 32 x 32, 8 x 8 and 4 x 4
    (ky, kz) meshes, n_layers device layers of the same cells and n_omega frequencies
    each; a quarter of each spectrum packed within 1 % of the band top.  The 24-DOF
    part alone is 1,048,576 instances."""
    from negfpy_amd.models.material_kspace import material_kspace_device, material_kspace_lead

    out = []
    for i, (cell, nk) in enumerate(UNEVEN3D_PARTS):
        params = diamond_params(cell, seed + i)
        build = material_kspace_lead(params).blocks_builder
        top = 0.0
        for ky in np.linspace(0.0, np.pi, 5):
            for kz in np.linspace(0.0, np.pi, 5):
                d00, d01, d10 = build((ky, kz))
                for kx in np.linspace(0.0, np.pi, 9):
                    dk = d00 + d01 * np.exp(1j * kx) + d10 * np.exp(-1j * kx)
                    top = max(top, float(np.linalg.eigvalsh(0.5 * (dk + dk.conj().T)).max()))
        wmax = float(np.sqrt(top))
        w = 1.1 * wmax * (np.arange(n_omega) + 0.5) / n_omega
        m = int(round(edge_fraction * n_omega))
        w[::max(n_omega // m, 1)][:m] = wmax * (1.0 + np.linspace(-0.01, 0.01, m))
        ks = [(-np.pi + 2 * np.pi * (a + 0.5) / nk, -np.pi + 2 * np.pi * (b + 0.5) / nk)
              for a in range(nk) for b in range(nk)]
        out.append(KspaceProblem(params=params, lead=material_kspace_lead(params),
                                 device=material_kspace_device(n_layers, params), kpoints=ks,
                                 omegas=np.sort(w), eta=eta, wmax=wmax))
    return out
