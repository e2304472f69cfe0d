"""Device Green's function and transmission on the GPU.

Same entry points, arguments, defaults, conventions and exceptions as the
reference ``src/negfpy/core/negf.py``:

* energy variable z = (omega + i eta)^2 for the leads and
  (omega + i eta_device)^2 for the device (negf.py:163-166, :245);
* default couplings V_L = d10_L (or d01_L^H) and V_R = d01_R, Sigma = V g V^H
  with both surface GFs from ``surface_gf(d00, d01, d10)`` (negf.py:192-201, :97);
* ``omega_scale`` divides omega, eta and every block by s, s^2; the default
  couplings are then divided by s^2 once more, exactly as negf.py:177-204 does;
* default contacts are the first / last n_c device DOFs (negf.py:208-217);
* T = Re Tr[Gamma_L G_LR Gamma_R G_LR^H] returned as a Python float (negf.py:356-357).

Where the reference factors the whole device with SuperLU per point, this
module packs the blocks of every k_par once into device tensors and solves all
(omega, k_par) points of a call in one batched launch sequence (decimation,
self-energies, block recursive Green's function sweep, FP64 trace).  Contacts
that reach beyond the first/last principal layer are handled by merging
layers into larger blocks.
"""

from __future__ import annotations

from collections.abc import Callable
from dataclasses import dataclass

import numpy as np

from . import _engine
from ._status import transmission_value
from .surface_gf import _prec, canonical_method, pack_lead, surface_route
from .types import Device1D, DeviceKSpace, KPar, LeadBlocks, LeadKSpace

Array = np.ndarray
LeadLike = LeadBlocks | LeadKSpace
DeviceLike = Device1D | DeviceKSpace
ContactIndices = slice | list[int] | tuple[int, ...] | Array | None
DeviceToLeadLike = Array | Callable[[KPar], Array]


def _resolve_lead_blocks(lead: LeadLike, kpar: KPar) -> LeadBlocks:
    """negf.py:22-32."""
    if isinstance(lead, LeadBlocks):
        return lead
    blocks = lead.blocks(kpar=kpar)
    if len(blocks) == 2:
        return LeadBlocks(d00=blocks[0], d01=blocks[1])
    if len(blocks) == 3:
        return LeadBlocks(d00=blocks[0], d01=blocks[1], d10=blocks[2])
    raise ValueError("LeadKSpace blocks_builder must return (d00,d01) or (d00,d01,d10).")


def _resolve_device(device: DeviceLike, kpar: KPar) -> Device1D:
    if isinstance(device, Device1D):
        return device
    return device.device(kpar=kpar)


def _resolve_coupling(coupling, kpar):
    if coupling is None:
        return None
    if callable(coupling):
        return np.asarray(coupling(kpar), dtype=np.complex128)
    return np.asarray(coupling, dtype=np.complex128)


def _normalize_contact_indices(indices: ContactIndices, *, dim: int, default: Array) -> Array:
    """negf.py:119-138."""
    if indices is None:
        idx = np.asarray(default, dtype=int).ravel()
    elif isinstance(indices, slice):
        idx = np.arange(dim, dtype=int)[indices]
    else:
        idx = np.asarray(indices, dtype=int).ravel()
    if idx.size == 0:
        raise ValueError("Contact indices must be non-empty.")
    if np.any(idx < 0) or np.any(idx >= dim):
        raise ValueError("Contact indices out of device range.")
    if np.unique(idx).size != idx.size:
        raise ValueError("Contact indices must be unique.")
    return idx


@dataclass
class KProblem:
    """Everything of one k_par point, resolved, scaled and laid out for the kernels."""

    lead_l: Array      # (3, nl, nl)
    lead_r: Array      # (3, nr, nr)
    v_l: Array         # (ncl, nl)
    v_r: Array         # (ncr, nr)
    pos_l: Array       # contact positions inside the first (merged) layer
    pos_r: Array       # contact positions inside the last (merged) layer
    onsite: Array      # (L', n', n')
    coupling: Array    # (L'-1, n', n')
    idx_l: Array       # contact indices in the original device numbering
    idx_r: Array
    scale: float       # omega_scale or 1
    dim: int

    def signature(self):
        return (self.onsite.shape, self.lead_l.shape, self.lead_r.shape, self.v_l.shape,
                self.v_r.shape, tuple(self.pos_l.tolist()), tuple(self.pos_r.tolist()))


def _merge_layers(onsite, coupling, m):
    """Group m consecutive layers into one block; pad the tail with decoupled zero DOFs."""
    nl = len(onsite)
    n = onsite[0].shape[0]
    nb = -(-nl // m)
    big = m * n
    ons = np.zeros((nb, big, big), dtype=np.complex128)
    cpl = np.zeros((max(nb - 1, 0), big, big), dtype=np.complex128)
    for i in range(nl):
        b, r = divmod(i, m)
        ons[b, r * n:(r + 1) * n, r * n:(r + 1) * n] = onsite[i]
    for i in range(nl - 1):
        b, r = divmod(i, m)
        blk = np.asarray(coupling[i], dtype=np.complex128)
        if r < m - 1:
            ons[b, r * n:(r + 1) * n, (r + 1) * n:(r + 2) * n] = blk
            ons[b, (r + 1) * n:(r + 2) * n, r * n:(r + 1) * n] = blk.conj().T
        else:
            cpl[b, (m - 1) * n:m * n, 0:n] = blk
    return ons, cpl


def _layout(onsite, coupling, idx_l, idx_r, max_block):
    nl = len(onsite)
    n = onsite[0].shape[0]
    dim = nl * n
    for m in range(1, nl + 1):
        nb = -(-nl // m)
        if m * n > max_block:
            break
        last0 = (nb - 1) * m * n
        if np.all(idx_l < m * n) and np.all(idx_r >= last0):
            if m == 1:
                ons = np.asarray(np.stack(onsite), dtype=np.complex128)
                cpl = (np.asarray(np.stack(coupling), dtype=np.complex128) if nl > 1
                       else np.zeros((0, n, n), dtype=np.complex128))
            else:
                ons, cpl = _merge_layers(onsite, coupling, m)
            return ons, cpl, idx_l.astype(np.int32), (idx_r - last0).astype(np.int32)
    raise NotImplementedError(
        f"contacts at {idx_l.tolist()} / {idx_r.tolist()} of a {dim}-DOF device need merged "
        f"blocks larger than the supported {max_block}")


def resolve_k(device, lead_left, lead_right, kpar, *, eta, eta_device, device_to_lead_left,
              device_to_lead_right, contact_left_indices, contact_right_indices, omega_scale,
              max_block) -> KProblem:
    """Resolve one k_par point like negf.py:141-243 (up to the self-energies)."""
    dev = _resolve_device(device, kpar)
    vl_in = _resolve_coupling(device_to_lead_left, kpar)
    vr_in = _resolve_coupling(device_to_lead_right, kpar)
    left = _resolve_lead_blocks(lead_left, kpar)
    right = _resolve_lead_blocks(lead_right, kpar)
    n = dev.dof_per_layer
    dim = dev.n_layers * n
    eta_dev = float(eta if eta_device is None else eta_device)
    if eta_dev < 0.0:
        raise ValueError("eta_device must be non-negative.")
    onsite = [np.asarray(b, dtype=np.complex128) for b in dev.onsite_blocks]
    coupling = [np.asarray(b, dtype=np.complex128) for b in dev.coupling_blocks]
    ld = [np.asarray(left.d00, dtype=np.complex128), np.asarray(left.d01, dtype=np.complex128),
          None if left.d10 is None else np.asarray(left.d10, dtype=np.complex128)]
    rd = [np.asarray(right.d00, dtype=np.complex128), np.asarray(right.d01, dtype=np.complex128),
          None if right.d10 is None else np.asarray(right.d10, dtype=np.complex128)]
    scale = 1.0
    if omega_scale is not None:
        scale = float(omega_scale)
        if scale <= 0.0:
            raise ValueError("omega_scale must be positive when provided.")
        s2 = scale * scale
        onsite = [b / s2 for b in onsite]
        coupling = [b / s2 for b in coupling]
        ld = [None if b is None else b / s2 for b in ld]
        rd = [None if b is None else b / s2 for b in rd]
    vl = (ld[2] if ld[2] is not None else ld[1].conj().T) if vl_in is None else vl_in
    vr = rd[1] if vr_in is None else vr_in
    if omega_scale is not None:
        vl = vl / (scale * scale)
        vr = vr / (scale * scale)
    ncl, ncr = vl.shape[0], vr.shape[0]
    idx_l = _normalize_contact_indices(contact_left_indices, dim=dim, default=np.arange(ncl))
    idx_r = _normalize_contact_indices(contact_right_indices, dim=dim,
                                       default=np.arange(dim - ncr, dim))
    if idx_l.size != ncl:
        raise ValueError("Left contact index count must match device_to_lead_left rows "
                         f"(got {idx_l.size}, expected {ncl}).")
    if idx_r.size != ncr:
        raise ValueError("Right contact index count must match device_to_lead_right rows "
                         f"(got {idx_r.size}, expected {ncr}).")
    for v, lead in ((vl, ld), (vr, rd)):
        if v.ndim != 2:
            raise ValueError("device_to_lead coupling must be a 2D array.")
        if v.shape[1] != lead[0].shape[0]:
            raise ValueError("device_to_lead coupling has incompatible lead dimension "
                             f"(got {v.shape[1]}, expected {lead[0].shape[0]}).")
    ons, cpl, pos_l, pos_r = _layout(onsite, coupling, idx_l, idx_r, max_block)
    return KProblem(lead_l=pack_lead(*ld), lead_r=pack_lead(*rd),
                    v_l=np.ascontiguousarray(vl, dtype=np.complex128),
                    v_r=np.ascontiguousarray(vr, dtype=np.complex128), pos_l=pos_l, pos_r=pos_r,
                    onsite=ons, coupling=cpl, idx_l=idx_l, idx_r=idx_r, scale=scale, dim=dim)


def point_jobs(kps: list[KProblem], k_of: Array, omega: Array, eta: Array, eta_dev: Array,
               *, precision: str, tol: float = 1e-12, maxiter: int = 200, route: int = 0):
    """Group instances by block layout: one ``(solver, set_of, omega, eta, eta_dev)`` job
    per layout (for ``TransmissionSolver.run`` or ``_engine.transmission_multi``) and
    the instance indices each job covers."""
    k_of = np.asarray(k_of, dtype=np.int64)
    omega = np.asarray(omega, dtype=float)
    eta = np.asarray(eta, dtype=float)
    eta_dev = np.asarray(eta_dev, dtype=float)
    groups: dict = {}
    for ki, kp in enumerate(kps):
        groups.setdefault(kp.signature(), []).append(ki)
    jobs, sels = [], []
    for members in groups.values():
        sel = np.nonzero(np.isin(k_of, members))[0]
        if sel.size == 0:
            continue
        local = {k: j for j, k in enumerate(members)}
        kg = [kps[k] for k in members]
        lead_l = np.stack([kp.lead_l for kp in kg])
        lead_r = np.stack([kp.lead_r for kp in kg])
        share = bool(np.array_equal(lead_l, lead_r))
        sets = _engine.TransportSets.from_arrays(
            lead_l, lead_r, np.stack([kp.v_l for kp in kg]), np.stack([kp.v_r for kp in kg]),
            kg[0].pos_l, kg[0].pos_r, np.stack([kp.onsite for kp in kg]),
            np.stack([kp.coupling for kp in kg]) if kg[0].onsite.shape[0] > 1 else None, share)
        scale = np.array([kps[k].scale for k in k_of[sel]])
        solver = _engine.TransmissionSolver(sets, precision=precision, tol=tol, maxiter=maxiter,
                                            surface_route=route)
        jobs.append((solver, np.array([local[k] for k in k_of[sel]], dtype=np.int32),
                     omega[sel] / scale, eta[sel] / scale, eta_dev[sel] / scale))
        sels.append(sel)
    return jobs, sels


def run_jobs(jobs, precision: str):
    """Run transmission jobs: several FP64 jobs (different block layouts) in one
    persistent launch, otherwise one staged solve per job."""
    if len(jobs) > 1 and precision == "fp64":
        return _engine.transmission_multi(jobs)
    return [job[0].run(*job[1:]) for job in jobs]


def solve_points(kps: list[KProblem], k_of: Array, omega: Array, eta: Array, eta_dev: Array,
                 *, precision: str, tol: float = 1e-12, maxiter: int = 200, route: int = 0):
    """Transmission of every instance (k_of[i] indexes ``kps``).  Returns (T, status) arrays.

    Instances are grouped by block layout.  One group is one staged batched solve;
    several groups (k points whose block sizes or contact layouts differ) in FP64 go
    through one persistent launch together (``_engine.transmission_multi``).
    """
    omega = np.asarray(omega, dtype=float)
    t_all = np.full(omega.shape, np.nan)
    st_all = np.zeros(omega.shape, dtype=np.int32)
    jobs, sels = point_jobs(kps, k_of, omega, eta, eta_dev, precision=precision, tol=tol,
                            maxiter=maxiter, route=route)
    for sel, res in zip(sels, run_jobs(jobs, precision)):
        t_all[sel] = res.t.cpu().numpy()
        st_all[sel] = res.status.cpu().numpy()
    return t_all, st_all


def _max_block(precision):
    from negfpy_amd import _lib

    return int(_lib.load().negf_max_block(_engine.precision_code(precision)))


def _max_block_green():
    """Largest block of the full Green's function and adjoint kernels."""
    from negfpy_amd import _lib

    return int(_lib.load().negf_max_block_green())


def _resolve_many(device, lead_left, lead_right, kpoints, *, eta, eta_device, device_to_lead_left,
                  device_to_lead_right, contact_left_indices, contact_right_indices, omega_scale,
                  precision):
    mb = _max_block(precision)
    return [resolve_k(device, lead_left, lead_right, k, eta=eta, eta_device=eta_device,
                      device_to_lead_left=device_to_lead_left,
                      device_to_lead_right=device_to_lead_right,
                      contact_left_indices=contact_left_indices,
                      contact_right_indices=contact_right_indices, omega_scale=omega_scale,
                      max_block=mb) for k in kpoints]


def device_green_function(omega: float, device: DeviceLike, lead_left: LeadLike,
                          lead_right: LeadLike, eta: float = 1e-8, eta_device: float | None = None,
                          kpar: KPar = None, device_to_lead_left: DeviceToLeadLike | None = None,
                          device_to_lead_right: DeviceToLeadLike | None = None,
                          contact_left_indices: ContactIndices = None,
                          contact_right_indices: ContactIndices = None,
                          surface_gf_method: str = "sancho_rubio",
                          omega_scale: float | None = None) -> tuple[Array, Array, Array]:
    """(G, Sigma_L, Sigma_R) of the finite device, each dim x dim (negf.py:253-303).

    G is every block of A^-1 from the GPU block sweep (FP64); the self-energies
    are embedded at the contact indices like _embed_self_energies (negf.py:100-112).
    """
    from ._status import raise_for_lead_status

    route = surface_route(surface_gf_method)
    kp = resolve_k(device, lead_left, lead_right, kpar, eta=eta, eta_device=eta_device,
                   device_to_lead_left=device_to_lead_left,
                   device_to_lead_right=device_to_lead_right,
                   contact_left_indices=contact_left_indices,
                   contact_right_indices=contact_right_indices, omega_scale=omega_scale,
                   max_block=_max_block_green())
    sets = _engine.TransportSets.from_arrays(
        kp.lead_l[None], kp.lead_r[None], kp.v_l[None], kp.v_r[None], kp.pos_l, kp.pos_r,
        kp.onsite[None], kp.coupling[None] if kp.onsite.shape[0] > 1 else None, False)
    eta_dev = float(eta if eta_device is None else eta_device)
    g, sl, sr, lst, dst = _engine.device_green(sets, [0], [float(omega) / kp.scale],
                                               [float(eta) / kp.scale], [eta_dev / kp.scale],
                                               route=route)
    raise_for_lead_status(int(lst[0].item()))
    code = int(dst[0].item())
    if code == 13:
        raise RuntimeError("Factor is exactly singular")
    dim = kp.dim
    gm = g[0].cpu().numpy()[:dim, :dim]
    if code == 12:
        gm = np.full((dim, dim), np.nan, dtype=np.complex128)
    sig_l = np.zeros((dim, dim), dtype=np.complex128)
    sig_r = np.zeros((dim, dim), dtype=np.complex128)
    sig_l[np.ix_(kp.idx_l, kp.idx_l)] = sl[0].cpu().numpy()[np.ix_(kp.pos_l, kp.pos_l)]
    sig_r[np.ix_(kp.idx_r, kp.idx_r)] = sr[0].cpu().numpy()[np.ix_(kp.pos_r, kp.pos_r)]
    return gm, sig_l, sig_r


def transmission(omega: float, device: DeviceLike, lead_left: LeadLike, lead_right: LeadLike,
                 eta: float = 1e-8, eta_device: float | None = None, kpar: KPar = None,
                 device_to_lead_left: DeviceToLeadLike | None = None,
                 device_to_lead_right: DeviceToLeadLike | None = None,
                 contact_left_indices: ContactIndices = None,
                 contact_right_indices: ContactIndices = None,
                 surface_gf_method: str = "sancho_rubio", omega_scale: float | None = None, *,
                 precision: str | None = None) -> float:
    """Coherent phonon transmission T(omega) (negf.py:306-357)."""
    route = surface_route(surface_gf_method)
    prec = _prec(precision)
    kp = _resolve_many(device, lead_left, lead_right, [kpar], eta=eta, eta_device=eta_device,
                       device_to_lead_left=device_to_lead_left,
                       device_to_lead_right=device_to_lead_right,
                       contact_left_indices=contact_left_indices,
                       contact_right_indices=contact_right_indices, omega_scale=omega_scale,
                       precision=prec)
    eta_dev = float(eta if eta_device is None else eta_device)
    t, st = solve_points(kp, [0], [float(omega)], [float(eta)], [eta_dev], precision=prec,
                         route=route)
    return transmission_value(int(st[0]), float(t[0]))


def kspace_sets(device, lead_left, lead_right, kpoints, *, device_to_lead_left=None,
                device_to_lead_right=None, contact_left_indices=None, contact_right_indices=None,
                omega_scale=None, precision: str = "fp64"):
    """TransportSets of every k point built on the GPU in one launch per model, when the
    leads and the device are force-constant models (``material_kspace_lead`` /
    ``material_kspace_device`` of ``negfpy_amd.models``, or the interface contacts of
    ``negfpy_amd.modeling.interface_contacts``) with default contacts and no
    ``omega_scale``; otherwise None (the caller resolves per k on the host, as the
    reference does)."""
    bl = getattr(lead_left, "blocks_builder", None)
    br = getattr(lead_right, "blocks_builder", None)
    tl = getattr(bl, "material_kspace", None)
    tr = getattr(br, "material_kspace", None)
    td = getattr(getattr(device, "device_builder", None), "material_kspace", None)
    if tl is None or tr is None or td is None:
        return None
    if any(x is not None for x in (contact_left_indices, contact_right_indices, omega_scale)):
        return None
    cl = cr = None
    for v, side in ((device_to_lead_left, "l"), (device_to_lead_right, "r")):
        if v is None:
            continue
        tag = getattr(v, "material_kspace_contact", None)
        if tag is None or (tag[1] is not tl and tag[1] is not tr):
            return None
        if side == "l":
            cl = tag
        else:
            cr = tag
    params, n_layers, layer_masses = td
    n = int(params.ndof)
    if int(tl.ndof) != n or int(tr.ndof) != n or n > _max_block(precision):
        return None
    from .kspace import kspace_blocks

    def lead_blocks(p, builder):
        lb = kspace_blocks(p, kpoints, lead=True).lead
        if getattr(builder, "d10_is_d01_h", False):
            lb[:, 2] = lb[:, 1].conj().transpose(-1, -2)
        return lb

    kb = kspace_blocks(params, kpoints, lead=False, n_layers=n_layers, layer_masses=layer_masses)
    share = tl is tr and bool(getattr(bl, "d10_is_d01_h", False)) == bool(
        getattr(br, "d10_is_d01_h", False))
    kl = lead_blocks(tl, bl)
    kr = kl if share else lead_blocks(tr, br)

    def coupling(tag, default):
        if tag is None:
            return default
        src = (kl if tag[1] is tl else kr)[:, 1]
        return src if tag[0] == "d01" else src.conj().transpose(-1, -2).contiguous()

    idx = np.arange(n)
    # default couplings: V_L = d10 of the left lead, V_R = d01 of the right lead
    return _engine.TransportSets.from_arrays(
        kl, kr, coupling(cl, kl[:, 2]), coupling(cr, kr[:, 1]), idx, idx, kb.onsite,
        kb.coupling if n_layers > 1 else None, share)


def _solve_sets(sets, omega, eta, eta_dev, precision, route=0):
    """(T, status) of instance i = k point i of ``sets``."""
    nk = sets.n_sets
    om = np.broadcast_to(np.asarray(omega, dtype=float), (nk,))
    solver = _engine.TransmissionSolver(sets, precision=precision, surface_route=route)
    res = solver.run(np.arange(nk, dtype=np.int32), om, np.full(nk, float(eta)),
                     np.full(nk, float(eta_dev)))
    return res.t.cpu().numpy(), res.status.cpu().numpy()


def transmission_kavg(omega: float, device: DeviceLike, lead_left: LeadLike, lead_right: LeadLike,
                      kpoints: list[tuple[float, ...]], eta: float = 1e-8,
                      eta_device: float | None = None,
                      device_to_lead_left: DeviceToLeadLike | None = None,
                      device_to_lead_right: DeviceToLeadLike | None = None,
                      contact_left_indices: ContactIndices = None,
                      contact_right_indices: ContactIndices = None,
                      surface_gf_method: str = "sancho_rubio", omega_scale: float | None = None,
                      *, precision: str | None = None) -> float:
    """k_par-averaged transmission (negf.py:360-397): all k points in one batch."""
    if len(kpoints) == 0:
        raise ValueError("kpoints must contain at least one k-point.")
    route = surface_route(surface_gf_method)
    prec = _prec(precision)
    eta_dev = float(eta if eta_device is None else eta_device)
    if eta_dev < 0.0:
        raise ValueError("eta_device must be non-negative.")
    sets = kspace_sets(device, lead_left, lead_right, list(kpoints),
                       device_to_lead_left=device_to_lead_left,
                       device_to_lead_right=device_to_lead_right,
                       contact_left_indices=contact_left_indices,
                       contact_right_indices=contact_right_indices, omega_scale=omega_scale,
                       precision=prec)
    if sets is not None:
        nk = sets.n_sets
        t, st = _solve_sets(sets, float(omega), eta, eta_dev, prec, route)
    else:
        kps = _resolve_many(device, lead_left, lead_right, list(kpoints), eta=eta,
                            eta_device=eta_device, device_to_lead_left=device_to_lead_left,
                            device_to_lead_right=device_to_lead_right,
                            contact_left_indices=contact_left_indices,
                            contact_right_indices=contact_right_indices,
                            omega_scale=omega_scale, precision=prec)
        nk = len(kps)
        t, st = solve_points(kps, np.arange(nk), np.full(nk, float(omega)),
                             np.full(nk, float(eta)), np.full(nk, eta_dev), precision=prec,
                             route=route)
    vals = [transmission_value(int(st[i]), float(t[i])) for i in range(nk)]
    return float(np.mean(vals))


def _kpar_tuple(kpar):
    return tuple(float(x) for x in kpar) if kpar is not None else ()


def _validate_adaptive(kpoints, eta_values, min_success_fraction, nonnegative_tolerance,
                       max_channel_factor):
    if len(kpoints) == 0:
        raise ValueError("kpoints must contain at least one k-point.")
    if len(eta_values) == 0:
        raise ValueError("eta_values must contain at least one eta value.")
    if not (0.0 <= min_success_fraction <= 1.0):
        raise ValueError("min_success_fraction must be in [0, 1].")
    if nonnegative_tolerance < 0.0:
        raise ValueError("nonnegative_tolerance must be non-negative.")
    if max_channel_factor is not None and max_channel_factor <= 0.0:
        raise ValueError("max_channel_factor must be positive when provided.")


def _classify(code, t, nonneg_tol, t_max):
    """Outcome of one (k, eta) attempt as the adaptive loops of negf.py:446-514 see it:
    ("ok", value) or (reason, t_or_None)."""
    try:
        tval = transmission_value(int(code), float(t))
    except Exception:
        return "exception", None
    if not np.isfinite(tval):
        return "non_finite", tval
    if tval < -float(nonneg_tol):
        return "negative", tval
    if t_max is not None and tval > t_max:
        return "over_channel_limit", tval
    return "ok", float(max(tval, 0.0))


def _resolve_each(device, lead_left, lead_right, kpoints, *, eta_device, precision, **kw):
    """Resolve every k point on its own; a k point whose resolution raises gets None
    (each of its attempts then counts as an "exception", as in negf.py:463/612)."""
    mb = _max_block(precision)
    out = []
    for k in kpoints:
        try:
            # eta only enters resolution through the eta_device >= 0 check, made per rung
            out.append(resolve_k(device, lead_left, lead_right, k, eta=1.0, eta_device=eta_device,
                                 max_block=mb, **kw))
        except NotImplementedError:
            raise     # a block size beyond the kernels is a limit of this build, not a failure
        except Exception:
            out.append(None)
    return out


class _KSolver:
    """Solves (k point, omega, eta, eta_device) instances of one k-point list in batches:
    on blocks built on the GPU for force-constant models, else on k points resolved on
    the host (a k point whose resolution raised is None: each of its attempts counts as
    an "exception", as in negf.py:463/612)."""

    def __init__(self, device, lead_left, lead_right, kpoints, *, eta_device, precision,
                 route=0, **kw):
        self.precision = precision
        self.route = int(route)
        self.kpoints = list(kpoints)
        self.sets = kspace_sets(device, lead_left, lead_right, self.kpoints, precision=precision,
                                **kw)
        self.kps = None
        if self.sets is None:
            self.kps = _resolve_each(device, lead_left, lead_right, self.kpoints,
                                     eta_device=eta_device, precision=precision, **kw)

    def live(self, k) -> bool:
        return self.sets is not None or self.kps[k] is not None

    def solve(self, k_idx, omega, eta, eta_dev):
        k_idx = np.asarray(k_idx, dtype=np.int64)
        if k_idx.size == 0:
            return np.zeros(0), np.zeros(0, dtype=np.int32)
        if self.sets is not None:
            res = _engine.TransmissionSolver(self.sets, precision=self.precision,
                                             surface_route=self.route).run(
                k_idx.astype(np.int32), omega, eta, eta_dev)
            return res.t.cpu().numpy(), res.status.cpu().numpy()
        uniq, local = np.unique(k_idx, return_inverse=True)
        return solve_points([self.kps[k] for k in uniq], local, omega, eta, eta_dev,
                            precision=self.precision, route=self.route)

    # Per-k eta ladders walked inside one persistent launch (set_eta_ladders_in_kernel)
    # instead of one batched solve per rung.  Off by default: on one MI355X the
    # rung-by-rung batches were faster (0.64x-0.92x for the in-launch form on
    # spectra, 0.82x-0.88x on single-omega calls), since a ticket's instances wait in
    # lockstep for the slowest one's rungs.
    in_kernel_ladders = False

    def ladder_ok(self, ladders, eta_device) -> bool:
        """Whether the eta ladders run inside the persistent launch: switched on, FP64,
        the decimation route, and no rung with a negative device eta (such a rung is
        an "exception" of its own, decided on the host)."""
        if not self.in_kernel_ladders or self.precision != "fp64" or self.route != 0:
            return False
        if eta_device is None:
            return all(float(e) >= 0.0 for lad in ladders for e in lad)
        return float(eta_device) >= 0.0

    def solve_ladder(self, k_idx, omega, ladders, eta_device, nonneg_tol, t_max):
        """Instance i (k point k_idx[i], omega[i]) tries the etas ladders[i] in order
        inside one persistent launch until one is accepted (status 0, finite T,
        T >= -nonneg_tol, T <= t_max[i]).  Returns (accepted rung or the rung count,
        T of every attempted rung, status of every attempted rung (-1: not tried))."""
        k_idx = np.asarray(k_idx, dtype=np.int64)
        omega = np.asarray(omega, dtype=float)
        n = k_idx.size
        max_rungs = max((len(lad) for lad in ladders), default=1)
        rung = np.zeros(n, dtype=np.int64)
        t_r = np.full((n, max_rungs), np.nan)
        st_r = np.full((n, max_rungs), -1, dtype=np.int32)
        if n == 0:
            return rung, t_r, st_r
        fixed = eta_device is not None
        ed = np.full(n, float(eta_device) if fixed else 0.0)
        if self.sets is not None:
            lad = _engine.EtaLadder(ladders, t_max=t_max, nonneg_tol=nonneg_tol,
                                    eta_dev_fixed=fixed)
            jobs = [(_engine.TransmissionSolver(self.sets, precision=self.precision),
                     k_idx.astype(np.int32), omega, None, ed, lad)]
            sels = [np.arange(n)]
        else:
            uniq, local = np.unique(k_idx, return_inverse=True)
            kps = [self.kps[k] for k in uniq]
            jobs, sels = point_jobs(kps, local, omega, np.zeros(n), ed, precision=self.precision,
                                    route=self.route)
            # omega_scale divides every eta of the point as it divides omega
            scale = np.array([kps[q].scale for q in local])
            jobs = [job + (_engine.EtaLadder([np.asarray(ladders[j], dtype=float) / scale[j]
                                              for j in sel],
                                             t_max=[t_max[j] for j in sel],
                                             nonneg_tol=nonneg_tol, eta_dev_fixed=fixed),)
                    for job, sel in zip(jobs, sels)]
        for sel, res in zip(sels, _engine.transmission_multi(jobs)):
            rung[sel] = res.ladder.rung.cpu().numpy()
            lt = res.ladder.t.cpu().numpy()
            t_r[sel, :lt.shape[1]] = lt
            st_r[sel, :lt.shape[1]] = res.ladder.status.cpu().numpy()
        return rung, t_r, st_r


def set_eta_ladders_in_kernel(flag: bool) -> None:
    """Per-k adaptive eta (``transmission_kavg_adaptive*``, ``eta_scheme="adaptive"``):
    True walks every point's eta ladder inside one persistent launch (FP64, decimation
    route; one launch, e.g. for capture in a hipGraph), False (default) solves one
    batch per rung.  Both give identical values, statistics and rejection records."""
    _KSolver.in_kernel_ladders = bool(flag)


def _t_max(lead_left, kpoints, max_channel_factor):
    if max_channel_factor is None:
        return [None] * len(kpoints)
    tagged = getattr(getattr(lead_left, "blocks_builder", None), "material_kspace", None)
    if tagged is not None:
        return [float(max_channel_factor) * float(tagged.ndof)] * len(kpoints)
    return [float(max_channel_factor) * float(_resolve_lead_blocks(lead_left, k).d00.shape[0])
            for k in kpoints]


def _adaptive_batch(omegas, solver, t_max, ladders, *, eta_device, global_eta,
                    min_success_fraction, nonnegative_tolerance, collect_rejected):
    """The per-k (negf.py:400-543) or global (:546-688) eta ladder for every omega of
    ``omegas`` at once: rung r solves all pending (omega, k) instances of all omegas in
    one batch.  ``ladders[i]`` is the eta tuple of omega i.  Returns, per omega, either
    ("ok", mean T, info) or ("error", message) with the reference's messages."""
    nw = len(omegas)
    nk = len(solver.kpoints)
    kk = np.arange(nk)
    outcome = [[None] * nk for _ in range(nw)]     # per-k: accepted value
    used_eta = [[None] * nk for _ in range(nw)]
    rejected = [[[] for _ in range(nk)] for _ in range(nw)]   # per-k: by k, then eta
    pending = [list(kk) for _ in range(nw)]
    done = [False] * nw                             # global: eta accepted
    g_rejected = [[] for _ in range(nw)]
    g_result = [None] * nw
    max_rungs = max((len(lad) for lad in ladders), default=0)
    if not global_eta and solver.ladder_ok(ladders, eta_device):
        # per-k ladders: every (omega, k) walks its ladder inside one persistent launch;
        # the attempts it reports are classified here exactly as the rung loop below would
        inst_w = np.repeat(np.arange(nw), nk)
        inst_k = np.tile(kk, nw)
        live = np.array([solver.live(int(k)) for k in inst_k], dtype=bool)
        sel = np.nonzero(live)[0]
        rung, t_r, st_r = solver.solve_ladder(
            inst_k[sel], np.asarray(omegas, float)[inst_w[sel]],
            [ladders[int(inst_w[j])] for j in sel], eta_device, nonnegative_tolerance,
            [t_max[int(inst_k[j])] for j in sel])
        pos = {int(j): q for q, j in enumerate(sel)}
        for j in range(inst_w.size):
            i, k = int(inst_w[j]), int(inst_k[j])
            lad = ladders[i]
            q = pos.get(j)
            tried = len(lad) if q is None else min(int(rung[q]) + 1, len(lad))
            for r in range(tried):
                reason, val = (("exception", None) if q is None else
                               _classify(st_r[q, r], t_r[q, r], nonnegative_tolerance, t_max[k]))
                if reason == "ok":
                    outcome[i][k] = val
                    used_eta[i][k] = float(lad[r])
                elif collect_rejected:
                    rejected[i][k].append({"omega": float(omegas[i]),
                                           "kpar": _kpar_tuple(solver.kpoints[k]),
                                           "eta": float(lad[r]),
                                           "t": None if val is None else float(val),
                                           "reason": reason})
        max_rungs = 0   # nothing left for the rung loop
    for r in range(max_rungs):
        inst_w, inst_k, inst_eta = [], [], []
        for i in range(nw):
            if r >= len(ladders[i]) or (global_eta and done[i]):
                continue
            ks = pending[i] if not global_eta else list(kk)
            for k in ks:
                inst_w.append(i)
                inst_k.append(k)
                inst_eta.append(float(ladders[i][r]))
        if not inst_w:
            continue
        inst_w = np.asarray(inst_w)
        inst_k = np.asarray(inst_k)
        inst_eta = np.asarray(inst_eta)
        inst_ed = inst_eta if eta_device is None else np.full(inst_eta.size, float(eta_device))
        live = np.array([solver.live(k) for k in inst_k]) & (inst_ed >= 0.0)
        t = np.full(inst_w.size, np.nan)
        st = np.zeros(inst_w.size, dtype=np.int32)
        sel = np.nonzero(live)[0]
        if sel.size:
            t[sel], st[sel] = solver.solve(inst_k[sel], np.asarray(omegas, float)[inst_w[sel]],
                                           inst_eta[sel], inst_ed[sel])
        outcomes = [_classify(st[j], t[j], nonnegative_tolerance, t_max[inst_k[j]])
                    if live[j] else ("exception", None) for j in range(inst_w.size)]
        if not global_eta:
            for j, (reason, val) in enumerate(outcomes):
                i, k = int(inst_w[j]), int(inst_k[j])
                if reason == "ok":
                    outcome[i][k] = val
                    used_eta[i][k] = float(inst_eta[j])
                else:
                    if collect_rejected:
                        rejected[i][k].append({"omega": float(omegas[i]),
                                               "kpar": _kpar_tuple(solver.kpoints[k]),
                                               "eta": float(inst_eta[j]),
                                               "t": None if val is None else float(val),
                                               "reason": reason})
            for i in range(nw):
                pending[i] = [k for k in pending[i] if outcome[i][k] is None]
        else:
            by_w: dict = {}
            for j, oc in enumerate(outcomes):
                by_w.setdefault(int(inst_w[j]), []).append((int(inst_k[j]), oc))
            for i, lst in by_w.items():
                eta = float(ladders[i][r])
                vals = [v for _, (reason, v) in lst if reason == "ok"]
                if collect_rejected:
                    g_rejected[i].extend({"omega": float(omegas[i]),
                                          "kpar": _kpar_tuple(solver.kpoints[k]), "eta": eta,
                                          "t": None if v is None else float(v), "reason": reason}
                                         for k, (reason, v) in lst if reason != "ok")
                n_success = len(vals)
                frac = float(n_success / nk)
                if n_success > 0 and frac >= min_success_fraction:
                    info = {"n_total": int(nk), "n_success": int(n_success),
                            "n_failed": int(nk - n_success), "success_fraction": frac,
                            "eta_selected": eta,
                            "eta_histogram": {float(e): (int(n_success) if float(e) == eta else 0)
                                              for e in ladders[i]}}
                    if collect_rejected:
                        info["n_rejected"] = int(len(g_rejected[i]))
                        info["rejected"] = list(g_rejected[i])
                    g_result[i] = ("ok", float(np.mean(vals)), info, np.asarray(vals))
                    done[i] = True
    results = []
    for i in range(nw):
        if global_eta:
            results.append(g_result[i] if g_result[i] is not None else (
                "error", "No eta value satisfied global adaptive k-averaging quality threshold "
                f"(required success_fraction >= {min_success_fraction:.3f})."))
            continue
        vals = [v for v in outcome[i] if v is not None]
        n_success = len(vals)
        frac = float(n_success / nk)
        if n_success == 0:
            results.append(("error", "No k-point converged for adaptive k-averaged transmission."))
            continue
        if frac < min_success_fraction:
            results.append(("error", "Adaptive k-averaged transmission failed success-fraction "
                            f"threshold (got {frac:.3f}, required >= {min_success_fraction:.3f})."))
            continue
        hist = {float(e): int(sum(1 for u in used_eta[i] if u == float(e))) for e in ladders[i]}
        info = {"n_total": int(nk), "n_success": int(n_success), "n_failed": int(nk - n_success),
                "success_fraction": frac, "eta_histogram": hist}
        if collect_rejected:
            flat = [rec for rs in rejected[i] for rec in rs]
            info["n_rejected"] = int(len(flat))
            info["rejected"] = flat
        results.append(("ok", float(np.mean(vals)), info))
    return results


def _adaptive_single(omega, device, lead_left, lead_right, kpoints, eta_values, eta_device,
                     min_success_fraction, device_to_lead_left, device_to_lead_right,
                     contact_left_indices, contact_right_indices, surface_gf_method, omega_scale,
                     nonnegative_tolerance, max_channel_factor, collect_rejected, precision,
                     global_eta):
    _validate_adaptive(kpoints, eta_values, min_success_fraction, nonnegative_tolerance,
                       max_channel_factor)
    canonical_method(surface_gf_method)
    kpoints = list(kpoints)
    prec = _prec(precision)
    t_max = _t_max(lead_left, kpoints, max_channel_factor)
    solver = _KSolver(device, lead_left, lead_right, kpoints, eta_device=eta_device,
                      route=surface_route(surface_gf_method), precision=prec, device_to_lead_left=device_to_lead_left,
                      device_to_lead_right=device_to_lead_right,
                      contact_left_indices=contact_left_indices,
                      contact_right_indices=contact_right_indices, omega_scale=omega_scale)
    res = _adaptive_batch([float(omega)], solver, t_max, [tuple(eta_values)],
                          eta_device=eta_device, global_eta=global_eta,
                          min_success_fraction=min_success_fraction,
                          nonnegative_tolerance=nonnegative_tolerance,
                          collect_rejected=collect_rejected)[0]
    if res[0] == "error":
        raise RuntimeError(res[1])
    return res[1], res[2]


def transmission_kavg_adaptive(omega: float, device: DeviceLike, lead_left: LeadLike,
                               lead_right: LeadLike, kpoints: list[tuple[float, ...]],
                               eta_values: tuple[float, ...] = (1e-8, 1e-7, 1e-6, 1e-5),
                               eta_device: float | None = None, min_success_fraction: float = 0.0,
                               device_to_lead_left: DeviceToLeadLike | None = None,
                               device_to_lead_right: DeviceToLeadLike | None = None,
                               contact_left_indices: ContactIndices = None,
                               contact_right_indices: ContactIndices = None,
                               surface_gf_method: str = "sancho_rubio",
                               omega_scale: float | None = None,
                               nonnegative_tolerance: float = 0.0,
                               max_channel_factor: float | None = None,
                               collect_rejected: bool = False, *, precision: str | None = None
                               ) -> tuple[float, dict[str, object]]:
    """Per-k eta ladder (negf.py:400-543).  Each rung is one batch over the k points
    that have not yet succeeded; outcomes, filters and statistics follow the reference."""
    return _adaptive_single(omega, device, lead_left, lead_right, kpoints, eta_values,
                            eta_device, min_success_fraction, device_to_lead_left,
                            device_to_lead_right, contact_left_indices, contact_right_indices,
                            surface_gf_method, omega_scale, nonnegative_tolerance,
                            max_channel_factor, collect_rejected, precision, False)


def transmission_kavg_adaptive_global_eta(
        omega: float, device: DeviceLike, lead_left: LeadLike, lead_right: LeadLike,
        kpoints: list[tuple[float, ...]], eta_values: tuple[float, ...] = (1e-8, 1e-7, 1e-6, 1e-5),
        eta_device: float | None = None, min_success_fraction: float = 0.0,
        device_to_lead_left: DeviceToLeadLike | None = None,
        device_to_lead_right: DeviceToLeadLike | None = None,
        contact_left_indices: ContactIndices = None, contact_right_indices: ContactIndices = None,
        surface_gf_method: str = "sancho_rubio", omega_scale: float | None = None,
        nonnegative_tolerance: float = 0.0, max_channel_factor: float | None = None,
        collect_rejected: bool = False, *, precision: str | None = None
) -> tuple[float, dict[str, object]]:
    """One eta per omega for all k points (negf.py:546-688); each rung is one batch."""
    return _adaptive_single(omega, device, lead_left, lead_right, kpoints, eta_values,
                            eta_device, min_success_fraction, device_to_lead_left,
                            device_to_lead_right, contact_left_indices, contact_right_indices,
                            surface_gf_method, omega_scale, nonnegative_tolerance,
                            max_channel_factor, collect_rejected, precision, True)


def transmission_kavg_adaptive_spectrum(
        omegas, device: DeviceLike, lead_left: LeadLike, lead_right: LeadLike,
        kpoints, eta_values=(1e-8, 1e-7, 1e-6, 1e-5), *, eta_scheme: str = "adaptive",
        eta_fixed: float = 1e-4, eta_device: float | None = None,
        min_success_fraction: float = 0.0, device_to_lead_left: DeviceToLeadLike | None = None,
        device_to_lead_right: DeviceToLeadLike | None = None,
        contact_left_indices: ContactIndices = None, contact_right_indices: ContactIndices = None,
        surface_gf_method: str = "sancho_rubio", omega_scale: float | None = None,
        nonnegative_tolerance: float = 0.0, max_channel_factor: float | None = None,
        collect_rejected: bool = False, precision: str | None = None):
    """k-averaged adaptive-eta transmission of a whole spectrum, all omegas at once.

    The batched form of the reference's per-omega spectrum driver
    (workflows/ifc_bulk.py:951-1029, :1059-1288): ``eta_scheme`` "adaptive" applies
    ``transmission_kavg_adaptive`` to every omega, "adaptive_global" and "fixed" (the
    single eta ``eta_fixed``) ``transmission_kavg_adaptive_global_eta``.  ``kpoints`` is
    one k-point list for all omegas; ``eta_values`` one ladder or one ladder per omega.
    Every eta rung is one batched solve over all pending (omega, k) instances.

    Returns (values, records): values[i] is the averaged T of omega i (0 for omega <= 0,
    NaN where the omega failed) and records[i] is {"info": the per-omega statistics
    dict} or {"error": the reference's exception message} (None for omega <= 0).
    """
    w = np.asarray(list(omegas), dtype=float)
    scheme = str(eta_scheme).lower().replace("-", "_")
    if scheme not in ("adaptive", "adaptive_global", "fixed"):
        raise ValueError("eta_scheme must be 'adaptive', 'adaptive_global' or 'fixed'.")
    if scheme == "fixed":
        ladders = [(float(eta_fixed),)] * w.size
    elif len(eta_values) > 0 and np.ndim(eta_values[0]) > 0:
        if len(eta_values) != w.size:
            raise ValueError("per-omega eta_values need one ladder per omega.")
        ladders = [tuple(float(e) for e in lad) for lad in eta_values]
    else:
        ladders = [tuple(float(e) for e in eta_values)] * w.size
    kpoints = list(kpoints)
    for lad in set(ladders):
        _validate_adaptive(kpoints, lad, min_success_fraction, nonnegative_tolerance,
                           max_channel_factor)
    canonical_method(surface_gf_method)
    prec = _prec(precision)
    t_max = _t_max(lead_left, kpoints, max_channel_factor)
    solver = _KSolver(device, lead_left, lead_right, kpoints, eta_device=eta_device,
                      route=surface_route(surface_gf_method), precision=prec, device_to_lead_left=device_to_lead_left,
                      device_to_lead_right=device_to_lead_right,
                      contact_left_indices=contact_left_indices,
                      contact_right_indices=contact_right_indices, omega_scale=omega_scale)
    pos = np.nonzero(w > 0.0)[0]
    res = _adaptive_batch(w[pos], solver, t_max, [ladders[i] for i in pos],
                          eta_device=eta_device, global_eta=(scheme != "adaptive"),
                          min_success_fraction=min_success_fraction,
                          nonnegative_tolerance=nonnegative_tolerance,
                          collect_rejected=collect_rejected)
    values = np.zeros(w.size)
    records: list = [None] * w.size
    for j, i in enumerate(pos):
        r = res[j]
        if r[0] == "ok":
            values[i] = r[1]
            records[i] = {"info": r[2]}
        else:
            values[i] = np.nan
            records[i] = {"error": r[1]}
    return values, records


def transmission_k_moments_spectrum(
        omegas, device: DeviceLike, lead_left: LeadLike, lead_right: LeadLike, kpoints,
        eta_values=(1e-8, 1e-7, 1e-6, 1e-5), *, eta_scheme: str = "adaptive_global",
        eta_fixed: float = 1e-4, eta_device: float | None = None,
        min_success_fraction: float = 0.0, device_to_lead_left: DeviceToLeadLike | None = None,
        device_to_lead_right: DeviceToLeadLike | None = None,
        surface_gf_method: str = "sancho_rubio", omega_scale: float | None = None,
        nonnegative_tolerance: float = 0.0, precision: str | None = None) -> dict:
    """k-moments <T>(omega) and <T^2>(omega) over the accepted k points of one eta per
    omega, for a whole spectrum in batched solves (one per eta rung).

    The batched form of the reference's FCS moment pass (workflows/ifc_bulk.py:1690-1785):
    for each omega > 0 the first eta (of ``eta_values``, or ``eta_fixed`` for
    ``eta_scheme="fixed"``) whose accepted k points (finite, T >= -tolerance, clamped
    at 0) reach ``min_success_fraction`` gives t_mean = mean(T), t2_mean = mean(T^2).
    Returns the reference's keys: t_mean, t2_mean (0 for omega <= 0, NaN where no eta
    passed), eta_selected, n_k_success, omega_failures.
    """
    w = np.asarray(list(omegas), dtype=float)
    scheme = str(eta_scheme).lower().replace("-", "_")
    if scheme not in ("adaptive_global", "fixed"):
        scheme = "adaptive_global"
    ladder = (float(eta_fixed),) if scheme == "fixed" else tuple(float(e) for e in eta_values)
    kpoints = list(kpoints)
    _validate_adaptive(kpoints, ladder, min_success_fraction, nonnegative_tolerance, None)
    canonical_method(surface_gf_method)
    prec = _prec(precision)
    solver = _KSolver(device, lead_left, lead_right, kpoints, eta_device=eta_device,
                      route=surface_route(surface_gf_method), precision=prec, device_to_lead_left=device_to_lead_left,
                      device_to_lead_right=device_to_lead_right, contact_left_indices=None,
                      contact_right_indices=None, omega_scale=omega_scale)
    n = w.size
    t_mean = np.full(n, np.nan)
    t2_mean = np.full(n, np.nan)
    eta_sel = np.full(n, np.nan)
    n_ok = np.zeros(n, dtype=int)
    failures = []
    pos = np.nonzero(w > 0.0)[0]
    t_mean[w <= 0.0] = 0.0
    t2_mean[w <= 0.0] = 0.0
    res = _adaptive_batch(w[pos], solver, [None] * len(kpoints), [ladder] * pos.size,
                          eta_device=eta_device, global_eta=True,
                          min_success_fraction=min_success_fraction,
                          nonnegative_tolerance=nonnegative_tolerance, collect_rejected=False)
    for j, i in enumerate(pos):
        r = res[j]
        if r[0] != "ok":
            failures.append({"omega": float(w[i]), "reason": "no_eta_passed_threshold"})
            continue
        acc = r[3]
        t_mean[i] = float(np.mean(acc))
        t2_mean[i] = float(np.mean(acc * acc))
        eta_sel[i] = float(r[2]["eta_selected"])
        n_ok[i] = int(acc.size)
    return {"t_mean": t_mean, "t2_mean": t2_mean, "eta_selected": eta_sel,
            "n_k_success": n_ok, "omega_failures": failures}
