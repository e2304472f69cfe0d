"""Differentiable transmission and thermal conductance.

The reference computes T(omega) and kappa(T) without gradients
(``src/negfpy/core/negf.py:306-357``, ``src/negfpy/workflows/ifc_bulk.py:408-438``).
Here both are ``torch.autograd.Function``s:

* ``transmission_torch`` -- T(omega) for a device given as torch tensors of onsite
  and coupling blocks.  Forward: the batched FP64 transmission kernels.  Backward:
  the adjoint kernel ``k_adjoint`` (one right- and one left-connected sweep per
  (omega) instance), weighted by the incoming dL/dT and summed over omega on the
  GPU.  Leads and device-to-lead couplings are constants here; the k-resolved
  functions (``transmission_sets_torch``, ``transmission_sum_torch``) also
  differentiate them (the lead adjoint kernel).
* ``thermal_conductance_torch`` -- kappa(T) of a GPU spectrum (the k_conductance
  kernel); backward multiplies by the trapezoid/Bose weights of
  ``k_conductance_weights``.

Chaining them gives d kappa / d(force constants) through any torch expression
that builds the blocks from force constants and masses.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _engine
from ._status import transmission_value
from .negf import _max_block_green, resolve_k
from .types import Device1D

__all__ = ["transmission_torch", "transmission_sets_torch", "transmission_sum_torch",
           "thermal_conductance_torch"]

# Device memory one backward launch may use for adjoint workspace and per-instance
# gradients; larger omega batches are processed in chunks of this size.  The
# k-resolved functions may use this fraction of the free device memory instead
# when that is more.
ADJOINT_CHUNK_BYTES = 2 << 30
ADJOINT_FREE_FRACTION = 0.25
# Layer checkpointing of the k-resolved backward: None recomputes the left-connected
# sweep from checkpoints every ~sqrt(L) layers only when the stored blocks would leave
# fewer than ADJOINT_MIN_CHUNK instances per launch (too few to fill the GPU); 0 never;
# K > 0 always, every K layers.
ADJOINT_CHECKPOINT = None
ADJOINT_MIN_CHUNK = 2048


def _merge_layers_torch(onsite: torch.Tensor, coupling: torch.Tensor, m: int):
    """Group m consecutive layers into one block (differentiable), zero-padding the tail;
    the same layout as the product's block merging for contacts deep in the device."""
    nl, n = int(onsite.shape[0]), int(onsite.shape[1])
    nb = -(-nl // m)
    big = m * n
    ons = onsite.new_zeros((nb, big, big))
    cpl = onsite.new_zeros((max(nb - 1, 0), big, big))
    for i in range(nl):
        b, r = divmod(i, m)
        ons[b, r * n:(r + 1) * n, r * n:(r + 1) * n] = onsite[i]
    for i in range(nl - 1):
        b, r = divmod(i, m)
        blk = coupling[i]
        if r < m - 1:
            ons[b, r * n:(r + 1) * n, (r + 1) * n:(r + 2) * n] = blk
            ons[b, (r + 1) * n:(r + 2) * n, r * n:(r + 1) * n] = blk.conj().transpose(0, 1)
        else:
            cpl[b, (m - 1) * n:m * n, 0:n] = blk
    return ons, cpl


class _Static:
    """Everything of one transmission problem except the device blocks."""

    def __init__(self, kp, omega, eta, eta_dev):
        s = kp.scale
        self.kp = kp
        self.omega = _engine.to_dev_f64(np.asarray(omega, dtype=float) / s)
        self.eta = _engine.to_dev_f64(np.full(self.omega.numel(), float(eta) / s))
        self.eta_dev = _engine.to_dev_f64(np.full(self.omega.numel(), float(eta_dev) / s))
        self.set_of = torch.zeros(self.omega.numel(), dtype=torch.int32,
                                  device=self.omega.device)

    def sets(self, ons: torch.Tensor, cpl: torch.Tensor) -> _engine.TransportSets:
        kp = self.kp
        share = bool(np.array_equal(kp.lead_l, kp.lead_r))
        return _engine.TransportSets.from_arrays(
            kp.lead_l[None], kp.lead_r[None], kp.v_l[None], kp.v_r[None], kp.pos_l, kp.pos_r,
            ons.detach()[None], cpl.detach()[None] if ons.shape[0] > 1 else None, share)


class _TransmissionFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ons, cpl, static):
        sets = static.sets(ons, cpl)
        solver = _engine.TransmissionSolver(sets, precision="fp64")
        res = solver.run(static.set_of, static.omega, static.eta, static.eta_dev)
        st = res.status.cpu().numpy()
        t = res.t
        if np.any(st != 0):
            tc = t.cpu().numpy()
            vals = [transmission_value(int(c), float(v)) for c, v in zip(st, tc)]
            t = torch.as_tensor(np.asarray(vals), dtype=torch.float64, device=t.device)
        ctx.static = static
        ctx.save_for_backward(ons, cpl)
        return t

    @staticmethod
    def backward(ctx, grad_t):
        ons, cpl = ctx.saved_tensors
        static = ctx.static
        sets = static.sets(ons, cpl)
        weight = grad_t.detach().to(torch.float64).contiguous()
        nl, n = int(ons.shape[0]), int(ons.shape[1])
        n_inst = int(static.omega.numel())
        per_inst = _adjoint_bytes_per_instance(sets, nl, n)
        chunk = max(1, min(n_inst, ADJOINT_CHUNK_BYTES // per_inst))
        g_ons = torch.zeros_like(ons)
        g_cpl = torch.zeros((max(nl - 1, 1), n, n), dtype=ons.dtype, device=ons.device)
        for a in range(0, n_inst, chunk):
            b = min(n_inst, a + chunk)
            res = _engine.adjoint(sets, static.set_of[a:b], static.omega[a:b], static.eta[a:b],
                                  static.eta_dev[a:b], weight=weight[a:b])
            g_ons += res.grad_onsite.sum(dim=0)
            g_cpl += res.grad_coupling.sum(dim=0)
        return g_ons, (g_cpl[:nl - 1] if nl > 1 else None), None


def transmission_torch(omega, onsite, coupling, lead_left, lead_right, *, eta: float = 1e-8,
                       eta_device: float | None = None, device_to_lead_left=None,
                       device_to_lead_right=None, contact_left_indices=None,
                       contact_right_indices=None, omega_scale: float | None = None
                       ) -> torch.Tensor:
    """T(omega) for every omega of ``omega`` as a differentiable GPU tensor.

    ``onsite`` is (L, n, n) and ``coupling`` (L-1, n, n), real or complex torch tensors
    (or arrays); gradients flow to both.  Arguments and conventions otherwise follow
    ``transmission`` (``negf.py:306-357``), with the leads and device-to-lead couplings
    held fixed.
    """
    c128 = torch.complex128
    dev = _engine.device()
    ons = onsite if isinstance(onsite, torch.Tensor) else torch.as_tensor(np.asarray(onsite))
    cpl = coupling if isinstance(coupling, torch.Tensor) else torch.as_tensor(np.asarray(coupling))
    ons = ons.to(device=dev, dtype=c128)
    cpl = cpl.to(device=dev, dtype=c128)
    if ons.ndim != 3 or ons.shape[1] != ons.shape[2]:
        raise ValueError("onsite must have shape (n_layers, n, n).")
    nl, n = int(ons.shape[0]), int(ons.shape[1])
    if cpl.ndim != 3 or int(cpl.shape[0]) != nl - 1 or (nl > 1 and tuple(cpl.shape[1:]) != (n, n)):
        raise ValueError("coupling must have shape (n_layers - 1, n, n).")
    w = np.atleast_1d(np.asarray(omega, dtype=float))
    host = Device1D(onsite_blocks=list(ons.detach().cpu().numpy()),
                    coupling_blocks=list(cpl.detach().cpu().numpy()))
    kp = resolve_k(host, lead_left, lead_right, None, eta=eta, eta_device=eta_device,
                   device_to_lead_left=device_to_lead_left,
                   device_to_lead_right=device_to_lead_right,
                   contact_left_indices=contact_left_indices,
                   contact_right_indices=contact_right_indices, omega_scale=omega_scale,
                   max_block=_max_block_green())
    if omega_scale is not None:
        s2 = float(omega_scale) ** 2
        ons = ons / s2
        cpl = cpl / s2
    m = int(kp.onsite.shape[1]) // n
    if m > 1:
        ons, cpl = _merge_layers_torch(ons, cpl, m)
    eta_dev = float(eta if eta_device is None else eta_device)
    return _TransmissionFn.apply(ons, cpl, _Static(kp, w, eta, eta_dev))


def _adjoint_blocks(nl: int, ckpt: int) -> int:
    """Workspace blocks of one adjoint instance (negf_adjoint_workspace_bytes)."""
    if ckpt <= 0:
        return 2 * nl - 1
    k = min(ckpt, nl)
    return nl + -(-nl // k) + k + 1


def _adjoint_bytes_per_instance(sets, nl: int, n: int, leads: bool = False, ckpt: int = 0,
                                per_set: bool = False) -> int:
    """Device bytes one instance of an adjoint chunk holds: the kernel's workspace
    blocks, the per-instance gradient blocks (none when summed per set on the GPU), and
    the lead surface GFs and self-energies of _engine.adjoint (complex128); with lead
    gradients also G_1N, the end-layer gradients and the lead and coupling gradients."""
    n_l = int(sets.lead_l.shape[2])
    n_r = int(sets.lead_r.shape[2])
    b = (_adjoint_blocks(nl, ckpt) + (0 if per_set else 2 * nl - 1) + 2) * n * n \
        + n_l * n_l + n_r * n_r
    if leads:
        b += 3 * n * n + 3 * (n_l * n_l + n_r * n_r) + int(sets.v_l.shape[1]) * n_l \
            + int(sets.v_r.shape[1]) * n_r
    return b * 16


def _plan(sets, nl: int, n: int, leads: bool, n_inst: int, dev):
    """(checkpoint interval, instances per launch) of a per-set backward."""
    stored = _chunk_instances(_adjoint_bytes_per_instance(sets, nl, n, leads, 0, True), n_inst,
                              dev)
    ck = ADJOINT_CHECKPOINT
    if ck == 0 or (ck is None and stored >= min(n_inst, ADJOINT_MIN_CHUNK)):
        return 0, stored
    k = int(ck) if ck else max(1, int(round(np.sqrt(nl))))
    return k, _chunk_instances(_adjoint_bytes_per_instance(sets, nl, n, leads, k, True), n_inst,
                               dev)


def _chunk_instances(per_inst: int, n_inst: int, dev) -> int:
    """Instances per backward launch: the per-instance adjoint memory within
    ADJOINT_CHUNK_BYTES, or within ADJOINT_FREE_FRACTION of the free device memory if
    that is more."""
    budget = ADJOINT_CHUNK_BYTES
    if dev.type == "cuda":
        budget = max(budget, int(torch.cuda.mem_get_info(dev)[0] * ADJOINT_FREE_FRACTION))
    return max(1, min(n_inst, budget // per_inst))


def _leads_wanted(ctx) -> bool:
    return any(ctx.needs_input_grad[2:6])


def _add_lead_grads(acc, k, res):
    """Per-set sums of the per-instance lead gradients (real view: float64 atomics)."""
    for a, g in zip(acc, (res.grad_lead_l, res.grad_lead_r, res.grad_v_l, res.grad_v_r)):
        torch.view_as_real(a).index_add_(0, k, torch.view_as_real(g))


class _SetsTransmissionFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ons, cpl, lead_l, lead_r, v_l, v_r, static):
        sets = static.sets(ons, cpl, lead_l, lead_r, v_l, v_r)
        res = _engine.TransmissionSolver(sets, precision="fp64").run(
            static.set_of, static.omega, static.eta, static.eta_dev)
        if bool((res.status != 0).any()):
            st = res.status.cpu().numpy()
            tc = res.t.cpu().numpy()
            for c, v in zip(st, tc):
                transmission_value(int(c), float(v))  # raises the reference's exception
        ctx.static = static
        ctx.save_for_backward(ons, cpl, lead_l, lead_r, v_l, v_r)
        return res.t

    @staticmethod
    def backward(ctx, grad_t):
        ons, cpl, lead_l, lead_r, v_l, v_r = ctx.saved_tensors
        static = ctx.static
        sets = static.sets(ons, cpl, lead_l, lead_r, v_l, v_r)
        leads = _leads_wanted(ctx)
        weight = grad_t.detach().to(torch.float64).contiguous()
        nl, n = int(ons.shape[1]), int(ons.shape[2])
        n_inst = int(static.omega.numel())
        ckpt, chunk = _plan(sets, nl, n, leads, n_inst, ons.device)
        g_ons = torch.zeros_like(ons)
        g_cpl = torch.zeros((ons.shape[0], max(nl - 1, 1), n, n), dtype=ons.dtype,
                            device=ons.device)
        g_lead = [torch.zeros_like(t) for t in (lead_l, lead_r, v_l, v_r)] if leads else None
        for a in range(0, n_inst, chunk):
            b = min(n_inst, a + chunk)
            k = static.set_of[a:b].long()
            # the device-block gradients come back summed per set (GPU atomics)
            res = _engine.adjoint(sets, static.set_of[a:b], static.omega[a:b], static.eta[a:b],
                                  static.eta_dev[a:b], weight=weight[a:b], lead_grads=leads,
                                  checkpoint=ckpt, per_set=True)
            g_ons += res.grad_onsite
            g_cpl += res.grad_coupling
            if leads:
                _add_lead_grads(g_lead, k, res)
            del res
        gl = g_lead if leads else [None] * 4
        return (g_ons, (g_cpl[:, :nl - 1] if nl > 1 else None), *gl, None)


def _lead_input(t):
    """A lead block or coupling tensor on the GPU as complex128, keeping its autograd
    history (numpy arrays become constants)."""
    return torch.as_tensor(t).to(device=_engine.device(), dtype=torch.complex128)


class _SetsStatic:
    """Instances of a k-resolved transmission problem (the leads and couplings are
    autograd inputs of the functions)."""

    def __init__(self, set_of, omega, eta, eta_dev):
        self.set_of = _engine.to_dev_i32(set_of)
        self.omega = _engine.to_dev_f64(omega)
        m = int(self.omega.numel())
        self.eta = _engine.to_dev_f64(np.broadcast_to(np.asarray(eta, dtype=float), (m,)).copy())
        self.eta_dev = _engine.to_dev_f64(
            np.broadcast_to(np.asarray(eta_dev, dtype=float), (m,)).copy())

    @staticmethod
    def sets(ons, cpl, lead_l, lead_r, v_l, v_r) -> _engine.TransportSets:
        n = int(ons.shape[2])
        idx = np.arange(n)
        lead_l, lead_r = lead_l.detach().contiguous(), lead_r.detach().contiguous()
        return _engine.TransportSets.from_arrays(
            lead_l, lead_r, v_l.detach().contiguous(), v_r.detach().contiguous(), idx, idx,
            ons.detach(), cpl.detach() if ons.shape[1] > 1 else None,
            bool(torch.equal(lead_l, lead_r)))


def transmission_sets_torch(omega, set_of, onsite, coupling, lead_left, lead_right, v_left,
                            v_right, *, eta: float = 1e-8,
                            eta_device: float | None = None) -> torch.Tensor:
    """T of every instance i = (omega[i], set set_of[i]) as a differentiable GPU tensor,
    for device blocks given per set (e.g. per k point): ``onsite`` (S, L, n, n) and
    ``coupling`` (S, L-1, n, n) torch tensors, gradients flow to both.

    The leads ``lead_left`` / ``lead_right`` are (S, 3, n, n) blocks (d00, d01, d10)
    per set, each passed to the decimation as given (``surface_gf.py:18-81``), and
    ``v_left`` / ``v_right`` (S, n, n) the device-to-lead couplings of the first and
    last layer (all of its DOFs are contacts); Sigma = V g V^H as in
    ``negf.py:72-97``.  Where they are torch tensors that require grad, gradients flow
    to them too (the lead adjoint: the implicit derivative of the surface GF through
    its Stein equation, ``negf_lead_adjoint_batched``; blocks up to 48 DOF).  The batched
    form of ``transmission`` (``negf.py:306-357``) over the points of a k mesh, for
    gradients with respect to force constants that enter every k point's blocks, the
    leads' included.
    """
    dev = _engine.device()
    ons = onsite.to(device=dev, dtype=torch.complex128)
    cpl = coupling.to(device=dev, dtype=torch.complex128)
    if ons.ndim != 4 or ons.shape[2] != ons.shape[3]:
        raise ValueError("onsite must have shape (n_sets, n_layers, n, n).")
    s, nl, n = int(ons.shape[0]), int(ons.shape[1]), int(ons.shape[2])
    if cpl.ndim != 4 or tuple(cpl.shape) != (s, nl - 1, n, n):
        raise ValueError("coupling must have shape (n_sets, n_layers - 1, n, n).")
    eta_dev = float(eta if eta_device is None else eta_device)
    static = _SetsStatic(set_of, omega, eta, eta_dev)
    leads = _check_leads(s, n, lead_left, lead_right, v_left, v_right)
    _check_set_of(static, s)
    return _SetsTransmissionFn.apply(ons, cpl, *leads, static)


def _check_leads(s, n, lead_left, lead_right, v_left, v_right):
    out = [_lead_input(t) for t in (lead_left, lead_right, v_left, v_right)]
    for name, t in (("lead_left", out[0]), ("lead_right", out[1])):
        if tuple(t.shape) != (s, 3, n, n):
            raise ValueError(f"{name} must have shape (n_sets, 3, n, n).")
    for name, t in (("v_left", out[2]), ("v_right", out[3])):
        if tuple(t.shape) != (s, n, n):
            raise ValueError(f"{name} must have shape (n_sets, n, n).")
    if any(t.requires_grad for t in out) and n > 48:
        raise NotImplementedError("gradients with respect to the leads take blocks of up to "
                                  f"48 DOF (got {n}); pass the leads as constants")
    return out


def _check_set_of(static, s):
    k = static.set_of
    if k.numel() != static.omega.numel():
        raise ValueError("set_of needs one entry per omega.")
    if k.numel() and (int(k.min()) < 0 or int(k.max()) >= s):
        raise ValueError("set_of entries must index the sets.")


class _SetsSumFn(torch.autograd.Function):
    """sum_i c_i T_i with its gradient from one adjoint pass (the weights c_i are known
    before T is): the forward runs the adjoint kernel and keeps the per-set gradient."""

    @staticmethod
    def forward(ctx, ons, cpl, lead_l, lead_r, v_l, v_r, static, coef):
        sets = static.sets(ons, cpl, lead_l, lead_r, v_l, v_r)
        leads = _leads_wanted(ctx)
        nl, n = int(ons.shape[1]), int(ons.shape[2])
        n_inst = int(static.omega.numel())
        ckpt, chunk = _plan(sets, nl, n, leads, n_inst, ons.device)
        g_ons = torch.zeros_like(ons)
        g_cpl = torch.zeros((ons.shape[0], max(nl - 1, 1), n, n), dtype=ons.dtype,
                            device=ons.device)
        g_lead = [torch.zeros_like(t) for t in (lead_l, lead_r, v_l, v_r)] if leads else []
        t_all = torch.empty((n_inst,), dtype=torch.float64, device=ons.device)
        st_all = torch.empty((n_inst,), dtype=torch.int32, device=ons.device)
        for a in range(0, n_inst, chunk):
            b = min(n_inst, a + chunk)
            k = static.set_of[a:b].long()
            res = _engine.adjoint(sets, static.set_of[a:b], static.omega[a:b], static.eta[a:b],
                                  static.eta_dev[a:b], weight=coef[a:b], lead_grads=leads,
                                  checkpoint=ckpt, per_set=True)
            t_all[a:b] = res.t
            st_all[a:b] = res.status
            g_ons += res.grad_onsite
            g_cpl += res.grad_coupling
            if leads:
                _add_lead_grads(g_lead, k, res)
            del res
        if bool((st_all != 0).any()):
            for c, v in zip(st_all.cpu().numpy(), t_all.cpu().numpy()):
                transmission_value(int(c), float(v))  # raises the reference's exception
        ctx.save_for_backward(g_ons, g_cpl, *g_lead)
        ctx.nl = nl
        return (coef * t_all).sum()

    @staticmethod
    def backward(ctx, grad):
        g_ons, g_cpl, *g_lead = ctx.saved_tensors
        nl = ctx.nl
        g = grad.to(g_ons.dtype)
        gl = [g * t for t in g_lead] if g_lead else [None] * 4
        return (g * g_ons, (g * g_cpl[:, :nl - 1] if nl > 1 else None), *gl, None, None)


def transmission_sum_torch(coef, omega, set_of, onsite, coupling, lead_left, lead_right, v_left,
                           v_right, *, eta: float = 1e-8,
                           eta_device: float | None = None) -> torch.Tensor:
    """sum_i coef[i] T_i over the instances of ``transmission_sets_torch`` (same
    arguments), differentiable in ``onsite`` and ``coupling`` (and in the leads and
    couplings where they require grad) but not in ``coef``.

    Because coef is known before T, one adjoint pass gives the sum and its gradient
    (the adjoint kernel weighted by coef); there is no separate forward solve.  For
    kappa(T) of a k-averaged spectrum, coef[i] = (conductance weight of omega_i) / n_k
    (``_engine.conductance_weights``).
    """
    dev = _engine.device()
    ons = onsite.to(device=dev, dtype=torch.complex128)
    cpl = coupling.to(device=dev, dtype=torch.complex128)
    if ons.ndim != 4 or ons.shape[2] != ons.shape[3]:
        raise ValueError("onsite must have shape (n_sets, n_layers, n, n).")
    s, nl, n = int(ons.shape[0]), int(ons.shape[1]), int(ons.shape[2])
    if cpl.ndim != 4 or tuple(cpl.shape) != (s, nl - 1, n, n):
        raise ValueError("coupling must have shape (n_sets, n_layers - 1, n, n).")
    eta_dev = float(eta if eta_device is None else eta_device)
    static = _SetsStatic(set_of, omega, eta, eta_dev)
    c = torch.as_tensor(coef, dtype=torch.float64, device=dev).detach().contiguous()
    if c.shape != static.omega.shape:
        raise ValueError("coef needs one entry per instance.")
    leads = _check_leads(s, n, lead_left, lead_right, v_left, v_right)
    _check_set_of(static, s)
    return _SetsSumFn.apply(ons, cpl, *leads, static, c)


class _ConductanceFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tvals, omega_rad_s, temps):
        ctx.save_for_backward(omega_rad_s, temps)
        return _engine.thermal_conductance_tensor(omega_rad_s, tvals, temps)

    @staticmethod
    def backward(ctx, grad_k):
        omega_rad_s, temps = ctx.saved_tensors
        wts = _engine.conductance_weights(omega_rad_s, temps)
        return grad_k.to(torch.float64) @ wts, None, None


def thermal_conductance_torch(omega_rad_s, tvals: torch.Tensor, temperatures) -> torch.Tensor:
    """kappa(T) = (1/2pi) int k_B T(w) x^2 e^-x / (1 - e^-x)^2 dw (w in rad/s), as a
    differentiable function of the spectrum ``tvals`` (``ifc_bulk.py:408-438``)."""
    dev = _engine.device()
    w = torch.as_tensor(np.asarray(omega_rad_s, dtype=float), device=dev)
    tk = torch.as_tensor(np.atleast_1d(np.asarray(temperatures, dtype=float)), device=dev)
    t = tvals.to(device=dev, dtype=torch.float64)
    if t.shape != w.shape:
        raise ValueError("tvals must have one entry per omega.")
    return _ConductanceFn.apply(t, w, tk)
