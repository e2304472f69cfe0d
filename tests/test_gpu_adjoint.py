"""GPU adjoint: dT/d(device blocks) and d kappa/d(force constants) through autograd.

Checked against (a) torch autograd through a dense complex128 solve on the CPU
(``adjoint_util``, itself pinned to the oracle by ``test_adjoint_reference.py``)
and (b) central finite differences of the GPU forward path, to 1e-6.
"""

import numpy as np
import pytest
import torch

from adjoint_util import dense_grads
from negfpy_amd.core import (
    LeadBlocks,
    thermal_conductance_torch,
    transmission,
    transmission_torch,
)
from negfpy_amd.core.types import Device1D
from negfpy_amd.models import multidof_lead_blocks
from negfpy_amd.workloads import OMEGA_UNIT_RAD_S, omega_max, random_lead, ribbon_params
from oracle import negf_oracle as o
from parity_util import oracle_spread
from test_reference_vectors import REF

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-9


def _md():
    return REF["md_onsite"], REF["md_coupling"], LeadBlocks(d00=REF["md_d00"], d01=REF["md_d01"])


def _w_edge():
    return float(REF["md_omegas"][-1] / 1.15)


# (product kwargs, oracle kwargs)
CASES = {
    "plain": (dict(eta=1e-6), dict(eta=1e-6)),
    "eta8": (dict(eta=1e-8, eta_device=1e-9), dict(eta=1e-8, eta_device=1e-9)),
    "scaled": (dict(eta=1e-6, omega_scale=1.7), dict(eta=1e-6, omega_scale=1.7)),
    "contacts": (dict(eta=1e-6, device_to_lead_left=REF["md_vl"],
                      device_to_lead_right=REF["md_vr"], contact_left_indices=[0, 2, 4],
                      contact_right_indices=[25, 27, 29]),
                 dict(eta=1e-6, v_left=REF["md_vl"], v_right=REF["md_vr"], idx_left=[0, 2, 4],
                      idx_right=[25, 27, 29])),
    "merged": (dict(eta=1e-6, device_to_lead_left=REF["md_vl"],
                    device_to_lead_right=REF["md_vr"], contact_left_indices=[0, 2, 7],
                    contact_right_indices=[20, 22, 28]),
               dict(eta=1e-6, v_left=REF["md_vl"], v_right=REF["md_vr"], idx_left=[0, 2, 7],
                    idx_right=[20, 22, 28])),
}


def _gpu_grads(omegas, ons, cps, lead_l, lead_r, coef, **kw):
    ons_t = torch.tensor(ons, dtype=torch.complex128, device="cuda", requires_grad=True)
    cps_t = torch.tensor(cps, dtype=torch.complex128, device="cuda", requires_grad=True)
    t = transmission_torch(omegas, ons_t, cps_t, lead_l, lead_r, **kw)
    (t * torch.as_tensor(coef, device="cuda")).sum().backward()
    g_c = cps_t.grad.cpu().numpy() if cps_t.grad is not None else np.zeros(np.shape(cps))
    return t.detach().cpu().numpy(), ons_t.grad.cpu().numpy(), g_c


def richardson(f, h):
    """Central difference f'(0) with one Richardson step: error O(h^4) + rounding / h.
    The GPU forward path rounds differently from point to point (about 1e-13 relative,
    more near band features), so the step is kept well above that noise."""
    d1 = (f(h) - f(-h)) / (2 * h)
    d2 = (f(h / 2) - f(-h / 2)) / h
    return (4 * d2 - d1) / 3


def _assert_grad(g, g_ref, what):
    scale = max(np.max(np.abs(g_ref)), 1e-12)
    err = np.max(np.abs(g - g_ref)) / scale
    assert err <= GRAD_TOL, (what, err)


@pytest.mark.parametrize("case", sorted(CASES))
def test_adjoint_matches_dense_autograd(gpu, case):
    ons, cps, lead = _md()
    kw, okw = CASES[case]
    omegas = np.array([0.15, 0.45, 0.7, 0.95]) * _w_edge()
    coef = np.array([1.0, -0.5, 0.25, 2.0])
    t, g_ons, g_cps = _gpu_grads(omegas, ons, cps, lead, lead, coef, **kw)
    tup = (REF["md_d00"], REF["md_d01"], None)
    ref_ons = np.zeros_like(g_ons)
    ref_cps = np.zeros_like(g_cps)
    for w, c, tg in zip(omegas, coef, t):
        tr, go, gc = dense_grads(w, ons, cps, tup, tup, **okw)
        assert abs(tg - tr) <= 1e-10 * max(abs(tr), 1e-3)
        ref_ons += c * go
        ref_cps += c * gc
    _assert_grad(g_ons, ref_ons, "onsite")
    _assert_grad(g_cps, ref_cps, "coupling")


def test_adjoint_two_leads(gpu):
    ons, cps, lead = _md()
    lead_r = LeadBlocks(d00=REF["md_d00_r"], d01=REF["md_d01_r"])
    omegas = np.array([0.2, 0.6]) * _w_edge()
    t, g_ons, g_cps = _gpu_grads(omegas, ons, cps, lead, lead_r, np.ones(2), eta=1e-6)
    tl = (REF["md_d00"], REF["md_d01"], None)
    tr_ = (REF["md_d00_r"], REF["md_d01_r"], None)
    refs = [dense_grads(w, ons, cps, tl, tr_, eta=1e-6) for w in omegas]
    _assert_grad(g_ons, sum(r[1] for r in refs), "onsite")
    _assert_grad(g_cps, sum(r[2] for r in refs), "coupling")


def test_adjoint_single_layer(gpu):
    ons, _, lead = _md()
    w = 0.4 * _w_edge()
    t, g_ons, _ = _gpu_grads([w], ons[:1], np.zeros((0, 6, 6)), lead, lead, [1.0], eta=1e-6)
    tup = (REF["md_d00"], REF["md_d01"], None)
    tr, go, _ = dense_grads(w, ons[:1], np.zeros((0, 6, 6)), tup, tup, eta=1e-6)
    assert abs(t[0] - tr) <= 1e-10 * max(abs(tr), 1e-3)
    _assert_grad(g_ons, go, "onsite")


@pytest.mark.parametrize("n,nl", [(24, 6), (48, 3), (60, 3), (90, 2), (96, 2)])
def test_adjoint_large_blocks(gpu, n, nl):
    d00, d01 = random_lead(40 + n, n)
    rng = np.random.default_rng(n)
    ons = np.stack([d00 + 0.05 * np.diag(rng.uniform(-1, 1, n)) for _ in range(nl)])
    cps = np.stack([d01 * (1.0 + 0.03 * rng.uniform(-1, 1)) for _ in range(nl - 1)])
    lead = LeadBlocks(d00=d00, d01=d01)
    w = 0.37 * omega_max(lead)
    t, g_ons, g_cps = _gpu_grads([w], ons, cps, lead, lead, [1.0], eta=1e-6)
    tup = (d00, d01, None)
    tr, go, gc = dense_grads(w, ons, cps, tup, tup, eta=1e-6)
    assert abs(t[0] - tr) <= 1e-9 * max(abs(tr), 1e-3)
    _assert_grad(g_ons, go, "onsite")
    _assert_grad(g_cps, gc, "coupling")


def test_adjoint_chunked_backward(gpu, monkeypatch):
    """Backward in chunks of one instance gives the same gradient as one launch."""
    from negfpy_amd.core import adjoint as adj

    ons, cps, lead = _md()
    omegas = np.array([0.15, 0.45, 0.7]) * _w_edge()
    coef = np.array([1.0, -0.5, 0.25])
    _, g1, c1 = _gpu_grads(omegas, ons, cps, lead, lead, coef, eta=1e-6)
    monkeypatch.setattr(adj, "ADJOINT_CHUNK_BYTES", 1)
    _, g2, c2 = _gpu_grads(omegas, ons, cps, lead, lead, coef, eta=1e-6)
    _assert_grad(g2, g1, "onsite")
    _assert_grad(c2, c1, "coupling")


def test_adjoint_matches_finite_differences(gpu):
    """Directional derivatives of the GPU forward path along random complex directions."""
    ons, cps, lead = _md()
    omegas = np.array([0.3, 0.8]) * _w_edge()
    coef = np.array([1.0, 1.0])
    _, g_ons, g_cps = _gpu_grads(omegas, ons, cps, lead, lead, coef, eta=1e-6)
    rng = np.random.default_rng(17)

    def total(o_, c_):
        dev = Device1D(onsite_blocks=list(o_), coupling_blocks=list(c_))
        return sum(transmission(w, dev, lead, lead, eta=1e-6) for w in omegas)

    # T(0.3 w_edge) has a large third derivative along random directions (the plain
    # central difference error is 3.4e7 h^2 relative there): h = 1e-5 with one
    # Richardson step leaves ~1e-8.
    h = 1e-5 * np.max(np.abs(ons))
    for _ in range(3):
        d_o = rng.standard_normal(ons.shape) + 1j * rng.standard_normal(ons.shape)
        d_c = rng.standard_normal(cps.shape) + 1j * rng.standard_normal(cps.shape)
        fd = richardson(lambda e: total(ons + e * d_o, cps + e * d_c), h)
        an = np.sum(np.real(np.conj(g_ons) * d_o)) + np.sum(np.real(np.conj(g_cps) * d_c))
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-3), (fd, an)


def _ifc_device(p_fc, p_mass, params, n_layers):
    """Device blocks from force constants and masses as a differentiable torch
    expression: D = M^-1/2 Phi M^-1/2 with Phi scaled by p_fc and layer masses by p_mass."""
    dof = params.dof_per_atom
    fc00 = torch.as_tensor(params.fc00, dtype=torch.float64, device="cuda") * p_fc
    fc01 = torch.as_tensor(params.fc01, dtype=torch.float64, device="cuda") * p_fc
    m = torch.as_tensor(np.repeat(params.masses, dof), dtype=torch.float64, device="cuda")
    ms = torch.stack([m * (1.0 + p_mass * (i % 2)) for i in range(n_layers)])
    inv = ms.rsqrt()
    ons = torch.stack([inv[i][:, None] * fc00 * inv[i][None, :] for i in range(n_layers)])
    cps = torch.stack([inv[i][:, None] * fc01 * inv[i + 1][None, :] for i in range(n_layers - 1)])
    return ons, cps


def test_kappa_gradient_matches_finite_differences(gpu):
    """d kappa(300 K) / d(force-constant scale, mass contrast) through transmission_torch and
    thermal_conductance_torch, against central differences of the same forward path."""
    params = ribbon_params(3, 2)
    lead = multidof_lead_blocks(params)
    nl = 5
    wq = np.linspace(0.01, 1.05, 48) * omega_max(lead)
    w_rad = wq * OMEGA_UNIT_RAD_S

    def kappa(p_fc, p_mass):
        ons, cps = _ifc_device(p_fc, p_mass, params, nl)
        t = transmission_torch(wq, ons, cps, lead, lead, eta=1e-6)
        return thermal_conductance_torch(w_rad, t, [300.0])[0]

    p_fc = torch.tensor(1.0, dtype=torch.float64, device="cuda", requires_grad=True)
    p_mass = torch.tensor(0.2, dtype=torch.float64, device="cuda", requires_grad=True)
    k = kappa(p_fc, p_mass)
    k.backward()
    # forward: T(w) against the oracle with the spread-aware bar, kappa against the
    # oracle's conductance of the same spectrum
    ons, cps = _ifc_device(torch.tensor(1.0, dtype=torch.float64, device="cuda"),
                           torch.tensor(0.2, dtype=torch.float64, device="cuda"), params, nl)
    with torch.no_grad():
        t_gpu = transmission_torch(wq, ons, cps, lead, lead, eta=1e-6).cpu().numpy()
    ons_np, cps_np = list(ons.cpu().numpy()), list(cps.cpu().numpy())
    tup = (lead.d00, lead.d01, None)
    for w, tg in zip(wq, t_gpu):
        tr = o.transmission_sparse(w, ons_np, cps_np, tup, tup, eta=1e-6)
        sp = oracle_spread(w, ons_np, cps_np, tup, tup, tr, eta=1e-6)
        assert abs(tg - tr) <= max(1e-10 * max(abs(tr), 1e-3), 50.0 * sp), (w, tg, tr, sp)
    k_or = o.thermal_conductance(w_rad / o.qe_omega_to_rad_s(1.0), t_gpu, np.array([300.0]))[0]
    assert abs(float(k.detach()) - k_or) <= 1e-12 * abs(k_or)
    for p, name, h in ((p_fc, "fc", 1e-4), (p_mass, "mass", 1e-4)):
        j = 0 if name == "fc" else 1

        def k_at(e):
            v = [1.0, 0.2]
            v[j] += e
            with torch.no_grad():
                return float(kappa(*[torch.tensor(x, dtype=torch.float64, device="cuda")
                                     for x in v]))

        fd = richardson(k_at, h)
        an = float(p.grad)
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-12 * abs(float(k.detach()))), (name, fd, an)


def test_conductance_weights_give_the_kernel_kappa(gpu):
    rng = np.random.default_rng(2)
    w = np.sort(rng.uniform(1e11, 9e13, 200))
    t = torch.tensor(rng.uniform(0, 3, 200), dtype=torch.float64, device="cuda",
                     requires_grad=True)
    temps = [0.0, 10.0, 300.0, 900.0]
    k = thermal_conductance_torch(w, t, temps)
    k_or = o.thermal_conductance(w / o.qe_omega_to_rad_s(1.0), t.detach().cpu().numpy(),
                                 np.array(temps))
    np.testing.assert_allclose(k.detach().cpu().numpy(), k_or, rtol=1e-12, atol=0)
    k.sum().backward()
    # kappa is linear in T: the gradient is sum_t of the weights, check via a dot product
    dt = rng.standard_normal(200)
    k2 = thermal_conductance_torch(w, t.detach() + torch.as_tensor(dt, device="cuda"), temps)
    lhs = float((k2 - k.detach()).sum())
    rhs = float((t.grad.cpu() * torch.as_tensor(dt)).sum())
    assert abs(lhs - rhs) <= 1e-10 * abs(float(k.detach().sum()))


def test_transmission_torch_argument_errors(gpu):
    ons, cps, lead = _md()
    with pytest.raises(ValueError, match="onsite must have shape"):
        transmission_torch([1.0], ons[0], cps, lead, lead)
    with pytest.raises(ValueError, match="coupling must have shape"):
        transmission_torch([1.0], ons, cps[:2], lead, lead)
    with pytest.raises(ValueError, match="thermal|tvals"):
        thermal_conductance_torch([1.0, 2.0], torch.zeros(3, dtype=torch.float64), [300.0])


@pytest.mark.parametrize("n,nl", [(6, 11), (24, 9), (60, 5)])
def test_checkpointed_adjoint_equals_stored(gpu, n, nl):
    """Layer checkpoints (the left-connected sweep kept every K layers and recomputed per
    segment) give the stored form's T and gradients bit for bit, for K = 1, 2, 3, sqrt(L)
    and L (one segment), in the workgroup classes and the pooled 96 class; per-set sums on
    the GPU (atomic adds) equal the per-instance gradients summed on the host."""
    from negfpy_amd.core import _engine

    d00, d01 = random_lead(70 + n, n)
    rng = np.random.default_rng(n + nl)
    sets_n = 2
    ons = np.stack([[d00 + 0.05 * np.diag(rng.uniform(-1, 1, n)) for _ in range(nl)]
                    for _ in range(sets_n)])
    cps = np.stack([[d01 * (1.0 + 0.03 * rng.uniform(-1, 1)) for _ in range(nl - 1)]
                    for _ in range(sets_n)])
    lb = np.stack([d00, d01, d01.conj().T])[None].repeat(sets_n, axis=0)
    sets = _engine.TransportSets.from_arrays(lb, lb, lb[:, 2], lb[:, 1], np.arange(n),
                                             np.arange(n), ons, cps, True)
    m = 12
    w = np.linspace(0.1, 1.0, m) * omega_max(LeadBlocks(d00=d00, d01=d01))
    set_of = np.arange(m, dtype=np.int32) % sets_n
    eta = np.full(m, 1e-6)
    weight = rng.uniform(0.5, 1.5, m)
    ref = _engine.adjoint(sets, set_of, w, eta, eta, weight)
    assert int((ref.status != 0).sum()) == 0
    for k in sorted({1, 2, 3, int(round(np.sqrt(nl))), nl}):
        res = _engine.adjoint(sets, set_of, w, eta, eta, weight, checkpoint=k)
        assert torch.equal(res.t, ref.t) and torch.equal(res.status, ref.status), k
        assert torch.equal(res.grad_onsite, ref.grad_onsite), k
        assert torch.equal(res.grad_coupling, ref.grad_coupling), k
    k_of = torch.as_tensor(set_of, device="cuda").long()
    per = _engine.adjoint(sets, set_of, w, eta, eta, weight, per_set=True, checkpoint=3)
    for got, full in ((per.grad_onsite, ref.grad_onsite), (per.grad_coupling, ref.grad_coupling)):
        want = torch.zeros_like(got)
        want.index_add_(0, k_of, full)
        err = (got - want).abs().max() / want.abs().max()
        assert float(err) < 1e-13, float(err)


def test_workspace_shrinks_with_checkpoints(gpu):
    from negfpy_amd import _lib

    lib = _lib.load()
    n, nl = 48, 64
    full = lib.negf_adjoint_workspace_bytes(n, nl, 1, 0)
    assert full == (2 * nl - 1) * n * n * 16
    k8 = lib.negf_adjoint_workspace_bytes(n, nl, 1, 8)
    assert k8 == (nl + 8 + 8 + 1) * n * n * 16 and k8 < 0.65 * full
    assert lib.negf_adjoint_workspace_bytes(n, nl, 1, 1000) == \
        lib.negf_adjoint_workspace_bytes(n, nl, 1, nl)
    assert lib.negf_adjoint_workspace_bytes(n, nl, 1, -1) == 0
