"""Gradients with respect to the leads (negf_lead_adjoint_batched through
transmission_sets_torch / transmission_sum_torch): dT/d(d00, d01, d10) of each lead and
dT/dV, against autograd through a dense CPU reference that differentiates the
decimation itself (complex128 torch: Sancho-Rubio iterations as in surface_gf.py:18-81,
Sigma = V g V^H as in negf.py:72-97, T from a dense solve as in negf.py:306-357), and
against Richardson central differences of the GPU forward."""

import numpy as np
import pytest
import torch

from negfpy_amd.core.adjoint import transmission_sets_torch, transmission_sum_torch
from negfpy_amd.core.types import LeadBlocks
from negfpy_amd.workloads import omega_max, random_lead

pytestmark = pytest.mark.gpu

C128 = torch.complex128


def _decimate(w, eta, d00, d01, d10, tol=1e-15, maxiter=400):
    """Differentiable Sancho-Rubio decimation (the oracle's iteration in torch)."""
    n = d00.shape[0]
    z = complex(w, eta) ** 2
    zi = z * torch.eye(n, dtype=C128)
    a, b, e, es = d01, d10, d00, d00
    for _ in range(maxiter):
        g = torch.linalg.inv(zi - e)
        ag, bg = a @ g, b @ g
        agb, bga = ag @ b, bg @ a
        es, e, a, b = es + agb, e + agb + bga, ag @ a, bg @ b
        if max(float(a.detach().abs().max()), float(b.detach().abs().max())) < tol:
            break
    else:
        raise RuntimeError("reference decimation did not converge")
    return torch.linalg.inv(zi - es)


def _dense_t(w, eta, ons, cpl, lead_l, lead_r, v_l, v_r):
    """T of one instance, all inputs torch (CPU complex128)."""
    nl, n = ons.shape[0], ons.shape[1]
    gl = _decimate(w, eta, lead_l[0], lead_l[1], lead_l[2])
    gr = _decimate(w, eta, lead_r[0], lead_r[1], lead_r[2])
    sl = v_l @ gl @ v_l.conj().T
    sr = v_r @ gr @ v_r.conj().T
    dim = nl * n
    blocks = [[torch.zeros((n, n), dtype=C128) for _ in range(nl)] for _ in range(nl)]
    for i in range(nl):
        blocks[i][i] = ons[i]
    for i in range(nl - 1):
        blocks[i][i + 1] = cpl[i]
        blocks[i + 1][i] = cpl[i].conj().T
    d = torch.cat([torch.cat(r, dim=1) for r in blocks], dim=0)
    a = complex(w, eta) ** 2 * torch.eye(dim, dtype=C128) - d
    emb_l = torch.zeros((dim, dim), dtype=C128)
    emb_l[:n, :n] = sl
    emb_r = torch.zeros((dim, dim), dtype=C128)
    emb_r[dim - n:, dim - n:] = sr
    a = a - emb_l - emb_r
    rhs = torch.zeros((dim, n), dtype=C128)
    rhs[dim - n:, :] = torch.eye(n, dtype=C128)
    glr = torch.linalg.solve(a, rhs)[:n, :]
    gam_l = 1j * (sl - sl.conj().T)
    gam_r = 1j * (sr - sr.conj().T)
    return torch.real(torch.trace(gam_l @ glr @ gam_r @ glr.conj().T))


def _problem(seed, n, n_layers, n_sets, n_w, eta):
    rng = np.random.default_rng(seed)
    leads, onss, cpls = [], [], []
    for side in range(2):
        per = []
        for k in range(n_sets):
            d00, d01 = random_lead(seed + 10 * side + k, n)
            per.append(np.stack([d00, d01, d01.conj().T]))
        leads.append(np.stack(per))
    for k in range(n_sets):
        d00, d01 = leads[0][k, 0], leads[0][k, 1]
        onss.append(np.stack([d00 + np.diag(rng.uniform(-0.1, 0.1, n)) for _ in range(n_layers)]))
        cpls.append(np.stack([d01 * (1 + 0.1 * rng.standard_normal()) for _ in range(n_layers - 1)])
                    if n_layers > 1 else np.zeros((0, n, n), np.complex128))
    v_l = np.stack([leads[0][k, 2] * (1 + 0.05 * rng.standard_normal()) for k in range(n_sets)])
    v_r = np.stack([leads[1][k, 1] * (1 + 0.05 * rng.standard_normal()) for k in range(n_sets)])
    wmax = omega_max(LeadBlocks(d00=leads[0][0, 0], d01=leads[0][0, 1]))
    omega = np.linspace(0.1, 1.05, n_w) * wmax
    set_of = np.arange(n_w) % n_sets
    weight = rng.uniform(0.5, 1.5, n_w)
    return dict(ons=np.stack(onss), cpl=np.stack(cpls), lead_l=leads[0], lead_r=leads[1],
                v_l=v_l, v_r=v_r, omega=omega, set_of=set_of, weight=weight, eta=eta)


NAMES = ("ons", "cpl", "lead_l", "lead_r", "v_l", "v_r")


def _leaves(p, device):
    return {k: torch.as_tensor(p[k], dtype=C128, device=device).clone().requires_grad_(True)
            for k in NAMES}


def _dense_grads(p):
    x = _leaves(p, "cpu")
    tot = 0.0
    for i, w in enumerate(p["omega"]):
        k = int(p["set_of"][i])
        tot = tot + p["weight"][i] * _dense_t(w, p["eta"], x["ons"][k], x["cpl"][k],
                                              x["lead_l"][k], x["lead_r"][k], x["v_l"][k],
                                              x["v_r"][k])
    tot.backward()
    return float(tot.detach()), {k: x[k].grad.numpy() for k in NAMES if x[k].grad is not None}


def _gpu_grads(p, one_pass=False):
    x = _leaves(p, "cuda")
    args = (p["omega"], p["set_of"], x["ons"], x["cpl"], x["lead_l"], x["lead_r"], x["v_l"],
            x["v_r"])
    if one_pass:
        tot = transmission_sum_torch(p["weight"], *args, eta=p["eta"])
    else:
        t = transmission_sets_torch(*args, eta=p["eta"])
        tot = (torch.as_tensor(p["weight"], device="cuda") * t).sum()
    tot.backward()
    return float(tot.detach()), {k: x[k].grad.cpu().numpy() for k in NAMES if x[k].grad is not None}


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("n,n_layers,one_pass", [(3, 3, False), (4, 1, False), (6, 4, True),
                                                 (12, 2, False)])
def test_lead_gradients_match_dense_autograd(gpu, n, n_layers, one_pass):
    p = _problem(40 + n, n, n_layers, 2, 10, 1e-3)
    t_ref, g_ref = _dense_grads(p)
    t_gpu, g_gpu = _gpu_grads(p, one_pass)
    assert abs(t_gpu - t_ref) <= 1e-10 * abs(t_ref)
    assert set(g_gpu) == set(g_ref) == set(NAMES) - ({"cpl"} if n_layers == 1 else set())
    for k in g_ref:
        err = _rel(g_gpu[k], g_ref[k])
        print(f"n={n} L={n_layers} {k}: rel {err:.2e} |g| {np.linalg.norm(g_ref[k]):.3e}",
              flush=True)
        assert err < 1e-9, (k, err)


def test_lead_gradients_match_finite_differences(gpu):
    """Richardson central differences of the GPU forward along random directions of every
    lead block and coupling, at eta = 1e-3 (T smooth on the steps taken)."""
    p = _problem(77, 6, 3, 1, 8, 1e-3)
    _, g = _gpu_grads(p)
    rng = np.random.default_rng(5)
    w = torch.as_tensor(p["weight"], device="cuda")

    def f(q):
        t = transmission_sets_torch(q["omega"], q["set_of"], *(torch.as_tensor(q[k], dtype=C128)
                                                                for k in NAMES), eta=q["eta"])
        return float((w * t).sum())

    for k in ("lead_l", "lead_r", "v_l", "v_r"):
        d = rng.standard_normal(p[k].shape) + 1j * rng.standard_normal(p[k].shape)
        d *= np.linalg.norm(p[k]) / np.linalg.norm(d)

        def fd(h):
            qp, qm = dict(p), dict(p)
            qp[k] = p[k] + h * d
            qm[k] = p[k] - h * d
            return (f(qp) - f(qm)) / (2 * h)

        h = 1e-4
        rich = (4 * fd(h / 2) - fd(h)) / 3
        ana = float(np.real(np.vdot(g[k], d)))
        print(f"{k}: adjoint {ana:.10e} fd {rich:.10e}", flush=True)
        assert abs(ana - rich) <= 1e-6 * max(abs(ana), 1e-12), (k, ana, rich)


def test_constant_leads_keep_the_device_path(gpu):
    """Leads without requires_grad: no lead adjoint, the device gradients unchanged."""
    p = _problem(91, 4, 3, 2, 6, 1e-3)
    x = _leaves(p, "cuda")
    t = transmission_sets_torch(p["omega"], p["set_of"], x["ons"], x["cpl"],
                                *(torch.as_tensor(p[k]) for k in ("lead_l", "lead_r", "v_l", "v_r")),
                                eta=p["eta"])
    (torch.as_tensor(p["weight"], device="cuda") * t).sum().backward()
    _, g_all = _gpu_grads(p)
    for k in ("ons", "cpl"):
        np.testing.assert_allclose(x[k].grad.cpu().numpy(), g_all[k], rtol=1e-12, atol=1e-14)


def test_kappa_gradient_through_leads_and_device(gpu):
    """d kappa(300 K) of a 3-D interface whose leads and device come from the same
    bulk force constants (``interface3d_lead_blocks``, ``interface3d_blocks``):
    along a global scale of every force constant and along a random symmetric change
    of the bulk same-layer constants that enters the leads only, against Richardson
    central differences of the GPU forward (eta = 1e-3, where kappa is smooth on the
    steps taken).  The lead part of the scale gradient is checked to be non-negligible."""
    from negfpy_amd.core import thermal_conductance_torch
    from negfpy_amd.workloads import (OMEGA_UNIT_RAD_S, interface3d_blocks,
                                      interface3d_lead_blocks, interface3d_problem)

    p = interface3d_problem(seed=3, cell=(2, 1, 1), n_layers=6, nk=2, n_omega=12,
                            n_perturbed=2)
    p.eta = 1e-3
    nk = len(p.kpoints)
    k_of = np.tile(np.arange(nk, dtype=np.int32), p.omegas.size)
    om = np.repeat(p.omegas, nk)
    dev = "cuda"
    fc00 = torch.tensor(p.fc00, device=dev)
    fc01 = torch.tensor(p.fc01, device=dev)
    b00 = torch.tensor(p.fc00[0], device=dev)
    b01 = torch.tensor(p.fc01[0], device=dev)

    def kappa(scale, lead_delta, leads_live=True):
        ons, cpl = interface3d_blocks(p, scale * fc00, scale * fc01)
        leads = interface3d_lead_blocks(p, scale * (b00 + lead_delta), scale * b01)
        if not leads_live:
            leads = [t.detach() for t in leads]
        t = transmission_sets_torch(om, k_of, ons, cpl, *leads, eta=p.eta)
        t = t.view(p.omegas.size, nk).mean(dim=1)
        return thermal_conductance_torch(p.omegas * OMEGA_UNIT_RAD_S, t, [300.0])[0]

    rng = np.random.default_rng(11)
    r = rng.standard_normal(p.fc00[0].shape)
    # keep Phi00(-d) = Phi00(d)^T: the same symmetric change for a term and its mirror
    terms = p.terms
    delta = np.zeros_like(p.fc00[0])
    for i, t in enumerate(terms):
        j = terms.index((-t[0], -t[1]))
        if j >= i:
            delta[i] = 0.05 * r[i] * p.fc00[0][i]
            delta[j] = delta[i].T if j != i else 0.5 * (delta[i] + delta[i].T)
    delta_t = torch.tensor(delta, device=dev)

    s = torch.tensor(1.0, dtype=torch.float64, device=dev, requires_grad=True)
    h = torch.tensor(0.0, dtype=torch.float64, device=dev, requires_grad=True)
    k0 = kappa(s, h * delta_t)
    k0.backward()
    ds, dh = float(s.grad), float(h.grad)
    s2 = torch.tensor(1.0, dtype=torch.float64, device=dev, requires_grad=True)
    kappa(s2, 0.0 * delta_t, leads_live=False).backward()
    ds_device_only = float(s2.grad)

    def fd(f, step):
        return (f(step) - f(-step)) / (2 * step)

    def richardson(f, step):
        return (4 * fd(f, step / 2) - fd(f, step)) / 3

    with torch.no_grad():
        fs = richardson(lambda e: float(kappa(1.0 + e, 0.0 * delta_t)), 1e-4)
        fh = richardson(lambda e: float(kappa(1.0, e * delta_t)), 1e-3)
    print(f"d kappa/d scale: adjoint {ds:.10e} fd {fs:.10e} (device only {ds_device_only:.4e}); "
          f"d kappa/d lead change: adjoint {dh:.10e} fd {fh:.10e}", flush=True)
    assert abs(ds - fs) <= 1e-6 * abs(fs)
    assert abs(dh - fh) <= 1e-6 * abs(fh)
    assert abs(ds - ds_device_only) > 1e-3 * abs(ds)
