"""k-resolved gradients: d kappa(300 K) / d(layer force constants) of a 3-D
mass-mismatch interface, through the device blocks of every k point
(``workloads.interface3d_blocks``, a torch expression), ``transmission_sets_torch``
(forward kernels, adjoint kernel, per-set sums) and ``thermal_conductance_torch``.

Checked against
- the oracle, per (omega, k) point, for the forward T;
- torch autograd through a dense complex128 solve, for dT/d(blocks) of single
  (omega, k) points, where the per-set sums of the backward are exercised;
- central finite differences of the whole GPU forward path (with one Richardson
  step), for d kappa along the gradient, scalings of the force constants and a
  random direction.
"""

import numpy as np
import pytest
import torch

from adjoint_util import dense_grads
from negfpy_amd.core import (_engine, thermal_conductance_torch, transmission_sets_torch,
                             transmission_sum_torch)
from negfpy_amd.workloads import OMEGA_UNIT_RAD_S, interface3d_blocks, interface3d_problem
from parity_util import assert_parity, oracle_t
from test_gpu_adjoint import richardson

pytestmark = pytest.mark.gpu


def _small():
    return interface3d_problem(seed=3, cell=(2, 1, 1), n_layers=6, nk=2, n_omega=12,
                               n_perturbed=2)


def _leaves(p):
    f00 = torch.tensor(p.fc00, dtype=torch.float64, device="cuda", requires_grad=True)
    f01 = torch.tensor(p.fc01, dtype=torch.float64, device="cuda", requires_grad=True)
    return f00, f01


def _instances(p):
    nk = len(p.kpoints)
    k_of = np.tile(np.arange(nk, dtype=np.int32), p.omegas.size)
    om = np.repeat(p.omegas, nk)
    return k_of, om


def _t(p, f00, f01):
    ons, cpl = interface3d_blocks(p, f00, f01)
    k_of, om = _instances(p)
    return transmission_sets_torch(om, k_of, ons, cpl, p.lead_left, p.lead_right, p.v_left,
                                   p.v_right, eta=p.eta)


def _kappa(p, f00, f01):
    t = _t(p, f00, f01).view(p.omegas.size, len(p.kpoints)).mean(dim=1)
    return thermal_conductance_torch(p.omegas * OMEGA_UNIT_RAD_S, t, [300.0])[0]


def _host_blocks(p, s):
    with torch.no_grad():
        ons, cpl = interface3d_blocks(p, torch.as_tensor(p.fc00, device="cuda"),
                                      torch.as_tensor(p.fc01, device="cuda"))
    return ons[s].cpu().numpy(), cpl[s].cpu().numpy()


def test_forward_matches_oracle(gpu):
    p = _small()
    f00, f01 = _leaves(p)
    with torch.no_grad():
        t = _t(p, f00, f01).cpu().numpy()
    k_of, om = _instances(p)
    sel = np.arange(0, om.size, 5)
    refs = []
    for i in sel:
        s = int(k_of[i])
        ons, cpl = _host_blocks(p, s)
        refs.append(oracle_t(om[i], list(ons), list(cpl), tuple(p.lead_left[s]),
                             tuple(p.lead_right[s]), eta=p.eta, v_left=p.v_left[s],
                             v_right=p.v_right[s]))
    assert_parity(t[sel], refs)


def test_block_gradients_match_dense_autograd(gpu):
    """dT_i/d(blocks of set s) for single instances: the backward's per-set sum leaves
    every other set at zero."""
    p = _small()
    k_of, om = _instances(p)
    ons0, cpl0 = interface3d_blocks(p, torch.as_tensor(p.fc00, device="cuda"),
                                    torch.as_tensor(p.fc01, device="cuda"))
    for i in (3, 14, 29):
        ons = ons0.detach().clone().requires_grad_(True)
        cpl = cpl0.detach().clone().requires_grad_(True)
        t = transmission_sets_torch(om, k_of, ons, cpl, p.lead_left, p.lead_right, p.v_left,
                                    p.v_right, eta=p.eta)
        t[i].backward()
        s = int(k_of[i])
        tr, go, gc = dense_grads(om[i], ons0[s].detach().cpu().numpy(),
                                 cpl0[s].detach().cpu().numpy(), tuple(p.lead_left[s]),
                                 tuple(p.lead_right[s]), eta=p.eta, v_left=p.v_left[s],
                                 v_right=p.v_right[s])
        assert abs(float(t[i].detach()) - tr) <= 1e-9 * max(abs(tr), 1e-3)
        for g, ref in ((ons.grad, go), (cpl.grad, gc)):
            g = g.cpu().numpy()
            scale = max(np.abs(ref).max(), 1e-12)
            assert np.abs(g[s] - ref).max() <= 1e-9 * scale, i
            others = np.delete(g, s, axis=0)
            assert np.all(others == 0)


def test_kappa_gradient_matches_finite_differences(gpu):
    """At eta = 1e-6 the channel steps of T(omega, k) cross the 12 frequencies of the
    small problem when the force constants move by 1e-4, so kappa is not smooth on
    the scale of a finite difference; eta = 1e-3 rounds the steps off over a width
    far above the steps taken (1e-5 of the largest constant)."""
    p = interface3d_problem(seed=3, cell=(2, 1, 1), n_layers=6, nk=2, n_omega=12,
                            n_perturbed=2, eta=1e-3)
    f00, f01 = _leaves(p)
    k = _kappa(p, f00, f01)
    k.backward()
    g00, g01 = f00.grad.cpu().numpy(), f01.grad.cpu().numpy()
    # the perturbed layers carry gradient, and the pristine transverse terms too
    assert np.abs(g00[-1]).max() > 0 and np.abs(g01).max() > 0
    # along the gradient itself, a scaling of the perturbed layers' same-layer
    # constants, a scaling of every next-layer constant, and a random direction
    gn = np.sqrt(np.sum(g00 ** 2) + np.sum(g01 ** 2))
    d_perturbed = np.zeros_like(p.fc00)
    d_perturbed[-2:] = p.fc00[-2:]
    rng = np.random.default_rng(5)
    for d00, d01 in ((g00 / gn, g01 / gn), (d_perturbed, np.zeros_like(p.fc01)),
                     (np.zeros_like(p.fc00), p.fc01),
                     (rng.standard_normal(p.fc00.shape), rng.standard_normal(p.fc01.shape))):
        h = 1e-5 * np.abs(p.fc00).max() / max(np.abs(d00).max(), np.abs(d01).max())

        def k_at(e):
            with torch.no_grad():
                a = torch.as_tensor(p.fc00 + e * d00, device="cuda")
                b = torch.as_tensor(p.fc01 + e * d01, device="cuda")
                return float(_kappa(p, a, b))

        fd = richardson(k_at, h)
        an = float(np.sum(g00 * d00) + np.sum(g01 * d01))
        print(f"d kappa: finite difference {fd:.10e}, adjoint {an:.10e}, "
              f"rel {abs(fd - an) / abs(an):.2e}", flush=True)
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_one_pass_kappa_equals_autograd_path(gpu):
    """kappa as sum_i c_i T_i with c_i = conductance weight / n_k: one adjoint pass
    gives the value and the gradient of the forward + backward path."""
    p = _small()
    f00, f01 = _leaves(p)
    k = _kappa(p, f00, f01)
    k.backward()
    g00, g01 = f00.grad.clone(), f01.grad.clone()
    f00.grad = f01.grad = None
    nk = len(p.kpoints)
    wts = _engine.conductance_weights(
        torch.as_tensor(p.omegas * OMEGA_UNIT_RAD_S, device="cuda"),
        torch.tensor([300.0], dtype=torch.float64, device="cuda"))[0]
    coef = wts.repeat_interleave(nk) / nk
    ons, cpl = interface3d_blocks(p, f00, f01)
    k_of, om = _instances(p)
    k1 = transmission_sum_torch(coef, om, k_of, ons, cpl, p.lead_left, p.lead_right, p.v_left,
                                p.v_right, eta=p.eta)
    k1.backward()
    assert abs(float(k1.detach()) - float(k.detach())) <= 1e-10 * abs(float(k.detach()))
    for a, b in ((f00.grad, g00), (f01.grad, g01)):
        assert float((a - b).abs().max()) <= 1e-10 * float(b.abs().max())
    with pytest.raises(ValueError):
        transmission_sum_torch(coef[:-1], om, k_of, ons, cpl, p.lead_left, p.lead_right,
                               p.v_left, p.v_right)
    # instances in another order (not omega-major with the k points repeated) give the
    # same value and gradient
    perm = np.random.default_rng(1).permutation(om.size)
    f00.grad = f01.grad = None
    ons, cpl = interface3d_blocks(p, f00, f01)
    k2 = transmission_sum_torch(coef[torch.as_tensor(perm, device="cuda")], om[perm], k_of[perm],
                                ons, cpl, p.lead_left, p.lead_right, p.v_left, p.v_right,
                                eta=p.eta)
    k2.backward()
    assert abs(float(k2.detach()) - float(k.detach())) <= 1e-10 * abs(float(k.detach()))
    for a, b in ((f00.grad, g00), (f01.grad, g01)):
        assert float((a - b).abs().max()) <= 1e-10 * float(b.abs().max())


def test_chunked_passes_equal_one_launch(gpu, monkeypatch):
    """Backward and one-pass sums in chunks of one instance give the gradients of a
    single launch (the per-set sums across chunks)."""
    from negfpy_amd.core import adjoint as adj

    p = _small()
    nk = len(p.kpoints)
    coef = torch.linspace(0.5, 1.5, p.omegas.size * nk, dtype=torch.float64, device="cuda")
    k_of, om = _instances(p)

    def grads(fn):
        f00, f01 = _leaves(p)
        ons, cpl = interface3d_blocks(p, f00, f01)
        fn(ons, cpl).backward()
        return f00.grad.clone(), f01.grad.clone()

    def via_sets(ons, cpl):
        t = transmission_sets_torch(om, k_of, ons, cpl, p.lead_left, p.lead_right, p.v_left,
                                    p.v_right, eta=p.eta)
        return (coef * t).sum()

    def via_sum(ons, cpl):
        return transmission_sum_torch(coef, om, k_of, ons, cpl, p.lead_left, p.lead_right,
                                      p.v_left, p.v_right, eta=p.eta)

    ref = {f: grads(f) for f in (via_sets, via_sum)}
    monkeypatch.setattr(adj, "ADJOINT_CHUNK_BYTES", 1)
    monkeypatch.setattr(adj, "ADJOINT_FREE_FRACTION", 0.0)
    for f in (via_sets, via_sum):
        for a, b in zip(grads(f), ref[f]):
            assert float((a - b).abs().max()) <= 1e-12 * float(b.abs().max())


def test_argument_errors(gpu):
    p = _small()
    f00, f01 = _leaves(p)
    ons, cpl = interface3d_blocks(p, f00, f01)
    k_of, om = _instances(p)
    with pytest.raises(ValueError):
        transmission_sets_torch(om, k_of, ons[0], cpl, p.lead_left, p.lead_right, p.v_left,
                                p.v_right)
    with pytest.raises(ValueError):
        transmission_sets_torch(om, k_of + 7, ons, cpl, p.lead_left, p.lead_right, p.v_left,
                                p.v_right)
    with pytest.raises(ValueError):
        transmission_sets_torch(om, k_of, ons, cpl, p.lead_left[:, :2], p.lead_right,
                                p.v_left, p.v_right)
