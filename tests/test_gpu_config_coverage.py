"""Parts of the benchmark configurations and drivers checked nowhere else:

- the 48- and 96-DOF parts of the 3-D uneven batch (simple-cubic cells on k meshes),
  through the staged launches and one persistent launch, against the oracle;
- the gradient workload at its real shape (48 DOF, 64 layers, 28 | 73 amu leads):
  forward T against the oracle and dT/d(device blocks) against autograd through a
  dense complex128 solve;
- the heat-current drivers heat_current_1d / _per_length_2d / _density_3d against
  the oracle's composition (thermal_current.py:70-178: T per frequency, k-averaged
  where the model is k-resolved, then heat_current_from_spectrum, / length or area);
- block-size classes chosen by the autotuner at 24 and 48 DOF against the oracle.
"""

import numpy as np
import pytest
import torch

from adjoint_util import dense_grads
from negfpy_amd import _lib
from negfpy_amd.core import (_engine, heat_current_1d, heat_current_density_3d,
                             heat_current_per_length_2d, transmission_sets_torch,
                             transmission_spectrum)
from negfpy_amd.core.negf import kspace_sets
from negfpy_amd.models import (ChainParams, CubicLatticeParams, SquareLatticeParams,
                               cubic_lattice_device, cubic_lattice_lead, device_perfect_chain,
                               lead_blocks, square_lattice_device, square_lattice_lead)
from negfpy_amd.workloads import interface3d_blocks, interface3d_problem, uneven3d_problems
from oracle import negf_oracle as o
from parity_util import assert_parity, oracle_t

pytestmark = pytest.mark.gpu


@pytest.fixture
def classic(gpu):
    """The persistent launch is compared with the staged path bit for bit elsewhere; here
    both go against the oracle, with the default kernel families."""
    return _lib.load()


@pytest.mark.parametrize("part", [1, 2])
def test_uneven3d_48_and_96_dof_parts_match_oracle(classic, part):
    p = uneven3d_problems(n_layers=4, n_omega=32)[part]
    n = p.params.masses.size * 3
    ks = [p.kpoints[0], p.kpoints[len(p.kpoints) // 2 + 1]]
    w = p.omegas[np.linspace(0, p.omegas.size - 1, 8).astype(int)]  # band-edge points included
    sets = kspace_sets(p.device, p.lead, p.lead, ks)
    nk = len(ks)
    k_of = np.tile(np.arange(nk, dtype=np.int32), w.size)
    om = np.repeat(w, nk)
    eta = np.full(om.size, p.eta)
    solver = _engine.TransmissionSolver(sets)
    staged = solver.run(k_of, om, eta).t.cpu().numpy()
    job = (solver, torch.tensor(k_of, device="cuda"), torch.tensor(om, device="cuda"),
           torch.tensor(eta, device="cuda"), None)
    multi = _engine.transmission_multi([job])[0].t.cpu().numpy()
    refs = []
    for kk, wi in zip(k_of, om):
        d00, d01, d10 = p.lead.blocks_builder(ks[kk])
        dv = p.device.device_builder(ks[kk])
        refs.append(oracle_t(wi, dv.onsite_blocks, dv.coupling_blocks, (d00, d01, d10),
                             (d00, d01, d10), eta=p.eta))
    print(f"{n}-DOF part: T range {staged.min():.3e}..{staged.max():.3e}", flush=True)
    assert_parity(staged, refs)
    assert_parity(multi, refs)


def test_interface_gradient_at_real_shape(gpu):
    """48 DOF x 64 layers, two leads of 28 and 73 amu, the last 4 layers perturbed: T of
    two (omega, k) instances against the oracle, dT/d(all 127 device blocks) against
    autograd through the dense 3072 x 3072 solve to 1e-9."""
    p = interface3d_problem(nk=2, n_omega=8)
    assert p.fc00.shape[-1] == 48 and p.fc00.shape[0] == 64
    nk = len(p.kpoints)
    k_of = np.tile(np.arange(nk, dtype=np.int32), p.omegas.size)
    om = np.repeat(p.omegas, nk)
    ons0, cpl0 = interface3d_blocks(p, torch.as_tensor(p.fc00, device="cuda"),
                                    torch.as_tensor(p.fc01, device="cuda"))
    with torch.no_grad():
        t_all = transmission_sets_torch(om, k_of, ons0, cpl0, p.lead_left, p.lead_right,
                                        p.v_left, p.v_right, eta=p.eta).cpu().numpy()
    # the two instances carrying the most transmission (most of the 64-layer interface's
    # (omega, k) points are evanescent, T ~ 1e-70)
    for i in np.argsort(t_all)[-2:]:
        i = int(i)
        s = int(k_of[i])
        ons = ons0.detach().clone().requires_grad_(True)
        cpl = cpl0.detach().clone().requires_grad_(True)
        t = transmission_sets_torch(om[i:i + 1], k_of[i:i + 1], ons, cpl, p.lead_left,
                                    p.lead_right, p.v_left, p.v_right, eta=p.eta)
        t[0].backward()
        ons_h, cpl_h = ons0[s].detach().cpu().numpy(), cpl0[s].detach().cpu().numpy()
        ref_t = oracle_t(om[i], list(ons_h), list(cpl_h), tuple(p.lead_left[s]),
                         tuple(p.lead_right[s]), eta=p.eta, v_left=p.v_left[s],
                         v_right=p.v_right[s])
        assert_parity([t[0].detach().item()], [ref_t])
        tr, go, gc = dense_grads(om[i], ons_h, cpl_h, tuple(p.lead_left[s]),
                                 tuple(p.lead_right[s]), eta=p.eta, v_left=p.v_left[s],
                                 v_right=p.v_right[s])
        assert abs(t[0].detach().item() - tr) <= 1e-9 * max(abs(tr), 1e-3)
        for g, ref in ((ons.grad, go), (cpl.grad, gc)):
            g = g[s].cpu().numpy()
            scale = max(np.abs(ref).max(), 1e-12)
            err = np.abs(g - ref).max() / scale
            print(f"instance {i}: T {tr:.6e}, gradient rel err {err:.2e}", flush=True)
            assert err <= 1e-9


def _oracle_spectrum(w, onsite_blocks, coupling_blocks, lead, eta):
    return np.array([o.transmission_sparse(wi, onsite_blocks, coupling_blocks, lead, lead, eta=eta)
                     for wi in w])


def test_heat_current_1d_matches_oracle_composition(gpu):
    k = 1.0e26  # spring / mass in s^-2: band top 2e13 rad/s
    params = ChainParams(mass=1.0, spring=k)
    lb = lead_blocks(params)
    dev = device_perfect_chain(6, params)
    w = np.linspace(0.05, 2.15, 41) * 1e13
    eta = 1e-6 * 2e13
    j = heat_current_1d(w, dev, lb, lb, 310.0, 290.0, eta=eta)
    t = _oracle_spectrum(w, dev.onsite_blocks, dev.coupling_blocks, (lb.d00, lb.d01, None), eta)
    jr = o.heat_current_from_spectrum(w, t, 310.0, 290.0)
    assert jr > 0 and abs(j - jr) <= 1e-10 * abs(jr)


@pytest.mark.parametrize("dim", [2, 3])
def test_heat_current_kspace_drivers_match_oracle_composition(gpu, dim):
    if dim == 2:
        params = SquareLatticeParams(mass=1.0, spring_x=1.0e26, spring_y=0.6e26)
        lead, dev = square_lattice_lead(params), square_lattice_device(5, params, 2, 1.4)
        ks = [(ky,) for ky in np.linspace(-np.pi, np.pi, 7, endpoint=False) + np.pi / 7]
    else:
        params = CubicLatticeParams(mass=1.0, spring_x=1.0e26, spring_y=0.5e26, spring_z=0.7e26)
        lead, dev = cubic_lattice_lead(params), cubic_lattice_device(5, params, 2, 1.4)
        ks = [(ky, kz) for ky in (-2.0, 0.3, 1.9) for kz in (-1.1, 0.8)]
    w = np.linspace(0.05, 3.4, 31) * 1e13
    eta = 1e-6 * 3e13
    size = 3.7e-10 if dim == 2 else 1.3e-19
    if dim == 2:
        j = heat_current_per_length_2d(w, dev, lead, lead, ks, size, 305.0, 295.0, eta=eta)
    else:
        j = heat_current_density_3d(w, dev, lead, lead, ks, size, 305.0, 295.0, eta=eta)
    tk = []
    for kp in ks:
        d00, d01 = lead.blocks_builder(kp)
        dk = dev.device_builder(kp)
        tk.append(_oracle_spectrum(w, dk.onsite_blocks, dk.coupling_blocks, (d00, d01, None), eta))
    jr = o.heat_current_from_spectrum(w, np.mean(tk, axis=0), 305.0, 295.0) / size
    assert jr > 0 and abs(j - jr) <= 1e-10 * abs(jr)
    # the spectrum route of the same drivers
    t = transmission_spectrum(w, dev, lead, lead, kpoints=ks, eta=eta)
    assert np.max(np.abs(t - np.mean(tk, axis=0))) <= 1e-9 * max(np.abs(t).max(), 1e-3)


@pytest.mark.parametrize("n", [24, 48])
def test_autotuned_class_matches_oracle(gpu, n):
    """Run the tuner at n = 24 and 48 (FP64), then a ribbon of that size with the class
    it recorded, against the oracle; the table is restored afterwards."""
    from negfpy_amd.core.autotune import (VECTOR, WAVE, block_class, kernel_family,
                                          product_form, reset_block_class, tune_block_class)
    from negfpy_amd.workloads import ribbon_params, omega_max
    from negfpy_amd.models.multidof import multidof_device, multidof_lead_blocks

    try:
        res = tune_block_class(n, "fp64", n_layers=3, n_inst=256, repeats=1)
        best = res["best"]
        if best == WAVE:
            assert kernel_family(n, "fp64") == WAVE
        elif isinstance(best, tuple):
            assert block_class(n, "fp64") == best[0] and product_form(n) == VECTOR
        else:
            assert block_class(n, "fp64") == best and product_form(n) == "matrix_core"
        print(f"n={n}: {res['timings_ms']} -> {best}", flush=True)
        params = ribbon_params(9, n // 3)
        lead = multidof_lead_blocks(params)
        rng = np.random.default_rng(n)
        dev = multidof_device(3, params, layer_masses=rng.uniform(10, 30, (3, n // 3)))
        w = np.linspace(0.05, 1.05, 6) * omega_max(lead)
        t = transmission_spectrum(w, dev, lead, lead, eta=1e-6)
        tup = (lead.d00, lead.d01, None)
        assert_parity(t, [oracle_t(wi, dev.onsite_blocks, dev.coupling_blocks, tup, tup,
                                   eta=1e-6) for wi in w])
    finally:
        reset_block_class(n, "fp64")
