"""GPU parity: the HIP path (through the C ABI) against the FP64 CPU oracle, with the
spread-aware bar of ``parity_util``."""

import numpy as np
import pytest
import torch

from negfpy_amd.core import _engine
from negfpy_amd.core import (
    Device1D,
    LeadBlocks,
    lead_surface_dos,
    surface_gf,
    thermal_conductance_from_spectrum,
    heat_current_from_spectrum,
    transmission,
    transmission_kavg,
    transmission_spectrum,
)
from negfpy_amd.core.surface_gf import pack_lead
from negfpy_amd.models import (
    ChainParams,
    MultiDofCellParams,
    SquareLatticeParams,
    device_perfect_chain,
    lead_blocks,
    multidof_device,
    multidof_lead_blocks,
    square_lattice_device,
    square_lattice_lead,
)
from negfpy_amd.workloads import omega_max, random_lead, ribbon_problem, sheet_problem
from oracle import negf_oracle as o
from parity_util import T_TOL, assert_parity, oracle_t, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 3, 6, 12, 16, 24, 30, 48])
@pytest.mark.parametrize("eta", [1e-6, 1e-7])
def test_surface_gf_batch_matches_oracle(gpu, n, eta):
    d00, d01 = random_lead(11 + n, n)
    wmax = omega_max(LeadBlocks(d00=d00, d01=d01))
    omegas = np.linspace(0.03, 1.15, 12) * wmax
    res = _engine.surface_gf_batch(pack_lead(d00, d01)[None], np.zeros(omegas.size, np.int32),
                                   omegas, np.full(omegas.size, eta))
    assert res.status.cpu().numpy().tolist() == [0] * omegas.size
    g = res.g.cpu().numpy()
    its = res.iters.cpu().numpy()
    dos = res.dos.cpu().numpy()
    for i, w in enumerate(omegas):
        gr, it = o.sancho_rubio(w, d00, d01, None, eta, return_iters=True)
        err = np.linalg.norm(g[i] - gr) / np.linalg.norm(gr)
        assert err < 1e-10, (n, w, err)
        assert abs(int(its[i]) - it) <= 1
        assert abs(dos[i] - o.surface_dos(w, gr)) <= 1e-10 * max(abs(o.surface_dos(w, gr)), 1e-3)


@pytest.mark.parametrize("eta", [1e-6, 1e-8])
def test_chain_surface_gf_closed_form(gpu, eta):
    lb = lead_blocks(ChainParams(mass=1.0, spring=1.0))
    for w in (0.3, 1.0, 1.7, 2.5):
        g = surface_gf(w, lb.d00, lb.d01, eta=eta)
        z = (w + 1j * eta) ** 2
        r = np.sqrt((z - 2) ** 2 - 4 + 0j)
        gc = min([((z - 2) + r) / 2, ((z - 2) - r) / 2], key=abs)
        assert abs(g[0, 0] - gc) / abs(gc) < 1e-11


def _ribbon_small(n_atoms, n_layers, seed=3):
    p = ribbon_problem(seed=seed, n_atoms=n_atoms, n_layers=n_layers, n_omega=24, eta=1e-6)
    return p


@pytest.mark.parametrize("n_atoms,n_layers", [(1, 5), (2, 9), (4, 6), (8, 8), (16, 3)])
def test_transmission_multidof_matches_oracle(gpu, n_atoms, n_layers):
    p = _ribbon_small(n_atoms, n_layers)
    t = transmission_spectrum(p.omegas, p.device, p.lead, p.lead, eta=p.eta)
    lead = (p.lead.d00, p.lead.d01, None)
    refs = [oracle_t(w, p.device.onsite_blocks, p.device.coupling_blocks, lead, lead, eta=p.eta)
            for w in p.omegas]
    assert_parity(t, refs)


def test_transmission_single_point_api(gpu):
    p = _ribbon_small(2, 6)
    lead = (p.lead.d00, p.lead.d01, None)
    for w in p.omegas[::6]:
        t = transmission(w, p.device, p.lead, p.lead, eta=1e-7, eta_device=3e-7)
        ref = oracle_t(w, p.device.onsite_blocks, p.device.coupling_blocks, lead, lead,
                       eta=1e-7, eta_device=3e-7)
        assert isinstance(t, float)
        assert_parity(t, [ref])


def test_transmission_different_leads(gpu):
    """Left and right leads differ: two decimations per point."""
    pl = _ribbon_small(2, 6, seed=4)
    pr = _ribbon_small(2, 6, seed=5)
    ll = (pl.lead.d00, pl.lead.d01, None)
    lr = (pr.lead.d00, pr.lead.d01, None)
    for w in pl.omegas[::5]:
        t = transmission(w, pl.device, pl.lead, pr.lead, eta=1e-6)
        assert_parity(t, [oracle_t(w, pl.device.onsite_blocks, pl.device.coupling_blocks, ll, lr,
                                   eta=1e-6)])


def test_contacts_and_couplings(gpu):
    """Explicit couplings, a 1-DOF contact inside a 2-DOF layer (rectangular V), and
    contacts that reach into the second layer (merged blocks)."""
    rng = np.random.default_rng(2)
    d00 = np.array([[2.2, -0.3], [-0.3, 1.9]])
    d01 = np.array([[-1.0, 0.1], [0.05, -0.9]])
    lead = LeadBlocks(d00=d00, d01=d01)
    ons = [d00 + np.diag(rng.uniform(-0.1, 0.1, 2)) for _ in range(6)]
    dev = Device1D(ons, [d01] * 5)
    ld = (d00, d01, None)
    w = 0.9
    for kw in ({"device_to_lead_left": d01.conj().T, "device_to_lead_right": d01},
               {"device_to_lead_left": d01.conj().T[:1], "device_to_lead_right": d01[1:],
                "contact_left_indices": [1], "contact_right_indices": [10]},
               {"device_to_lead_left": np.vstack([d01.conj().T, d01.conj().T[:1]]),
                "contact_left_indices": [0, 1, 3]}):
        okw = {"v_left": kw.get("device_to_lead_left"), "v_right": kw.get("device_to_lead_right"),
               "idx_left": kw.get("contact_left_indices"),
               "idx_right": kw.get("contact_right_indices")}
        t = transmission(w, dev, lead, lead, eta=1e-6, **kw)
        assert_parity(t, [oracle_t(w, ons, [d01] * 5, ld, ld, eta=1e-6, **okw)])


def test_omega_scale_parity(gpu):
    p = _ribbon_small(2, 5)
    lead = (p.lead.d00, p.lead.d01, None)
    for w in p.omegas[::8]:
        t = transmission(w, p.device, p.lead, p.lead, eta=1e-6, omega_scale=2.5)
        assert_parity(t, [oracle_t(w, p.device.onsite_blocks, p.device.coupling_blocks, lead,
                                   lead, eta=1e-6, omega_scale=2.5)])


def test_kspace_lattice_matches_oracle(gpu):
    sp = SquareLatticeParams(mass=1.0, spring_x=1.0, spring_y=0.6)
    lead = square_lattice_lead(sp)
    dev = square_lattice_device(25, sp, defect_index=7, defect_mass=1.7)
    ks = [(-np.pi,), (-1.0,), (0.0,), (0.4,), (np.pi / 2,)]
    for w in (0.5, 1.2, 2.1):
        tk = transmission_kavg(w, dev, lead, lead, ks, eta=1e-7)
        vals = []
        for k in ks:
            d00, d01 = lead.blocks(k)
            dv = dev.device(k)
            vals.append(o.transmission_sparse(w, dv.onsite_blocks, dv.coupling_blocks,
                                              (d00, d01, None), (d00, d01, None), eta=1e-7))
        assert rel(tk, np.mean(vals)) < T_TOL


def test_sheet_kspace_spectrum_matches_oracle(gpu):
    """Graphene-like sheet between interface-contact leads: the k-averaged spectrum
    (blocks of every k built on the GPU, all (omega, k) in one batch) against the
    oracle's mean over k, and the pristine sheet's integer channel counts."""
    p = sheet_problem(n_layers=16, n_k=8, n_omega=24)
    w = p.omegas[1::3]
    t = transmission_spectrum(w, p.device, p.lead_left, p.lead_right, kpoints=p.kpoints,
                              eta=p.eta, device_to_lead_left=p.v_left,
                              device_to_lead_right=p.v_right)
    refs = []
    for wi in w:
        per_k = []
        for k in p.kpoints:
            dv = p.device.device_builder(k)
            per_k.append(oracle_t(wi, dv.onsite_blocks, dv.coupling_blocks,
                                  p.lead_left.blocks_builder(k), p.lead_right.blocks_builder(k),
                                  eta=p.eta, v_left=p.v_left(k), v_right=p.v_right(k)))
        refs.append((float(np.mean([a for a, _ in per_k])), max(b for _, b in per_k)))
    assert_parity(t, refs)
    # every (omega, k) of a pristine sheet carries a whole number of channels
    # (within the eta broadening of slow channels near band edges, up to ~1e-2 here)
    assert np.all(np.abs(t * len(p.kpoints) - np.round(t * len(p.kpoints))) < 5e-2)


def test_cubic_kspace_spectrum_matches_oracle(gpu):
    """3-D simple-cubic cells with (ky, kz) transverse terms (the uneven 3-D batch's
    24-DOF part, reduced): the k-averaged spectrum, blocks built on the GPU, against
    the oracle's mean over k."""
    from negfpy_amd.workloads import uneven3d_problems

    p = uneven3d_problems(n_layers=4, n_omega=16)[0]
    ks = p.kpoints[::257][:4]
    w = p.omegas[1::4]
    t = transmission_spectrum(w, p.device, p.lead, p.lead, kpoints=ks, eta=p.eta)
    refs = []
    for wi in w:
        per_k = []
        for k in ks:
            d00, d01, d10 = p.lead.blocks_builder(k)
            dv = p.device.device_builder(k)
            per_k.append(oracle_t(wi, dv.onsite_blocks, dv.coupling_blocks, (d00, d01, d10),
                                  (d00, d01, d10), eta=p.eta))
        refs.append((float(np.mean([a for a, _ in per_k])), max(b for _, b in per_k)))
    assert_parity(t, refs)


def test_identity_blocks_channel_count(gpu):
    params = MultiDofCellParams(masses=np.ones(4), fc00=2.0 * np.eye(12), fc01=-np.eye(12))
    lead = multidof_lead_blocks(params)
    dev = multidof_device(30, params)
    t = transmission(1.0, dev, lead, lead, eta=1e-8)
    assert abs(t - 12.0) < 1e-4


def test_error_paths(gpu):
    lb = lead_blocks(ChainParams(mass=1.0, spring=1.0))
    dev = device_perfect_chain(5, ChainParams(mass=1.0, spring=1.0))
    with pytest.raises(RuntimeError, match="did not converge"):
        surface_gf(1.0, lb.d00, lb.d01, eta=1e-8, maxiter=1)
    with pytest.raises(RuntimeError):
        surface_gf(1.0, np.array([[np.nan]]), lb.d01, eta=1e-8)
    with pytest.raises(ValueError):
        transmission(1.0, dev, lb, lb, surface_gf_method="qz")
    with pytest.raises(ValueError):
        lead_surface_dos(-1.0, lb)


def test_surface_dos_closed_form(gpu):
    lb = lead_blocks(ChainParams(mass=1.0, spring=1.0))
    for w in (0.2, 0.9, 1.5, 1.9):
        assert abs(lead_surface_dos(w, lb, eta=1e-10) - w * w * np.sqrt(4 - w * w) / np.pi) < 1e-7


def test_thermal_kernels_match_oracle(gpu):
    import os

    g = os.path.join(os.path.dirname(__file__), "golden", "study_kappa")
    tr = np.loadtxt(os.path.join(g, "graphene_transmission.tsv"), comments="#")
    kp = np.loadtxt(os.path.join(g, "graphene_thermal_conductance.tsv"), comments="#")
    gk = thermal_conductance_from_spectrum(tr[:, 0], tr[:, 3], kp[:, 0])
    pos = kp[:, 1] > 0
    assert np.all(gk[~pos] == 0.0)
    assert np.max(np.abs(gk[pos] - kp[pos, 1]) / kp[pos, 1]) < 1e-11
    ref = o.thermal_conductance(tr[:, 0], tr[:, 3], kp[:, 0])
    assert np.max(np.abs(gk[pos] - ref[pos]) / ref[pos]) < 1e-12
    w = o.qe_omega_to_rad_s(tr[:, 0])
    for tl, trr in ((300.0, 290.0), (50.0, 400.0)):
        j = heat_current_from_spectrum(w, tr[:, 3], tl, trr)
        jr = o.heat_current_from_spectrum(w, tr[:, 3], tl, trr)
        assert abs(j - jr) <= 1e-12 * abs(jr)


def test_ribbon_workload_parity(gpu):
    """A 256-omega subset of the 24-DOF, 32-layer ribbon benchmark, FP64 mode."""
    p = ribbon_problem(n_omega=4096)
    sel = np.arange(0, 4096, 16)
    t = transmission_spectrum(p.omegas[sel], p.device, p.lead, p.lead, eta=p.eta)
    lead = (p.lead.d00, p.lead.d01, None)
    chk = sel[::8]
    refs = [oracle_t(w, p.device.onsite_blocks, p.device.coupling_blocks, lead, lead, eta=p.eta)
            for w in p.omegas[chk]]
    assert_parity(t[::8], refs)
    assert np.all(t > -1e-9)


def test_mixed_precision_error_bound(gpu):
    """Mixed mode (certified FP32 decimation, FP64 sweep and trace) against FP64 on an
    eighth of the ribbon: <= 1e-6 relative at every point (every frequency:
    tests/test_gpu_mixed_precision.py)."""
    p = ribbon_problem(n_omega=4096)
    sel = np.arange(0, 4096, 8)
    t64 = transmission_spectrum(p.omegas[sel], p.device, p.lead, p.lead, eta=p.eta, precision="fp64")
    t32 = transmission_spectrum(p.omegas[sel], p.device, p.lead, p.lead, eta=p.eta,
                                precision="mixed")
    err = np.abs(t32 - t64) / np.maximum(np.abs(t64), 1e-3)
    print(f"mixed vs fp64: max {err.max():.3e} median {np.median(err):.3e}", flush=True)
    assert np.all(np.isfinite(t32))
    assert err.max() <= 1e-6


@pytest.mark.parametrize("case", ["chain", "multidof", "contacts", "merged"])
def test_device_green_function_matches_oracle(gpu, case):
    from negfpy_amd.core import device_green_function

    if case == "chain":
        params = ChainParams(mass=1.0, spring=1.0)
        lb = lead_blocks(params)
        dev = device_perfect_chain(9, params)
        ons, cps, ll = dev.onsite_blocks, dev.coupling_blocks, (lb.d00, lb.d01, None)
        kw, okw, w = {}, {}, 1.1
    else:
        p = _ribbon_small(2, 5)
        lb, dev = p.lead, p.device
        ons, cps, ll = dev.onsite_blocks, dev.coupling_blocks, (lb.d00, lb.d01, None)
        w = float(p.omegas[7])
        kw, okw = {}, {}
        if case == "contacts":
            v = np.asarray(lb.d01).conj().T[:3]
            kw = {"device_to_lead_left": v, "contact_left_indices": [0, 2, 4]}
            okw = {"v_left": v, "idx_left": [0, 2, 4]}
        if case == "merged":
            v = np.asarray(lb.d01).conj().T[:3]
            kw = {"device_to_lead_left": v, "contact_left_indices": [0, 5, 8]}
            okw = {"v_left": v, "idx_left": [0, 5, 8]}
    g, sl, sr = device_green_function(w, dev, lb, lb, eta=1e-6, **kw)
    gr, slr, srr = o.device_green_dense(w, ons, cps, ll, ll, eta=1e-6, **okw)
    assert g.shape == gr.shape and sl.shape == slr.shape
    assert np.linalg.norm(g - gr) <= 1e-9 * np.linalg.norm(gr)
    assert np.linalg.norm(sl - slr) <= 1e-10 * np.linalg.norm(slr)
    assert np.linalg.norm(sr - srr) <= 1e-10 * np.linalg.norm(srr)
