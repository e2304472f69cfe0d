"""Oracle and host logic against vectors written by the reference package itself.

The vectors (``tests/golden/reference_vectors.npz``, ``reference_adaptive.json``)
come from ``tests/golden/make_reference_vectors.py``, which ran the reference's
own functions on seeded inputs.  They fix what closed forms cannot: the
Sancho-Rubio truncation and SuperLU rounding on random blocks, the
``omega_scale`` convention for default couplings, the eigen routes' fallback at
eta = 0, the DOS normalisation, the device Green's function embedding and the
adaptive drivers' decisions and statistics.
"""

import json
import os

import numpy as np
import pytest

from oracle import negf_oracle as o

HERE = os.path.dirname(os.path.abspath(__file__))
REF = np.load(os.path.join(HERE, "golden", "reference_vectors.npz"))
with open(os.path.join(HERE, "golden", "reference_adaptive.json")) as _f:
    ADAPTIVE = json.load(_f)

# The reference's own values are reproducible to rounding; this bar is relative
# to max(|T|, 1e-3) like the GPU parity bar.
TOL = 1e-10


def _rel(a, b, floor=1e-3):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def md_inputs():
    ons = list(REF["md_onsite"])
    cps = list(REF["md_coupling"])
    lead = (REF["md_d00"], REF["md_d01"], None)
    return ons, cps, lead


CASES = {
    "md_t_sr": dict(eta=1e-6),
    "md_t_eta8": dict(eta=1e-8, eta_device=1e-9),
    "md_t_scaled": dict(eta=1e-6, omega_scale=1.7),
    "md_t_generalized_eigen": dict(eta=1e-6, method="generalized_eigen"),
    "md_t_generalized_eigen_svd": dict(eta=1e-6, method="generalized_eigen_svd"),
    "md_t_legacy_eigen_svd": dict(eta=1e-6, method="legacy_eigen_svd"),
}


@pytest.mark.parametrize("key", sorted(CASES))
def test_oracle_transmission_matches_reference(key):
    ons, cps, lead = md_inputs()
    t = [o.transmission_sparse(w, ons, cps, lead, lead, **CASES[key]) for w in REF["md_omegas"]]
    assert _rel(t, REF[key]) < TOL


def test_omega_scale_changes_default_couplings():
    """The reference divides the default device-to-lead couplings by s^2 twice, so the
    scaled transmission is not the unscaled one."""
    assert _rel(REF["md_t_scaled"], REF["md_t_sr"]) > 1e-2


def test_oracle_contacts_and_two_leads_match_reference():
    ons, cps, lead = md_inputs()
    t = [o.transmission_sparse(w, ons, cps, lead, lead, eta=1e-6, v_left=REF["md_vl"],
                               v_right=REF["md_vr"], idx_left=[0, 2, 4], idx_right=[1, 3, 5])
         for w in REF["md_omegas"]]
    assert _rel(t, REF["md_t_contacts"]) < TOL
    lead_r = (REF["md_d00_r"], REF["md_d01_r"], None)
    t = [o.transmission_sparse(w, ons, cps, lead, lead_r, eta=1e-6) for w in REF["md_omegas"]]
    assert _rel(t, REF["md_t_two_leads"]) < TOL


@pytest.mark.parametrize("method", ["sancho_rubio", "generalized_eigen", "generalized_eigen_svd",
                                    "legacy_eigen_svd"])
def test_oracle_surface_gf_matches_reference(method):
    for w, g_ref in zip(REF["sg_omegas"], REF[f"sg_{method}"]):
        g = o.surface_gf(w, REF["md_d00"], REF["md_d01"], eta=1e-7, method=method)
        assert np.max(np.abs(g - g_ref)) <= 1e-10 * np.max(np.abs(g_ref))


@pytest.mark.parametrize("method", ["generalized_eigen", "generalized_eigen_svd",
                                    "legacy_eigen_svd"])
def test_oracle_eigen_fallback_at_zero_eta(method):
    w_edge = float(REF["md_omegas"][-1] / 1.15)
    g = o.surface_gf(0.5 * w_edge, REF["md_d00"], REF["md_d01"], eta=0.0, method=method)
    g_ref = REF[f"sg0_{method}"]
    assert np.max(np.abs(g - g_ref)) <= 1e-8 * np.max(np.abs(g_ref))


def test_oracle_dos_matches_reference():
    for w, d_ref, dn_ref in zip(REF["md_omegas"], REF["dos"], REF["dos_norm"]):
        g = o.surface_gf(w, REF["md_d00"], REF["md_d01"], eta=1e-7)
        assert abs(o.surface_dos(w, g) - d_ref) <= 1e-10 * max(abs(d_ref), 1e-3)
        assert abs(o.surface_dos(w, g, normalize_per_mode=True) - dn_ref) <= \
            1e-10 * max(abs(dn_ref), 1e-3)


def test_oracle_heat_current_matches_reference():
    j = [o.heat_current_from_spectrum(REF["hc_omegas"], REF["hc_t"], 310.0, 290.0),
         o.heat_current_from_spectrum(REF["hc_omegas"], REF["hc_t"], 5.0, 1.0)]
    assert np.max(np.abs(np.array(j) - REF["hc_j"]) / np.abs(REF["hc_j"])) < 1e-12


def test_oracle_device_green_matches_reference():
    ons, cps, lead = md_inputs()
    g, sl, sr = o.device_green_dense(float(REF["dg_omega"]), ons, cps, lead, lead, eta=1e-6,
                                     v_left=REF["md_vl"], v_right=REF["md_vr"],
                                     idx_left=[0, 2, 4], idx_right=[25, 27, 29])
    scale = np.max(np.abs(REF["dg_g"]))
    assert np.max(np.abs(g - REF["dg_g"])) <= 1e-10 * scale
    assert np.max(np.abs(sl - REF["dg_sigma_l"])) <= 1e-12 * np.max(np.abs(REF["dg_sigma_l"]))
    assert np.max(np.abs(sr - REF["dg_sigma_r"])) <= 1e-12 * np.max(np.abs(REF["dg_sigma_r"]))


def test_host_builders_match_reference():
    from negfpy_amd.models import (
        CubicLatticeParams,
        MultiDofCellParams,
        SquareLatticeParams,
        cubic_lattice_device,
        cubic_lattice_lead,
        multidof_device,
        multidof_lead_blocks,
        square_lattice_device,
        square_lattice_lead,
    )

    p = MultiDofCellParams(masses=REF["md_masses"], fc00=REF["md_fc00"], fc01=REF["md_fc01"],
                           dof_per_atom=3)
    lead = multidof_lead_blocks(p)
    dev = multidof_device(5, p, layer_masses=REF["md_layer_masses"])
    assert np.array_equal(lead.d00, REF["md_d00"]) and np.array_equal(lead.d01, REF["md_d01"])
    assert np.array_equal(np.stack(dev.onsite_blocks), REF["md_onsite"])
    assert np.array_equal(np.stack(dev.coupling_blocks), REF["md_coupling"])
    pr = MultiDofCellParams(masses=REF["md_masses"] * 1.6, fc00=REF["md_fc00"],
                            fc01=REF["md_fc01"], dof_per_atom=3)
    assert np.array_equal(multidof_lead_blocks(pr).d00, REF["md_d00_r"])

    sq = SquareLatticeParams(mass=1.3, spring_x=0.9, spring_y=0.6)
    cu = CubicLatticeParams(mass=0.8, spring_x=1.1, spring_y=0.5, spring_z=0.7,
                            onsite_pinning=0.05)
    sq_lead, cu_lead = square_lattice_lead(sq), cubic_lattice_lead(cu)
    sq_dev = square_lattice_device(4, sq, defect_index=2, defect_mass=2.0)
    cu_dev = cubic_lattice_device(3, cu, defect_index=1, defect_mass=0.5)
    for i, k in enumerate(REF["kb_ky"]):
        np.testing.assert_allclose(np.stack(sq_lead.blocks_builder((k,))[:2]),
                                   REF["kb_sq_lead"][i], rtol=1e-15, atol=1e-15)
        np.testing.assert_allclose(np.stack(sq_dev.device_builder((k,)).onsite_blocks),
                                   REF["kb_sq_dev_onsite"][i], rtol=1e-15, atol=1e-15)
        np.testing.assert_allclose(np.stack(cu_lead.blocks_builder((k, 0.4 * k))[:2]),
                                   REF["kb_cu_lead"][i], rtol=1e-15, atol=1e-15)
        np.testing.assert_allclose(np.stack(cu_dev.device_builder((k, 0.4 * k)).onsite_blocks),
                                   REF["kb_cu_dev_onsite"][i], rtol=1e-15, atol=1e-15)


def _ribbon_checksum(rp):
    import hashlib

    h = hashlib.sha256()
    for a in (rp.lead.d00, rp.lead.d01, *rp.device.onsite_blocks, *rp.device.coupling_blocks,
              rp.omegas):
        h.update(np.ascontiguousarray(np.asarray(a, dtype=np.complex128)).tobytes())
    return h.digest()


def test_ribbon_workload_is_the_reference_input():
    """The 24-DOF x 32-layer ribbon that the benchmark solves is the input the reference
    was run on, bit for bit, and the oracle reproduces the reference's T on it."""
    from negfpy_amd.workloads import ribbon_problem

    rp = ribbon_problem(seed=0, n_omega=4096)
    assert _ribbon_checksum(rp) == REF["rb_checksum"].tobytes()
    ons, cps = rp.device.onsite_blocks, rp.device.coupling_blocks
    lead = (rp.lead.d00, rp.lead.d01, None)
    for i, t_ref in zip(REF["rb_index"], REF["rb_t"]):
        t = o.transmission_sparse(rp.omegas[i], ons, cps, lead, lead, eta=rp.eta)
        assert abs(t - t_ref) <= TOL * max(abs(t_ref), 1e-3), (i, t, t_ref)


# --------------------------------------------------------------------------
# adaptive drivers: host logic with the oracle standing in for the batched solve
# --------------------------------------------------------------------------

def _oracle_solve_points(kps, k_of, omega, eta, eta_dev, *, precision, tol=1e-12, maxiter=200,
                         route=0):
    t_all = np.full(len(k_of), np.nan)
    st = np.zeros(len(k_of), dtype=np.int32)
    for j, k in enumerate(k_of):
        kp = kps[k]
        nb, nbig = kp.onsite.shape[0], kp.onsite.shape[1]
        lead_l = tuple(kp.lead_l)
        lead_r = tuple(kp.lead_r)
        try:
            t_all[j] = o.transmission_sparse(
                omega[j] / kp.scale, list(kp.onsite), list(kp.coupling), lead_l, lead_r,
                eta=eta[j] / kp.scale, eta_device=eta_dev[j] / kp.scale, v_left=kp.v_l,
                v_right=kp.v_r, idx_left=kp.pos_l, idx_right=(nb - 1) * nbig + kp.pos_r)
        except RuntimeError:
            st[j] = 1
    return t_all, st


@pytest.fixture
def oracle_backed(monkeypatch):
    from negfpy_amd.core import negf

    monkeypatch.setattr(negf, "solve_points", _oracle_solve_points)
    monkeypatch.setattr(negf, "_max_block", lambda precision: 48)
    # the host's rung-by-rung eta loop (the in-launch ladder needs the GPU)
    monkeypatch.setattr(negf._KSolver, "in_kernel_ladders", False)
    return negf


def _sq_problem():
    from negfpy_amd.models import SquareLatticeParams, square_lattice_device, square_lattice_lead

    sq = SquareLatticeParams(mass=1.3, spring_x=0.9, spring_y=0.6)
    return (square_lattice_device(4, sq, defect_index=2, defect_mass=2.0),
            square_lattice_lead(sq), [(float(k),) for k in REF["sq_kpts"]])


def _norm(info):
    """JSON-read info -> comparable form (float keys, lists for k tuples)."""
    out = dict(info)
    out["eta_histogram"] = {float(k): v for k, v in info["eta_histogram"].items()}
    if "rejected" in info:
        out["rejected"] = [dict(r, kpar=list(r["kpar"])) for r in info["rejected"]]
    return out


def assert_info_matches(info, ref, t_tol=1e-9):
    got = _norm(json.loads(json.dumps({**info, "eta_histogram":
                                       {repr(k): v for k, v in info["eta_histogram"].items()}})))
    want = _norm(ref)
    for key in ("n_total", "n_success", "n_failed", "eta_selected", "n_rejected"):
        assert got.get(key) == want.get(key), key
    assert abs(got["success_fraction"] - want["success_fraction"]) < 1e-15
    assert got["eta_histogram"] == want["eta_histogram"]
    for a, b in zip(got.get("rejected", []), want.get("rejected", [])):
        assert (a["reason"], a["eta"], a["omega"]) == (b["reason"], b["eta"], b["omega"])
        assert np.allclose(a["kpar"], b["kpar"], rtol=0, atol=1e-15)
        assert (a["t"] is None) == (b["t"] is None)
        if a["t"] is not None:
            assert abs(a["t"] - b["t"]) <= t_tol * max(abs(b["t"]), 1e-3)
    assert len(got.get("rejected", [])) == len(want.get("rejected", []))


def test_adaptive_matches_reference(oracle_backed):
    dev, lead, kpts = _sq_problem()
    ref = ADAPTIVE["sq_default"]
    t, info = oracle_backed.transmission_kavg_adaptive(
        ref["omega"], dev, lead, lead, kpoints=kpts, eta_values=(1e-8, 1e-6),
        nonnegative_tolerance=1e-6, collect_rejected=True)
    assert abs(t - ref["t"]) <= TOL * max(abs(ref["t"]), 1e-3)
    assert_info_matches(info, ref["info"])


def test_adaptive_channel_limit_matches_reference(oracle_backed):
    dev, lead, kpts = _sq_problem()
    ref = ADAPTIVE["sq_channel_limit"]
    t, info = oracle_backed.transmission_kavg_adaptive(
        ref["omega"], dev, lead, lead, kpoints=kpts, eta_values=(1e-8, 1e-7),
        max_channel_factor=0.5, nonnegative_tolerance=1e-6, collect_rejected=True)
    assert abs(t - ref["t"]) <= TOL * max(abs(ref["t"]), 1e-3)
    assert_info_matches(info, ref["info"])


def test_adaptive_global_eta_matches_reference(oracle_backed):
    dev, lead, kpts = _sq_problem()
    ref = ADAPTIVE["sq_global_eta"]
    t, info = oracle_backed.transmission_kavg_adaptive_global_eta(
        ref["omega"], dev, lead, lead, kpoints=kpts, eta_values=(1e-8, 1e-7),
        min_success_fraction=0.2, max_channel_factor=0.5, nonnegative_tolerance=1e-6,
        collect_rejected=True)
    assert abs(t - ref["t"]) <= TOL * max(abs(ref["t"]), 1e-3)
    assert_info_matches(info, ref["info"])


def test_adaptive_chain_defaults_match_reference(oracle_backed):
    from negfpy_amd.models import ChainParams, device_perfect_chain, lead_blocks

    chain = ChainParams(mass=1.0, spring=1.0)
    ref = ADAPTIVE["chain_plain"]
    t, info = oracle_backed.transmission_kavg_adaptive(
        ref["omega"], device_perfect_chain(8, chain), lead_blocks(chain), lead_blocks(chain),
        kpoints=[(), ()])
    assert abs(t - ref["t"]) <= TOL * max(abs(ref["t"]), 1e-3)
    assert_info_matches(info, ref["info"])


@pytest.mark.parametrize("tag", ["h", "x"])
def test_host_material_kspace_builder_matches_reference(tag):
    """The general force-constant k-space builder (material_kspace.py:99-193), implicit
    (Phi_10 = Phi_01^H) and explicit 10 terms, pinning, per-layer masses, kz."""
    from negfpy_amd.models import (MaterialKspaceParams, material_kspace_device,
                                   material_kspace_lead)

    mp = material_params(MaterialKspaceParams, explicit_10=(tag == "x"))
    ml = material_kspace_lead(mp)
    md = material_kspace_device(4, mp, layer_masses=REF["mk_layer_masses"])
    for i, k in enumerate(MK_K):
        np.testing.assert_array_equal(np.stack(ml.blocks_builder(k)), REF[f"mk_{tag}_lead"][i])
        d = md.device_builder(k)
        np.testing.assert_array_equal(np.stack(d.onsite_blocks), REF[f"mk_{tag}_onsite"][i])
        np.testing.assert_array_equal(np.stack(d.coupling_blocks), REF[f"mk_{tag}_coupling"][i])


MK_K = [(0.3,), (1.1, -0.7), (), (-2.0, 0.5)]


def material_params(cls, explicit_10=False):
    a0, a1, a2 = REF["mk_a"]
    b0, b1, b2 = REF["mk_b"]
    fc00 = {(0, 0): a0, (1, 0): a1, (-1, 0): a1.T, (0, 1): a2, (0, -1): a2.T}
    fc01 = {(0, 0): b0, (1, 0): b1, (0, 1): b2}
    fc10 = {(0, 0): b0.T, (-1, 0): b1.T, (0, -1): b2.T} if explicit_10 else None
    return cls(masses=REF["mk_masses"], fc00_terms=fc00, fc01_terms=fc01, fc10_terms=fc10,
               dof_per_atom=3, onsite_pinning=0.03)


def test_oracle_material_kavg_matches_reference():
    ons = REF["mk_h_onsite"]
    cps = REF["mk_h_coupling"]
    lead = REF["mk_h_lead"]
    for w, t_ref in zip(REF["mk_omegas"], REF["mk_t_kavg"]):
        ts = [o.transmission_sparse(w, list(ons[i]), list(cps[i]), tuple(lead[i]), tuple(lead[i]),
                                    eta=1e-6) for i in range(len(MK_K))]
        assert abs(np.mean(ts) - t_ref) <= TOL * max(abs(t_ref), 1e-3)


@pytest.mark.parametrize("scheme", ["adaptive", "adaptive_global"])
def test_adaptive_spectrum_equals_per_omega_calls(oracle_backed, scheme):
    """The batched spectrum driver gives, omega by omega, what the per-omega functions
    give (values, statistics, rejection lists, and the exceptions as error records)."""
    dev, lead, kpts = _sq_problem()
    ws = [0.9, 1.1, 0.0, 1.6, -0.3, 0.4]
    kw = dict(eta_values=(1e-8, 1e-7), nonnegative_tolerance=1e-6, max_channel_factor=0.5,
              min_success_fraction=0.2, collect_rejected=True)
    vals, recs = oracle_backed.transmission_kavg_adaptive_spectrum(ws, dev, lead, lead, kpts,
                                                                   eta_scheme=scheme, **kw)
    single = (oracle_backed.transmission_kavg_adaptive if scheme == "adaptive"
              else oracle_backed.transmission_kavg_adaptive_global_eta)
    n_err = 0
    for w, v, rec in zip(ws, vals, recs):
        if w <= 0.0:
            assert v == 0.0 and rec is None
            continue
        try:
            t, info = single(w, dev, lead, lead, kpoints=kpts, **kw)
        except RuntimeError as exc:
            n_err += 1
            assert np.isnan(v) and rec == {"error": str(exc)}
            continue
        assert v == t
        assert rec["info"] == info
    assert n_err >= 1  # the channel limit makes at least one omega fail


def _moments_by_reference_loop(ws, dev, lead, kpts, ladder, min_frac, tol):
    """The reference's moment loop (workflows/ifc_bulk.py:1723-1776) on oracle T."""
    from negfpy_amd.core.negf import resolve_k

    out = []
    for w in ws:
        if w <= 0.0:
            out.append((0.0, 0.0, np.nan, 0))
            continue
        got = None
        for eta in ladder:
            vals = []
            for k in kpts:
                kp = resolve_k(dev, lead, lead, k, eta=eta, eta_device=None,
                               device_to_lead_left=None, device_to_lead_right=None,
                               contact_left_indices=None, contact_right_indices=None,
                               omega_scale=None, max_block=48)
                nb, nbig = kp.onsite.shape[0], kp.onsite.shape[1]
                t = o.transmission_sparse(w, list(kp.onsite), list(kp.coupling), tuple(kp.lead_l),
                                          tuple(kp.lead_r), eta=eta, v_left=kp.v_l,
                                          v_right=kp.v_r, idx_left=kp.pos_l,
                                          idx_right=(nb - 1) * nbig + kp.pos_r)
                if not np.isfinite(t) or t < -tol:
                    continue
                vals.append(max(t, 0.0))
            if vals and len(vals) / len(kpts) >= min_frac:
                got = (float(np.mean(vals)), float(np.mean(np.square(vals))), eta, len(vals))
                break
        out.append(got if got is not None else (np.nan, np.nan, np.nan, 0))
    return out


def test_k_moments_spectrum_matches_reference_loop(oracle_backed):
    dev, lead, kpts = _sq_problem()
    ws = [0.9, 1.1, 0.0, 1.6, 0.4]
    res = oracle_backed.transmission_k_moments_spectrum(ws, dev, lead, lead, kpts,
                                                        eta_values=(1e-8, 1e-7),
                                                        min_success_fraction=0.7,
                                                        nonnegative_tolerance=1e-8)
    ref = _moments_by_reference_loop(ws, dev, lead, kpts, (1e-8, 1e-7), 0.7, 1e-8)
    for i, (t1, t2, eta, nk) in enumerate(ref):
        if np.isnan(t1):
            assert np.isnan(res["t_mean"][i]) and res["n_k_success"][i] == 0
            continue
        assert abs(res["t_mean"][i] - t1) <= 1e-13 * max(t1, 1e-3)
        assert abs(res["t2_mean"][i] - t2) <= 1e-13 * max(t2, 1e-3)
        assert res["n_k_success"][i] == nk
        if ws[i] > 0:
            assert res["eta_selected"][i] == eta
