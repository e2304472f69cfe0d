"""The GPU path against vectors written by the reference package itself.

Same vectors as ``test_reference_vectors.py`` (there they pin the oracle); here
the drop-in functions, running the HIP kernels, must reproduce them.  The bar is
the spread-aware one of ``parity_util``: 1e-10 relative to max(|T|, 1e-3), widened
to 50x the oracle's own one-ulp spread where decimation is ill-conditioned.
"""

import json
import os

import numpy as np
import pytest

from negfpy_amd.core import (
    Device1D,
    LeadBlocks,
    device_green_function,
    lead_surface_dos,
    surface_gf,
    transmission,
    transmission_kavg_adaptive,
    transmission_kavg_adaptive_global_eta,
    transmission_spectrum,
)
from negfpy_amd.models import SquareLatticeParams, square_lattice_device, square_lattice_lead
from negfpy_amd.workloads import ribbon_problem
from oracle import negf_oracle as o
from parity_util import T_TOL, oracle_spread
from test_reference_vectors import ADAPTIVE, CASES, REF, assert_info_matches

pytestmark = pytest.mark.gpu


def _md():
    ons, cps = list(REF["md_onsite"]), list(REF["md_coupling"])
    return (ons, cps, Device1D(onsite_blocks=ons, coupling_blocks=cps),
            LeadBlocks(d00=REF["md_d00"], d01=REF["md_d01"]))


def _assert_vs_ref(t_gpu, t_ref, spreads):
    for tg, tr, sp in zip(t_gpu, t_ref, spreads):
        bound = max(T_TOL * max(abs(tr), 1e-3), 50.0 * sp)
        assert abs(tg - tr) <= bound, (tg, tr, sp)


@pytest.mark.parametrize("key", sorted(CASES))
def test_transmission_matches_reference(gpu, key):
    ons, cps, dev, lead = _md()
    kw = dict(CASES[key])
    method = kw.pop("method", "sancho_rubio")
    t = [transmission(w, dev, lead, lead, surface_gf_method=method, **kw) for w in REF["md_omegas"]]
    okw = dict(CASES[key])
    tup = (REF["md_d00"], REF["md_d01"], None)
    sp = [oracle_spread(w, ons, cps, tup, tup, tr, **okw)
          for w, tr in zip(REF["md_omegas"], REF[key])]
    _assert_vs_ref(t, REF[key], sp)


def test_contacts_and_two_leads_match_reference(gpu):
    ons, cps, dev, lead = _md()
    t = [transmission(w, dev, lead, lead, eta=1e-6, device_to_lead_left=REF["md_vl"],
                      device_to_lead_right=REF["md_vr"], contact_left_indices=[0, 2, 4],
                      contact_right_indices=[1, 3, 5]) for w in REF["md_omegas"]]
    _assert_vs_ref(t, REF["md_t_contacts"], [0.0] * len(t))
    lead_r = LeadBlocks(d00=REF["md_d00_r"], d01=REF["md_d01_r"])
    t = [transmission(w, dev, lead, lead_r, eta=1e-6) for w in REF["md_omegas"]]
    tl = (REF["md_d00"], REF["md_d01"], None)
    tr_ = (REF["md_d00_r"], REF["md_d01_r"], None)
    sp = [oracle_spread(w, ons, cps, tl, tr_, tr, eta=1e-6)
          for w, tr in zip(REF["md_omegas"], REF["md_t_two_leads"])]
    _assert_vs_ref(t, REF["md_t_two_leads"], sp)


@pytest.mark.parametrize("method", ["sancho_rubio", "generalized_eigen", "generalized_eigen_svd",
                                    "legacy_eigen_svd"])
def test_surface_gf_matches_reference(gpu, method):
    for w, g_ref in zip(REF["sg_omegas"], REF[f"sg_{method}"]):
        g = surface_gf(w, REF["md_d00"], REF["md_d01"], eta=1e-7, method=method)
        assert np.max(np.abs(g - g_ref)) <= 1e-9 * np.max(np.abs(g_ref)), w


@pytest.mark.parametrize("method", ["generalized_eigen", "generalized_eigen_svd",
                                    "legacy_eigen_svd"])
def test_eigen_fallback_at_zero_eta_is_refused(gpu, method):
    """eta = 0 inside the band: the pencil has eigenvalues of modulus 1 (to rounding)
    between the decaying and the growing ones, and the reference's n-smallest-|lambda|
    fallback (surface_gf.py:123-124) keeps some of them in the order LAPACK's rounding
    gives them -- a choice no spectral projector reproduces.  The GPU route refuses that
    case with NotImplementedError instead of returning another solution; the reference's
    own vector for it is therefore not compared.  Out of the band (all modes evanescent)
    eta = 0 decimates normally."""
    d00, d01 = REF["md_d00"], REF["md_d01"]
    w_in = float(REF["sg_omegas"][1])  # half the band top: the reference's sg0_* point
    with pytest.raises(NotImplementedError, match="eta = 0"):
        surface_gf(w_in, d00, d01, eta=0.0, method=method)
    w_out = float(REF["sg_omegas"][3])  # 1.3 x the band top
    g = surface_gf(w_out, d00, d01, eta=0.0, method=method)
    g_ref = o.sancho_rubio(w_out, d00, d01, None, 0.0)
    assert np.max(np.abs(g - g_ref)) <= 1e-10 * np.max(np.abs(g_ref))


def test_legacy_route_refuses_dropped_modes(gpu):
    """At eta = 1e-7 the propagating modes have 1 - |lambda| ~ 1e-7: a tolerance of 1e-3
    makes the reference drop them; the default 1e-10 of the dispatcher keeps all n."""
    w = float(REF["sg_omegas"][1])
    with pytest.raises(NotImplementedError, match="unit_circle_tol"):
        surface_gf(w, REF["md_d00"], REF["md_d01"], eta=1e-7, method="legacy_eigen_svd",
                   unit_circle_tol=1e-3)
    g = surface_gf(w, REF["md_d00"], REF["md_d01"], eta=1e-7, method="legacy_eigen_svd")
    g_ref = REF["sg_legacy_eigen_svd"][1]
    assert np.max(np.abs(g - g_ref)) <= 1e-9 * np.max(np.abs(g_ref))


def test_dos_matches_reference(gpu):
    lead = LeadBlocks(d00=REF["md_d00"], d01=REF["md_d01"])
    for w, d_ref, dn_ref in zip(REF["md_omegas"], REF["dos"], REF["dos_norm"]):
        assert abs(lead_surface_dos(w, lead, eta=1e-7) - d_ref) <= 1e-9 * max(abs(d_ref), 1e-3)
        assert abs(lead_surface_dos(w, lead, eta=1e-7, normalize_per_mode=True) - dn_ref) <= \
            1e-9 * max(abs(dn_ref), 1e-3)


def test_device_green_function_matches_reference(gpu):
    _, _, dev, lead = _md()
    g, sl, sr = device_green_function(float(REF["dg_omega"]), dev, lead, lead, eta=1e-6,
                                      device_to_lead_left=REF["md_vl"],
                                      device_to_lead_right=REF["md_vr"],
                                      contact_left_indices=[0, 2, 4],
                                      contact_right_indices=[25, 27, 29])
    assert np.max(np.abs(g - REF["dg_g"])) <= 1e-10 * np.max(np.abs(REF["dg_g"]))
    assert np.max(np.abs(sl - REF["dg_sigma_l"])) <= 1e-10 * np.max(np.abs(REF["dg_sigma_l"]))
    assert np.max(np.abs(sr - REF["dg_sigma_r"])) <= 1e-10 * np.max(np.abs(REF["dg_sigma_r"]))


def _sq():
    sq = SquareLatticeParams(mass=1.3, spring_x=0.9, spring_y=0.6)
    return (square_lattice_device(4, sq, defect_index=2, defect_mass=2.0), square_lattice_lead(sq),
            [(float(k),) for k in REF["sq_kpts"]])


def test_square_lattice_k_resolved_matches_reference(gpu):
    dev, lead, kpts = _sq()
    for wi, w in enumerate((0.4, 1.1, 1.6)):
        t = [transmission(w, dev, lead, lead, eta=1e-7, kpar=k) for k in kpts]
        _assert_vs_ref(t, REF["sq_t"][wi], [0.0] * len(t))


def test_adaptive_drivers_match_reference(gpu):
    dev, lead, kpts = _sq()
    ref = ADAPTIVE["sq_default"]
    t, info = transmission_kavg_adaptive(ref["omega"], dev, lead, lead, kpoints=kpts,
                                         eta_values=(1e-8, 1e-6), nonnegative_tolerance=1e-6,
                                         collect_rejected=True)
    assert abs(t - ref["t"]) <= T_TOL * max(abs(ref["t"]), 1e-3)
    assert_info_matches(info, ref["info"])
    ref = ADAPTIVE["sq_channel_limit"]
    t, info = transmission_kavg_adaptive(ref["omega"], dev, lead, lead, kpoints=kpts,
                                         eta_values=(1e-8, 1e-7), max_channel_factor=0.5,
                                         nonnegative_tolerance=1e-6, collect_rejected=True)
    assert abs(t - ref["t"]) <= T_TOL * max(abs(ref["t"]), 1e-3)
    assert_info_matches(info, ref["info"])
    ref = ADAPTIVE["sq_global_eta"]
    t, info = transmission_kavg_adaptive_global_eta(
        ref["omega"], dev, lead, lead, kpoints=kpts, eta_values=(1e-8, 1e-7),
        min_success_fraction=0.2, max_channel_factor=0.5, nonnegative_tolerance=1e-6,
        collect_rejected=True)
    assert abs(t - ref["t"]) <= T_TOL * max(abs(ref["t"]), 1e-3)
    assert_info_matches(info, ref["info"])


def test_ribbon_matches_reference(gpu):
    """The benchmark's 24-DOF x 32-layer ribbon: the GPU's T at the reference's points."""
    rp = ribbon_problem(seed=0, n_omega=4096)
    idx = REF["rb_index"]
    t = transmission_spectrum(rp.omegas[idx], rp.device, rp.lead, rp.lead, eta=rp.eta)
    ons, cps = rp.device.onsite_blocks, rp.device.coupling_blocks
    tup = (rp.lead.d00, rp.lead.d01, None)
    sp = [oracle_spread(rp.omegas[i], ons, cps, tup, tup, tr, eta=rp.eta)
          for i, tr in zip(idx, REF["rb_t"])]
    _assert_vs_ref(t, REF["rb_t"], sp)
