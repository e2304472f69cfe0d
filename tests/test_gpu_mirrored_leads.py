"""Mirrored lead pairs: the right lead is the left one with d01 and d10 exchanged, as
the interface contacts of a bulk force-constant model are (ifc_bulk.py:690-742).
The reference decimates each lead on its own (surface_gf.py:18-81, called per lead
from negf.py:72-97); here one decimation of the left lead also sums b g a and gives
the right lead's surface (share_leads = NEGF_LEADS_MIRRORED).  Checked against two
independent decimations of the same problem and against the oracle, in FP64 and
mixed precision, for every block class of the mirrored kernel (8, 16, 24, 32, 48),
with decimation failures and the eigenmode fallback.
"""

import dataclasses

import numpy as np
import pytest
import torch

from negfpy_amd.core import _engine
from negfpy_amd.models.multidof import multidof_lead_blocks
from negfpy_amd.workloads import omega_max, ribbon_params, sheet_problem
from parity_util import assert_parity, oracle_t

pytestmark = pytest.mark.gpu


def _mirror_problem(seed, n, n_layers):
    """Right lead (d00, A, A^H) of a ribbon cell (random pair springs: A != P A^H P^T
    for any permutation P), left lead (d00, A^H, A), couplings V_R = A, V_L = A^H,
    and a device of n_layers cells with perturbed onsite blocks."""
    lb = multidof_lead_blocks(ribbon_params(seed, n // 3))
    d00, a = np.asarray(lb.d00), np.asarray(lb.d01)
    ah = a.conj().T
    rng = np.random.default_rng(seed + 3)
    ons = [d00 + np.diag(rng.uniform(-0.3, 0.3, n)) for _ in range(n_layers)]
    cps = [a] * (n_layers - 1)
    left, right = (d00, ah, a), (d00, a, ah)
    sets = _engine.TransportSets.from_arrays(
        np.stack(left)[None], np.stack(right)[None], ah[None], a[None], np.arange(n),
        np.arange(n), np.stack(ons)[None], np.stack(cps)[None], False)
    wm = omega_max(lb)
    return sets, wm, (left, right, ah, a, ons, cps)


def _run(sets, w, *, precision="fp64", maxiter=200, route=0, eta=1e-6):
    m = w.size
    eta = np.full(m, 1e-6) if eta is None else np.array(np.broadcast_to(np.asarray(eta, float), (m,)))
    res = _engine.TransmissionSolver(sets, precision=precision, maxiter=maxiter,
                                     surface_route=route).run(
        np.zeros(m, dtype=np.int32), w, eta, eta)
    return {k: getattr(res, k).cpu().numpy().copy()
            for k in ("t", "status", "iters_l", "iters_r", "dos_l", "dos_r")}


@pytest.mark.parametrize("n", [6, 12, 24, 30, 48])
def test_mirrored_equals_independent_and_oracle(gpu, n):
    sets, wm, (left, right, vl, vr, ons, cps) = _mirror_problem(40 + n, n, 4)
    assert sets.mirror_leads and not sets.share_leads
    ind = dataclasses.replace(sets, mirror_leads=False)
    w = np.linspace(0.04, 1.08, 23) * wm
    mir, two = _run(sets, w), _run(ind, w)
    assert np.all(mir["status"] == 0) and np.all(two["status"] == 0)
    np.testing.assert_array_equal(mir["iters_l"], mir["iters_r"])
    assert np.max(np.abs(mir["iters_r"] - two["iters_r"])) <= 1
    floor = np.maximum(np.abs(two["t"]), 1e-3)
    rel = np.abs(mir["t"] - two["t"]) / floor
    print(f"n={n}: mirrored vs two decimations: max {rel.max():.2e} median {np.median(rel):.2e}",
          flush=True)
    assert np.median(rel) < 1e-12
    for k in ("dos_l", "dos_r"):
        np.testing.assert_allclose(mir[k], two[k], rtol=1e-8, atol=1e-10 * np.abs(two[k]).max())
    sel = np.arange(0, w.size, 4)
    refs = [oracle_t(w[i], ons, cps, left, right, eta=1e-6, v_left=vl, v_right=vr) for i in sel]
    assert_parity(mir["t"][sel], refs)
    assert_parity(two["t"][sel], refs)


@pytest.mark.parametrize("n", [12, 24, 48])
@pytest.mark.parametrize("eta_rel", [None, 1e-2])
def test_mirrored_mixed_precision(gpu, n, eta_rel):
    """FP32 decimation of both surfaces at once, each completed by the FP64 Newton step
    against its own lead and certified; what either certificate rejects is decimated
    again in FP64.  At eta = 1e-6 most instances go to the FP64 redo; at eta = 1e-2
    omega the FP32 pass is certified (n <= 32; 48-DOF blocks always take the FP64
    redo) -- both must meet the 1e-6 bar against FP64."""
    sets, wm, _ = _mirror_problem(70 + n, n, 4)
    ind = dataclasses.replace(sets, mirror_leads=False)
    w = np.linspace(0.04, 1.08, 31) * wm
    eta = None if eta_rel is None else eta_rel * w
    ref = _run(sets, w, eta=eta)
    mir, two = (_run(sets, w, precision="mixed", eta=eta),
                _run(ind, w, precision="mixed", eta=eta))
    np.testing.assert_array_equal(mir["status"], ref["status"])
    np.testing.assert_array_equal(two["status"], ref["status"])
    for got in (mir, two):
        err = np.abs(got["t"] - ref["t"]) / np.maximum(np.abs(ref["t"]), 1e-3)
        assert np.all(np.isfinite(got["t"])) and np.max(err) <= 1e-6, np.max(err)


def test_mirrored_failures_and_eigen_fallback(gpu):
    """Iteration limit 3: both surfaces fail together (one decimation) with the status
    of two failing decimations; with the eigenmode route every instance takes the
    fallback for each lead, and T equals the two-decimation run's bit for bit."""
    sets, wm, _ = _mirror_problem(91, 12, 3)
    ind = dataclasses.replace(sets, mirror_leads=False)
    w = np.linspace(0.1, 0.9, 9) * wm
    for prec in ("fp64", "mixed"):
        mir, two = _run(sets, w, precision=prec, maxiter=3), _run(ind, w, precision=prec,
                                                                  maxiter=3)
        np.testing.assert_array_equal(mir["status"], two["status"])
        assert np.all(mir["status"] != 0)
    mir, two = _run(sets, w, maxiter=3, route=1), _run(ind, w, maxiter=3, route=1)
    assert np.all(mir["status"] == 0)
    for k in ("t", "status", "dos_l", "dos_r"):
        np.testing.assert_array_equal(mir[k], two[k], err_msg=k)


def test_sheet_kspace_sets_are_mirrored(gpu):
    """The GPU k-space builder's interface-contact leads are exact mirrors, so the sheet
    takes one decimation per (omega, k); T as with two decimations."""
    from negfpy_amd.core.negf import kspace_sets

    p = sheet_problem(n_layers=6, n_k=8, n_omega=16)
    sets = kspace_sets(p.device, p.lead_left, p.lead_right, p.kpoints,
                       device_to_lead_left=p.v_left, device_to_lead_right=p.v_right)
    assert sets is not None and sets.mirror_leads
    nk = len(p.kpoints)
    k_of = torch.arange(nk, dtype=torch.int32, device="cuda").repeat(p.omegas.size)
    om = torch.tensor(np.repeat(p.omegas, nk), device="cuda")
    eta = torch.full_like(om, p.eta)
    a = _engine.TransmissionSolver(sets).run(k_of, om, eta)
    b = _engine.TransmissionSolver(dataclasses.replace(sets, mirror_leads=False)).run(k_of, om,
                                                                                     eta)
    assert torch.equal(a.status, b.status) and int(a.status.abs().sum()) == 0
    rel = (a.t - b.t).abs() / b.t.abs().clamp_min(1e-3)
    assert float(rel.median()) < 1e-12 and float(rel.max()) < 1e-8
