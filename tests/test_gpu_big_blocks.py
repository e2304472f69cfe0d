"""Principal layers of 49..96 DOF: the pooled block-size class 96 (FP64).

Surface GF, surface DOS and transmission at n = 60, 72, 90, 96 against the
oracle (the same conditioning-aware bar as every FP64 parity test), explicit
device-to-lead couplings with two different leads, a batch larger than the
pool's slot count (every workgroup then walks several instances), a 40-DOF
problem forced into class 96, which must give the class-48 result to rounding, and every
block of the device Green's function against the oracle's dense inverse (the
adjoint at these sizes is in test_gpu_adjoint.py).
"""

import numpy as np
import pytest

from negfpy_amd import _lib
from negfpy_amd.core import (
    LeadBlocks,
    device_green_function,
    lead_surface_dos,
    surface_gf,
    transmission,
    transmission_spectrum,
)
from negfpy_amd.core.autotune import block_class
from negfpy_amd.core.types import Device1D
from negfpy_amd.workloads import omega_max, random_lead
from oracle import negf_oracle as o
from parity_util import assert_parity, oracle_t

pytestmark = pytest.mark.gpu


def _problem(seed, n, n_layers):
    d00, d01 = random_lead(seed, n)
    rng = np.random.default_rng(seed + 7)
    ons = [d00 + np.diag(rng.uniform(-0.05, 0.05, n)) for _ in range(n_layers)]
    cps = [d01 * (1.0 + 0.1 * rng.standard_normal()) for _ in range(n_layers - 1)]
    lead = LeadBlocks(d00=d00, d01=d01)
    return lead, ons, cps, omega_max(lead)


def test_classes(gpu):
    assert block_class(60) == 96 and block_class(96) == 96 and block_class(48) == 48
    assert _lib.load().negf_max_block(_lib.FP64) == 96


@pytest.mark.parametrize("n", [60, 72, 90, 96])
def test_surface_gf_and_dos(gpu, n):
    lead, _, _, wm = _problem(100 + n, n, 1)
    for f in (0.2, 0.55, 0.9):
        w = f * wm
        g = surface_gf(w, lead.d00, lead.d01, eta=1e-6)
        g_or = o.sancho_rubio(w, lead.d00, lead.d01, eta=1e-6)
        scale = np.max(np.abs(g_or))
        print(f"n={n} w={f:.2f}wmax |dg|/|g| = {np.max(np.abs(g - g_or)) / scale:.2e}", flush=True)
        assert np.max(np.abs(g - g_or)) <= 1e-9 * scale
        dos = lead_surface_dos(w, lead, eta=1e-6)
        dos_or = o.surface_dos(w, g_or)
        assert abs(dos - dos_or) <= 1e-9 * max(abs(dos_or), 1e-12)


@pytest.mark.parametrize("n,n_layers", [(60, 1), (72, 2), (90, 5), (96, 3)])
def test_transmission_matches_oracle(gpu, n, n_layers):
    lead, ons, cps, wm = _problem(200 + n, n, n_layers)
    dev = Device1D(onsite_blocks=ons, coupling_blocks=cps)
    for f in (0.15, 0.5, 0.85):
        w = f * wm
        t = transmission(w, dev, lead, lead, eta=1e-6)
        ref = oracle_t(w, ons, cps, (lead.d00, lead.d01, None), (lead.d00, lead.d01, None), eta=1e-6)
        print(f"n={n} L={n_layers} w={f:.2f}wmax T={t:.12e} oracle={ref[0]:.12e}", flush=True)
        assert_parity([t], [ref])


def test_two_leads_and_explicit_couplings(gpu):
    n = 72
    lead, ons, cps, wm = _problem(300, n, 3)
    d00r, d01r = random_lead(301, n)
    lead_r = LeadBlocks(d00=d00r, d01=d01r)
    rng = np.random.default_rng(5)
    vl = lead.d01.conj().T * (1.0 + 0.05 * rng.standard_normal((n, n)))
    vr = d01r * (1.0 + 0.05 * rng.standard_normal((n, n)))
    dev = Device1D(onsite_blocks=ons, coupling_blocks=cps)
    wr = min(wm, omega_max(lead_r))
    for f in (0.3, 0.7):
        w = f * wr
        t = transmission(w, dev, lead, lead_r, eta=1e-6, device_to_lead_left=vl,
                         device_to_lead_right=vr)
        ref = oracle_t(w, ons, cps, (lead.d00, lead.d01, None), (d00r, d01r, None), eta=1e-6,
                       v_left=vl, v_right=vr)
        assert_parity([t], [ref])


def test_batch_larger_than_pool(gpu):
    """600 omegas in one batch (more than the resident slots): each instance equals
    its own single-instance solve bit for bit."""
    n = 60
    lead, ons, cps, wm = _problem(400, n, 4)
    dev = Device1D(onsite_blocks=ons, coupling_blocks=cps)
    w = np.linspace(0.02, 0.98, 600) * wm
    t_all = transmission_spectrum(w, dev, lead, lead, eta=1e-6)
    for i in (0, 1, 257, 511, 599):
        assert t_all[i] == transmission(w[i], dev, lead, lead, eta=1e-6)


def test_class_96_equals_class_48(gpu):
    """A 40-DOF problem solved in the pooled class and in the LDS class 48: the same T
    to rounding (class 48 runs its products on the matrix cores, class 96 on the vector
    ALUs, so the two agree to the problem's conditioning, not bit for bit), and both
    within the oracle bar."""
    from negfpy_amd.core.autotune import reset_block_class

    n = 40
    lead, ons, cps, wm = _problem(500, n, 4)
    dev = Device1D(onsite_blocks=ons, coupling_blocks=cps)
    w = np.linspace(0.05, 0.95, 32) * wm
    t48 = transmission_spectrum(w, dev, lead, lead, eta=1e-6)
    lib = _lib.load()
    try:
        assert lib.negf_set_block_class(_lib.FP64, n, 96) == 0
        t96 = transmission_spectrum(w, dev, lead, lead, eta=1e-6)
    finally:
        reset_block_class(n)
    rel = np.abs(t96 - t48) / np.maximum(np.abs(t48), 1e-3)
    assert np.median(rel) < 1e-12 and rel.max() < 1e-7, (np.median(rel), rel.max())
    ld = (lead.d00, lead.d01, None)
    sel = np.argsort(rel)[-4:]  # the points where the two classes differ most
    refs = [oracle_t(w[i], ons, cps, ld, ld, eta=1e-6) for i in sel]
    assert_parity(t48[sel], refs)
    assert_parity(t96[sel], refs)


@pytest.mark.parametrize("n,n_layers", [(60, 3), (96, 2)])
def test_device_green_function(gpu, n, n_layers):
    """Every block of G for 60- and 96-DOF layers against the oracle's dense inverse."""
    lead, ons, cps, wm = _problem(600 + n, n, n_layers)
    dev = Device1D(onsite_blocks=ons, coupling_blocks=cps)
    w = 0.5 * wm
    g, sl, sr = device_green_function(w, dev, lead, lead, eta=1e-6)
    tup = (lead.d00, lead.d01, None)
    g_ref, sl_ref, sr_ref = o.device_green_dense(w, ons, cps, tup, tup, eta=1e-6)
    err = np.max(np.abs(g - g_ref)) / np.max(np.abs(g_ref))
    print(f"n={n} L={n_layers}: |G - G_ref| / |G| = {err:.2e}", flush=True)
    assert err <= 1e-9
    assert np.max(np.abs(sl - sl_ref)) <= 1e-10 * np.max(np.abs(sl_ref))
    assert np.max(np.abs(sr - sr_ref)) <= 1e-10 * np.max(np.abs(sr_ref))
