"""Matrix-core products (negfpy_amd/csrc/mm.hpp: v_mfma_f64_4x4x4f64) in the FP64
workgroup decimation and sweep, against the oracle and against the vector-ALU form of
the same kernels (negf_set_matrix_core_products 0 / 1), for every workgroup class
(8, 16, 24, 32, 48 DOF) and block sizes that are not multiples of 4 (padded K steps).
The one-wavefront family is switched off so that every size runs the workgroup
kernels."""

import dataclasses

import numpy as np
import pytest

from negfpy_amd import _lib
from negfpy_amd.core import _engine, transmission_spectrum
from negfpy_amd.core.surface_gf import pack_lead
from negfpy_amd.core.types import LeadBlocks
from negfpy_amd.workloads import omega_max, random_lead, ribbon_problem
from oracle import negf_oracle as o
from parity_util import assert_parity, oracle_t

pytestmark = pytest.mark.gpu


@pytest.fixture
def mm_mode(gpu):
    lib = _lib.load()
    prev_wave = lib.negf_set_wave_kernels(0)
    prev_mm = lib.negf_set_matrix_core_products(1)

    def set_mode(m):
        assert lib.negf_set_matrix_core_products(m) in (0, 1)

    yield set_mode
    lib.negf_set_matrix_core_products(prev_mm)
    lib.negf_set_wave_kernels(prev_wave)


def _surface(d00, d01, omegas, eta):
    res = _engine.surface_gf_batch(pack_lead(d00, d01)[None], np.zeros(omegas.size, np.int32),
                                   omegas, np.full(omegas.size, eta))
    return (res.g.cpu().numpy(), res.iters.cpu().numpy(), res.dos.cpu().numpy(),
            res.status.cpu().numpy())


@pytest.mark.parametrize("n", [1, 5, 8, 11, 16, 21, 24, 30, 32, 41, 48])
def test_mm_decimation_matches_oracle(mm_mode, n):
    d00, d01 = random_lead(131 + n, n)
    wmax = omega_max(LeadBlocks(d00=d00, d01=d01))
    omegas = np.linspace(0.03, 1.15, 10) * wmax
    g, its, dos, st = _surface(d00, d01, omegas, 1e-6)
    assert st.tolist() == [0] * omegas.size
    rng = np.random.default_rng(9)
    for i, w in enumerate(omegas):
        gr, it = o.sancho_rubio(w, d00, d01, None, 1e-6, return_iters=True)
        err = np.linalg.norm(g[i] - gr) / np.linalg.norm(gr)
        # the oracle's own change under ~1-ulp perturbations of the lead blocks
        spread = max(np.linalg.norm(o.sancho_rubio(
            w, d00 * (1 + 2.2e-16 * rng.standard_normal(d00.shape)),
            d01 * (1 + 2.2e-16 * rng.standard_normal(d01.shape)), None, 1e-6) - gr)
            for _ in range(3)) / np.linalg.norm(gr)
        assert err < max(1e-10, 10 * spread), (n, w, err, spread)
        assert abs(int(its[i]) - it) <= 1
        ref_dos = o.surface_dos(w, gr)
        assert abs(dos[i] - ref_dos) <= max(1e-10, 10 * spread) * max(abs(ref_dos), 1e-3)


@pytest.mark.parametrize("n", [6, 13, 24, 32, 48])
def test_mm_and_vector_decimation_agree(mm_mode, n):
    d00, d01 = random_lead(151 + n, n)
    wmax = omega_max(LeadBlocks(d00=d00, d01=d01))
    omegas = np.linspace(0.02, 1.1, 16) * wmax
    g1, it1, d1, s1 = _surface(d00, d01, omegas, 1e-3)
    mm_mode(0)
    g0, it0, d0, s0 = _surface(d00, d01, omegas, 1e-3)
    assert s1.tolist() == s0.tolist() == [0] * omegas.size
    assert np.max(np.abs(it1 - it0)) <= 1
    err = np.linalg.norm(g1 - g0, axis=(1, 2)) / np.linalg.norm(g0, axis=(1, 2))
    assert np.median(err) < 1e-12 and err.max() < 1e-9, (np.median(err), err.max())


@pytest.mark.parametrize("n_atoms,n_layers", [(1, 4), (2, 7), (3, 5), (5, 3), (8, 9), (10, 4),
                                              (16, 3)])
def test_mm_transmission_matches_oracle(mm_mode, n_atoms, n_layers):
    p = ribbon_problem(seed=21 + n_atoms, n_atoms=n_atoms, n_layers=n_layers, n_omega=16,
                       eta=1e-6)
    t = transmission_spectrum(p.omegas, p.device, p.lead, p.lead, eta=p.eta)
    lead = (p.lead.d00, p.lead.d01, None)
    refs = [oracle_t(w, p.device.onsite_blocks, p.device.coupling_blocks, lead, lead, eta=p.eta)
            for w in p.omegas]
    assert_parity(t, refs)


def test_mm_single_layer_and_two_leads(mm_mode):
    """One device layer (G_1N = g_1) and different left and right leads."""
    pl = ribbon_problem(seed=61, n_atoms=4, n_layers=1, n_omega=12, eta=1e-6)
    pr = ribbon_problem(seed=62, n_atoms=4, n_layers=1, n_omega=12, eta=1e-6)
    t = transmission_spectrum(pl.omegas, pl.device, pl.lead, pr.lead, eta=1e-6)
    ll = (pl.lead.d00, pl.lead.d01, None)
    lr = (pr.lead.d00, pr.lead.d01, None)
    refs = [oracle_t(w, pl.device.onsite_blocks, pl.device.coupling_blocks, ll, lr, eta=1e-6)
            for w in pl.omegas]
    assert_parity(t, refs)


def test_mm_ribbon_equals_vector_path(mm_mode):
    """The benchmark ribbon (24 DOF, 32 layers): matrix-core and vector products give the
    same T to rounding at 128 frequencies."""
    p = ribbon_problem(seed=0, n_omega=128, eta=1e-6)
    t1 = transmission_spectrum(p.omegas, p.device, p.lead, p.lead, eta=p.eta)
    mm_mode(0)
    t0 = transmission_spectrum(p.omegas, p.device, p.lead, p.lead, eta=p.eta)
    rel = np.abs(t1 - t0) / np.maximum(np.abs(t0), 1e-3)
    print(f"ribbon mm vs vector: max {rel.max():.2e} median {np.median(rel):.2e}", flush=True)
    assert np.median(rel) < 1e-12 and rel.max() < 1e-7
