"""The one-wavefront MFMA kernel family (negfpy_amd/csrc/wpi.hpp) against the oracle
and against the workgroup family, forced on for every block size it covers
(negf_set_wave_kernels(2)): surface Green's functions, FP32 decimation, the device
sweep, and whole transmissions with the oracle's spread-aware FP64 bar."""

import numpy as np
import pytest
import torch

from negfpy_amd import _lib
from negfpy_amd.core import _engine
from negfpy_amd.core.surface_gf import pack_lead
from negfpy_amd.core.types import LeadBlocks
from negfpy_amd.workloads import omega_max, random_lead
from oracle import negf_oracle as o
from parity_util import assert_parity, oracle_t

pytestmark = pytest.mark.gpu


@pytest.fixture
def wave_mode(gpu):
    lib = _lib.load()
    prev = lib.negf_set_wave_kernels(2)

    def set_mode(m):
        lib.negf_set_wave_kernels(m)

    yield set_mode
    lib.negf_set_wave_kernels(prev)


@pytest.mark.parametrize("n", [1, 3, 8, 12, 16, 20, 24, 30, 32])
def test_wave_decimation_matches_oracle(wave_mode, n):
    d00, d01 = random_lead(31 + n, n)
    wmax = omega_max(LeadBlocks(d00=d00, d01=d01))
    omegas = np.linspace(0.03, 1.15, 10) * wmax
    res = _engine.surface_gf_batch(pack_lead(d00, d01)[None], np.zeros(omegas.size, np.int32),
                                   omegas, np.full(omegas.size, 1e-6))
    assert res.status.cpu().numpy().tolist() == [0] * omegas.size
    g = res.g.cpu().numpy()
    its = res.iters.cpu().numpy()
    dos = res.dos.cpu().numpy()
    rng = np.random.default_rng(5)
    for i, w in enumerate(omegas):
        gr, it = o.sancho_rubio(w, d00, d01, None, 1e-6, return_iters=True)
        err = np.linalg.norm(g[i] - gr) / np.linalg.norm(gr)
        # the oracle's own change under ~1-ulp perturbations of the lead blocks: near
        # band features decimation is conditioned like 1/eta
        spread = max(np.linalg.norm(o.sancho_rubio(
            w, d00 * (1 + 2.2e-16 * rng.standard_normal(d00.shape)),
            d01 * (1 + 2.2e-16 * rng.standard_normal(d01.shape)), None, 1e-6) - gr)
            for _ in range(3)) / np.linalg.norm(gr)
        assert err < max(1e-10, 10 * spread), (n, w, err, spread)
        assert abs(int(its[i]) - it) <= 1
        ref_dos = o.surface_dos(w, gr)
        assert abs(dos[i] - ref_dos) <= max(1e-10, 10 * spread) * max(abs(ref_dos), 1e-3)


@pytest.mark.parametrize("n", [6, 16, 24])
def test_wave_fp32_decimation_surface_term(wave_mode, n):
    """The FP32 pass leaves e_s (complex128) for the FP64 completion: close to the FP64
    decimation's e_s = z - g^-1 to FP32 accuracy away from band features.  An instance
    whose slowest mode the pass cannot resolve stops early (status 1, left to the FP64
    redo; at eta = 1e-2 omega only the lowest frequencies can); the complete mixed
    decimation then gives status 0 everywhere."""
    d00, d01 = random_lead(77 + n, n)
    wmax = omega_max(LeadBlocks(d00=d00, d01=d01))
    omegas = np.linspace(0.05, 1.1, 8) * wmax
    eta = 1e-2 * omegas  # well-conditioned points
    lib = _lib.load()
    lead = _engine.to_dev_c128(pack_lead(d00, d01)[None])
    m = omegas.size
    es = torch.empty((m, n, n), dtype=torch.complex128, device="cuda")
    iters = torch.zeros(m, dtype=torch.int32, device="cuda")
    status = torch.zeros(m, dtype=torch.int32, device="cuda")
    so = torch.zeros(m, dtype=torch.int32, device="cuda")
    om = torch.tensor(omegas, device="cuda")
    et = torch.tensor(eta, device="cuda")
    _lib.check(lib.negf_sancho_rubio_batched(
        _lib.stream(), _lib.MIXED, n, m, lead.data_ptr(), so.data_ptr(), om.data_ptr(),
        et.data_ptr(), 1e-12, 200, -1, es.data_ptr(), None, iters.data_ptr(), status.data_ptr()))
    es = es.cpu().numpy()
    st = status.cpu().numpy()
    assert set(st.tolist()) <= {0, 1} and int((st == 0).sum()) >= m - 2, st
    assert np.all(st[2:] == 0), st  # the stall test stops only the lowest frequencies
    st2 = torch.zeros(m, dtype=torch.int32, device="cuda")
    g_full = torch.empty((m, n, n), dtype=torch.complex128, device="cuda")
    _lib.check(lib.negf_sancho_rubio_batched(
        _lib.stream(), _lib.MIXED, n, m, lead.data_ptr(), so.data_ptr(), om.data_ptr(),
        et.data_ptr(), 1e-12, 200, 2, g_full.data_ptr(), None, iters.data_ptr(),
        st2.data_ptr()))
    assert st2.cpu().numpy().tolist() == [0] * m
    for i, w in enumerate(omegas):
        if st[i] != 0:
            continue
        gr = o.sancho_rubio(w, d00, d01, None, eta[i])
        z = (w + 1j * eta[i]) ** 2
        es_ref = z * np.eye(n) - np.linalg.inv(gr)
        assert np.linalg.norm(es[i] - es_ref) <= 1e-4 * np.linalg.norm(es_ref), (n, w)


@pytest.mark.parametrize("n,n_layers", [(4, 6), (12, 5), (24, 4), (32, 3)])
def test_wave_transmission_matches_oracle_and_classic(wave_mode, n, n_layers):
    d00, d01 = random_lead(500 + n, n)
    rng = np.random.default_rng(n)
    ons = [d00 + np.diag(rng.uniform(-0.05, 0.05, n)) for _ in range(n_layers)]
    cps = [d01 * (1.0 + 0.1 * rng.standard_normal()) for _ in range(n_layers - 1)]
    lb = np.stack([d00, d01, d01.conj().T])[None]
    sets = _engine.TransportSets.from_arrays(lb, lb, d01.conj().T[None], d01[None], np.arange(n),
                                             np.arange(n), np.stack(ons)[None],
                                             np.stack(cps)[None], True)
    wmax = omega_max(LeadBlocks(d00=d00, d01=d01))
    omegas = np.linspace(0.04, 1.05, 12) * wmax
    m = omegas.size
    args = (torch.zeros(m, dtype=torch.int32, device="cuda"),
            torch.tensor(omegas, device="cuda"), torch.full((m,), 1e-6, dtype=torch.float64, device="cuda"))
    t_wave = _engine.TransmissionSolver(sets).run(*args).t.cpu().numpy()
    wave_mode(0)
    t_classic = _engine.TransmissionSolver(sets).run(*args).t.cpu().numpy()
    tup = (d00, d01, None)
    refs = [oracle_t(w, ons, cps, tup, tup, eta=1e-6) for w in omegas]
    assert_parity(t_wave, refs)
    assert_parity(t_classic, refs)
