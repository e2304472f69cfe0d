"""negf_transmission_sweep_batched: the contact self-energies, the device sweep and the
trace in one launch (k_rgf_mm<NMAX, true>) against the four separate kernels
(negf_self_energy_batched x 2, negf_merge_status, negf_rgf_g1n_batched,
negf_transmission_trace_batched) on the same synthetic inputs, for every FP64 workgroup
class, contact subsets, leads larger and smaller than the device block, one layer, lead
failures, the fallback form (vector-ALU products) and a lead larger than the device
block (the launch takes the lead's class).
The physics of the composition is pinned against the oracle by the transmission parity
tests, which run through this entry."""

import numpy as np
import pytest
import torch

from negfpy_amd import _lib

pytestmark = pytest.mark.gpu

C128, I32, F64 = torch.complex128, torch.int32, torch.float64


@pytest.fixture
def lib(gpu):
    lib = _lib.load()
    prev_wave = lib.negf_set_wave_kernels(0)
    prev_mm = lib.negf_set_matrix_core_products(1)
    yield lib
    lib.negf_set_matrix_core_products(prev_mm)
    lib.negf_set_wave_kernels(prev_wave)


def _cplx(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _case(seed, n, layers, nll, ncl, nlr, ncr, n_inst=24, n_sets=2):
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda")
    herm = lambda a: 0.5 * (a + np.conj(np.swapaxes(a, -1, -2)))  # noqa: E731
    c = {
        "n": n, "layers": layers, "n_inst": n_inst, "nll": nll, "ncl": ncl, "nlr": nlr, "ncr": ncr,
        "onsite": herm(_cplx(rng, n_sets, layers, n, n)) / np.sqrt(n),
        "coupling": 0.5 * _cplx(rng, n_sets, max(layers - 1, 1), n, n) / np.sqrt(n),
        "v_l": 0.5 * _cplx(rng, n_sets, ncl, nll) / np.sqrt(nll),
        "v_r": 0.5 * _cplx(rng, n_sets, ncr, nlr) / np.sqrt(nlr),
        "g_l": _cplx(rng, n_inst, nll, nll) / np.sqrt(nll),
        "g_r": _cplx(rng, n_inst, nlr, nlr) / np.sqrt(nlr),
        "idx_l": np.sort(rng.choice(n, ncl, replace=False)),
        "idx_r": np.sort(rng.choice(n, ncr, replace=False)),
        "set_of": rng.integers(0, n_sets, n_inst),
        "omega": rng.uniform(1.2, 2.0, n_inst),
        "eta": np.full(n_inst, 1e-3),
        "st_l": np.zeros(n_inst, np.int32), "st_r": np.zeros(n_inst, np.int32),
    }
    c["st_l"][3], c["st_r"][3] = 1, 2       # both leads failed: the left status wins
    c["st_r"][5] = 3                        # right lead only
    t = {}
    for k, v in c.items():
        if isinstance(v, np.ndarray):
            dt = C128 if np.iscomplexobj(v) else (I32 if v.dtype.kind in "iu" else F64)
            t[k] = torch.as_tensor(v, dtype=dt, device=dev).contiguous()
        else:
            t[k] = v
    return t


def _outputs(c):
    n, ni = c["n"], c["n_inst"]
    z = lambda *s: torch.zeros(s, dtype=C128, device="cuda")  # noqa: E731
    return {"sig_l": z(ni, n, n), "sig_r": z(ni, n, n), "g1n": z(ni, n, n),
            "status": torch.full((ni,), -9, dtype=I32, device="cuda"),
            "t": torch.zeros(ni, dtype=F64, device="cuda"),
            "st_l": c["st_l"].clone(), "st_r": c["st_r"].clone()}


def _fused(lib, c):
    o = _outputs(c)
    _lib.check(lib.negf_transmission_sweep_batched(
        _lib.stream(), c["n"], c["layers"], c["n_inst"], c["onsite"].data_ptr(),
        c["coupling"].data_ptr(), c["set_of"].data_ptr(), c["omega"].data_ptr(),
        c["eta"].data_ptr(), c["nll"], c["ncl"], c["v_l"].data_ptr(), c["g_l"].data_ptr(),
        c["idx_l"].data_ptr(), o["st_l"].data_ptr(), c["nlr"], c["ncr"], c["v_r"].data_ptr(),
        c["g_r"].data_ptr(), c["idx_r"].data_ptr(), o["st_r"].data_ptr(), o["sig_l"].data_ptr(),
        o["sig_r"].data_ptr(), o["g1n"].data_ptr(), o["status"].data_ptr(), o["t"].data_ptr()))
    torch.cuda.synchronize()
    return o


def _separate(lib, c):
    o = _outputs(c)
    st, n, ni = _lib.stream(), c["n"], c["n_inst"]
    for side in "lr":
        _lib.check(lib.negf_self_energy_batched(
            st, c["nl" + side], c["nc" + side], n, ni, c["v_" + side].data_ptr(),
            c["set_of"].data_ptr(), c["g_" + side].data_ptr(), c["idx_" + side].data_ptr(),
            o["st_" + side].data_ptr(), o["sig_" + side].data_ptr()))
    _lib.check(lib.negf_merge_status(st, ni, o["st_l"].data_ptr(), o["st_r"].data_ptr()))
    _lib.check(lib.negf_rgf_g1n_batched(
        st, _lib.FP64, n, c["layers"], ni, c["onsite"].data_ptr(), c["coupling"].data_ptr(),
        c["set_of"].data_ptr(), c["omega"].data_ptr(), c["eta"].data_ptr(),
        o["sig_l"].data_ptr(), o["sig_r"].data_ptr(), o["st_l"].data_ptr(), o["g1n"].data_ptr(),
        o["status"].data_ptr()))
    _lib.check(lib.negf_transmission_trace_batched(
        st, n, ni, o["sig_l"].data_ptr(), o["sig_r"].data_ptr(), o["g1n"].data_ptr(),
        o["status"].data_ptr(), o["t"].data_ptr()))
    torch.cuda.synchronize()
    return o


def _rel(a, b, ok):
    a, b = a[ok].cpu().numpy(), b[ok].cpu().numpy()
    num = np.linalg.norm((a - b).reshape(a.shape[0], -1), axis=1)
    return num / np.maximum(np.linalg.norm(b.reshape(b.shape[0], -1), axis=1), 1e-300)


def _compare(f, s, tol=1e-11):
    assert f["status"].tolist() == s["status"].tolist()
    st = f["status"].cpu().numpy()
    assert st[3] == 1 and st[5] == 3
    ok = torch.as_tensor(st == 0, device="cuda")
    assert int(ok.sum()) == st.size - 2
    tf, ts = f["t"].cpu().numpy(), s["t"].cpu().numpy()
    assert np.all(np.isnan(tf[st != 0])) and np.all(np.isnan(ts[st != 0]))
    for k in ("sig_l", "sig_r", "g1n"):
        r = _rel(f[k], s[k], ok)
        assert r.max() < tol, (k, r.max())
    okn = st == 0
    scale = np.maximum(np.abs(ts[okn]), 1e-3 * np.abs(ts[okn]).max())
    assert np.max(np.abs(tf[okn] - ts[okn]) / scale) < 100 * tol


CASES = [  # n, layers, lead l, contact l, lead r, contact r
    (6, 5, 6, 6, 6, 6),        # class 8
    (13, 3, 9, 4, 13, 13),     # class 16, a 4-DOF contact of a 9-DOF lead
    (24, 32, 24, 24, 24, 24),  # class 24, the ribbon's shape
    (24, 4, 17, 24, 12, 5),    # class 24, leads smaller than the block, rectangular V
    (31, 1, 20, 7, 31, 31),    # class 32, one layer (G_1N = g_1)
    (45, 3, 45, 45, 40, 40),   # class 48
]


@pytest.mark.parametrize("case", CASES)
def test_fused_equals_separate_kernels(lib, case):
    c = _case(sum(case), *case)
    f, s = _fused(lib, c), _separate(lib, c)
    print(f"{case}: T {f['t'][:3].tolist()}", flush=True)
    _compare(f, s)


def test_fallback_forms(lib):
    """Vector-ALU products run the four kernels, whose results equal the fused form's to
    rounding; a lead larger than the device block's class runs fused in the lead's class
    and equals the four kernels."""
    c = _case(7, 24, 6, 24, 24, 24, 24)
    fused = _fused(lib, c)
    assert lib.negf_set_matrix_core_products(0) == 1
    _compare(_fused(lib, c), fused, tol=1e-10)
    assert lib.negf_set_matrix_core_products(1) == 0
    big = _case(8, 24, 4, 30, 24, 24, 24)   # 30-DOF lead: class 32, the block's is 24
    _compare(_fused(lib, big), _separate(lib, big))


def test_bad_arguments(lib):
    c = _case(9, 8, 2, 8, 8, 8, 8)
    rc = lib.negf_transmission_sweep_batched(
        _lib.stream(), 8, 2, c["n_inst"], c["onsite"].data_ptr(), c["coupling"].data_ptr(),
        c["set_of"].data_ptr(), c["omega"].data_ptr(), c["eta"].data_ptr(), 8, 9,
        c["v_l"].data_ptr(), c["g_l"].data_ptr(), c["idx_l"].data_ptr(), c["st_l"].data_ptr(), 8,
        8, c["v_r"].data_ptr(), c["g_r"].data_ptr(), c["idx_r"].data_ptr(), c["st_r"].data_ptr(),
        None, None, None, None, None)
    assert rc == 1 and b"contact" in lib.negf_last_error()
