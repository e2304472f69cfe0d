"""negf_transmission_batched_noalloc (workspace in the library's file-scope device
scratch, no allocation) against negf_transmission_batched (caller workspace) through
the C ABI: identical T, status, decimation steps and surface DOS on two calls in a row,
the second one smaller, in FP64 and mixed precision."""

import ctypes

import numpy as np
import pytest
import torch

from negfpy_amd import _lib
from negfpy_amd.core import _engine
from negfpy_amd.workloads import ribbon_problem
from negfpy_amd.core.negf import resolve_k

pytestmark = pytest.mark.gpu


def _solver(precision):
    p = ribbon_problem(n_omega=4096)
    kp = resolve_k(p.device, p.lead, p.lead, None, eta=p.eta, eta_device=None,
                   device_to_lead_left=None, device_to_lead_right=None, contact_left_indices=None,
                   contact_right_indices=None, omega_scale=None, max_block=64)
    sets = _engine.TransportSets.from_arrays(kp.lead_l[None], kp.lead_r[None], kp.v_l[None],
                                             kp.v_r[None], kp.pos_l, kp.pos_r, kp.onsite[None],
                                             kp.coupling[None], True)
    return _engine.TransmissionSolver(sets, precision=precision), p


def _call(solver, omegas, eta, noalloc):
    lib = _lib.load()
    m = omegas.size
    so = torch.zeros(m, dtype=torch.int32, device="cuda")
    om = torch.tensor(omegas, device="cuda")
    et = torch.full((m,), eta, dtype=torch.float64, device="cuda")
    out = solver.alloc_result(m)
    prob = solver._problem(m, so, om, et, et, out)
    if noalloc:
        _lib.check(lib.negf_transmission_batched_noalloc(_lib.stream(), ctypes.byref(prob)))
    else:
        need = int(lib.negf_transmission_workspace_bytes(ctypes.byref(prob)))
        ws = torch.empty(need, dtype=torch.uint8, device="cuda")
        _lib.check(lib.negf_transmission_batched(_lib.stream(), ctypes.byref(prob), ws.data_ptr(),
                                                 need))
    torch.cuda.synchronize()
    return {k: getattr(out, k).cpu().numpy().copy() for k in ("t", "status", "iters_l", "dos_l")}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_equals_workspace_entry(gpu, precision):
    solver, p = _solver(precision)
    for sel in (np.arange(0, 4096, 3), np.arange(5, 4096, 17)):  # 1366, then 241 instances
        w = p.omegas[sel]
        ref = _call(solver, w, p.eta, noalloc=False)
        got = _call(solver, w, p.eta, noalloc=True)
        for k in ref:
            np.testing.assert_array_equal(got[k], ref[k], err_msg=f"{precision} {k} m={w.size}")
        assert np.all(got["status"] == 0)


def test_noalloc_rejects_oversized_batches(gpu):
    solver, p = _solver("fp64")
    lib = _lib.load()
    m = 4096  # 4096 x 4 x 24^2 x 16 B > 64 MiB
    so = torch.zeros(m, dtype=torch.int32, device="cuda")
    om = torch.tensor(p.omegas, device="cuda")
    et = torch.full((m,), p.eta, dtype=torch.float64, device="cuda")
    out = solver.alloc_result(m)
    prob = solver._problem(m, so, om, et, et, out)
    assert lib.negf_transmission_batched_noalloc(_lib.stream(), ctypes.byref(prob)) == 1
    assert b"scratch" in lib.negf_last_error()
