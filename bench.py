"""Benchmark: (omega, k) transmission solves per second on one MI355X.

Workload: the quasi-1D multi-DOF ribbon (8 atoms x 3 DOF = 24 DOF per principal
layer, 32 device layers, 4096 omega on (0, 1.1 w_max), eta = 1e-6, seeded
masses in [10, 30] amu and pair force constants in [-5, 5] eV/A^2 with the
acoustic sum rule).  One step is one full transmission solve of all 4096
(omega, k) instances: lead decimation, self-energies, recursive Green's function
sweep and FP64 trace, with all inputs resident in HBM.

Both paths are timed in the same run: FP64, and mixed (FP32 decimation, FP64
Newton completion with FP32 Stein corrections and a certificate, FP64 decimation
of the uncertified instances, FP64 sweep and trace).  The metric counts mixed
solves only at <= 1e-6 relative error in T; the headline `value` is the mixed
rate when its error against the oracle (every frequency, committed fixture
tests/golden/ribbon24_oracle.npz) meets that bar and the FP64 rate otherwise
(`headline_path`, `mixed_meets_1e-6`), with the other path's numbers beside it.
The CPU baseline is the FP64 NumPy/SciPy oracle (decimation + SuperLU, the
reference algorithm) on a bounded sample of the same instances, over a fork pool
of the host's cores (at most 16, the CPU share of one GPU on the benchmark hosts)
with one BLAS thread each.
"""

from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "(ω,k) transmission solves/s, mixed vs FP64 GPU path at ≤1e-6 rel err in T"
PEAK_TFLOPS = {"fp64": 78.6, "mixed": 157.3}  # MI355X dense matrix peaks (FP64, FP32)
# HBM bytes per launch of the FP64 kernels on this workload, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE scaled by a copy-kernel calibration (profiles/ribbon24_hbm_counters.md)
TRAFFIC: dict = {"sweep": 347.0e6, "sancho_rubio": 85.3e6}


def _cpu_worker(args):
    idx, = args
    from threadpoolctl import threadpool_limits

    from oracle import negf_oracle as o

    p = _CPU_PROBLEM
    lead = (p.lead.d00, p.lead.d01, None)
    out = []
    with threadpool_limits(limits=1):
        for i in idx:
            out.append(o.transmission_sparse(float(p.omegas[i]), p.device.onsite_blocks,
                                             p.device.coupling_blocks, lead, lead, eta=p.eta))
    return out


_CPU_PROBLEM = None


def host_description() -> dict:
    """CPU model, logical CPUs, and the NumPy / SciPy / BLAS builds of this host."""
    import platform

    import scipy

    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    blas = ""
    try:
        from threadpoolctl import threadpool_info

        blas = ", ".join(sorted({f"{i.get('internal_api')} {i.get('version')}"
                                 for i in threadpool_info() if i.get("user_api") == "blas"}))
    except Exception:  # threadpoolctl missing or unable to inspect: report nothing
        blas = ""
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "usable_cpus": affinity,
            "numpy": np.__version__, "scipy": scipy.__version__, "blas": blas}


def cpu_baseline(problem, n_sample: int):
    """Time the oracle on n_sample evenly spaced instances over a fork pool of the
    host's usable cores, capped at 16 (the CPU share of one GPU on the benchmark
    hosts)."""
    import multiprocessing as mp

    global _CPU_PROBLEM
    _CPU_PROBLEM = problem
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    workers = max(1, min(usable or 1, 16))
    sel = np.linspace(0, problem.omegas.size - 1, n_sample).astype(int)
    chunks = [sel[i::workers] for i in range(workers)]
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(workers) as pool:
        parts = pool.map(_cpu_worker, [(c,) for c in chunks])
    wall = time.perf_counter() - t0
    vals = np.empty(n_sample)
    for c, part in zip(chunks, parts):
        for i, v in zip(c, part):
            vals[np.searchsorted(sel, i)] = v
    return sel, vals, wall, workers


def uneven_batch(steps: int, warmup: int) -> dict:
    """The uneven batch (24-, 48- and 96-DOF ribbons, band-edge clustering) solved
    by the staged launches problem after problem and by one persistent launch."""
    import torch

    from negfpy_amd.core import _engine
    from negfpy_amd.workloads import uneven_problems

    jobs = []
    for p in uneven_problems():
        lead, dev = p.lead, p.device
        lb = np.stack([lead.d00, lead.d01, np.conj(lead.d01).T])[None]
        n = lead.d00.shape[0]
        sets = _engine.TransportSets.from_arrays(
            lb, lb, np.conj(lead.d01).T[None], lead.d01[None], np.arange(n), np.arange(n),
            np.stack(dev.onsite_blocks)[None], np.stack(dev.coupling_blocks)[None], True)
        m = p.omegas.size
        jobs.append((_engine.TransmissionSolver(sets),
                     torch.zeros(m, dtype=torch.int32, device="cuda"),
                     torch.tensor(p.omegas, dtype=torch.float64, device="cuda"),
                     torch.full((m,), p.eta, dtype=torch.float64, device="cuda"), None))
    n_inst = sum(int(j[2].numel()) for j in jobs)
    outs_s = [j[0].alloc_result(int(j[2].numel())) for j in jobs]
    outs_p = [j[0].alloc_result(int(j[2].numel())) for j in jobs]

    def staged():
        for j, o in zip(jobs, outs_s):
            j[0].run(*j[1:], out=o)

    def persistent():
        _engine.transmission_multi(jobs, outs_p)

    times = {}
    for name, fn in (("staged", staged), ("persistent", persistent)):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        times[name] = (time.perf_counter() - t0) / steps
    same = all(torch.equal(a.t.nan_to_num(-1.0), b.t.nan_to_num(-1.0)) and
               torch.equal(a.iters_l, b.iters_l) for a, b in zip(outs_s, outs_p))
    iters = np.concatenate([o.iters_l.cpu().numpy() for o in outs_p])
    return {"workload": "uneven batch: ribbons of 24/48/96 DOF per layer, 8 layers, "
                        "4096/1024/256 omega, a quarter within 1% of the band top, eta=1e-6",
            "instances": n_inst, "decimation_steps": [int(iters.min()), int(iters.max())],
            "staged_solves_per_s": round(n_inst / times["staged"], 1),
            "persistent_solves_per_s": round(n_inst / times["persistent"], 1),
            "persistent_over_staged": round(times["staged"] / times["persistent"], 3),
            "identical_outputs": bool(same)}


def uneven_batch_3d(steps: int, warmup: int) -> dict:
    """The 3-D uneven batch (workloads.uneven3d_problems, silicon-like diamond cells):
    24-DOF cells on a 32 x 32
    k mesh (1,048,576 instances), 48-DOF on 8 x 8 and 96-DOF on 4 x 4, 1024 omega
    each with a quarter within 1 % of the band top, 8 layers; k-point blocks built
    on the GPU.  Solved by the staged launches problem after problem and by one
    persistent launch."""
    import torch

    from negfpy_amd.core import _engine
    from negfpy_amd.core.negf import kspace_sets
    from negfpy_amd.workloads import uneven3d_problems

    jobs = []
    for p in uneven3d_problems():
        sets = kspace_sets(p.device, p.lead, p.lead, p.kpoints)
        nk, nw = len(p.kpoints), p.omegas.size
        m = nk * nw
        jobs.append((_engine.TransmissionSolver(sets),
                     torch.arange(nk, dtype=torch.int32, device="cuda").repeat(nw),
                     torch.tensor(np.repeat(p.omegas, nk), dtype=torch.float64, device="cuda"),
                     torch.full((m,), p.eta, dtype=torch.float64, device="cuda"), None))
    sizes = [int(j[2].numel()) for j in jobs]
    n_inst = sum(sizes)
    outs_s = [j[0].alloc_result(m) for j, m in zip(jobs, sizes)]
    outs_p = [j[0].alloc_result(m) for j, m in zip(jobs, sizes)]

    def staged():
        for j, o in zip(jobs, outs_s):
            j[0].run(*j[1:], out=o)

    def persistent():
        _engine.transmission_multi(jobs, outs_p)

    times = {}
    for name, fn in (("staged", staged), ("persistent", persistent)):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        times[name] = (time.perf_counter() - t0) / steps
    same = all(torch.equal(a.t.nan_to_num(-1.0), b.t.nan_to_num(-1.0)) and
               torch.equal(a.iters_l, b.iters_l) and torch.equal(a.status, b.status)
               for a, b in zip(outs_s, outs_p))
    iters = torch.cat([o.iters_l for o in outs_p])
    failed = int(sum(int((o.status != 0).sum()) for o in outs_p))
    return {"workload": "3-D uneven batch: silicon-like diamond cells (1, 2 and 4 conventional "
                        "cubes) of 24/48/96 DOF on 32x32/8x8/4x4 k meshes, 8 layers, 1024 "
                        "omega each (a quarter within 1% of the band top), eta=1e-6, FP64",
            "instances": n_inst, "instances_per_part": sizes,
            "decimation_steps": [int(iters.min()), int(iters.max())],
            "failed_instances": failed,
            "staged_solves_per_s": round(n_inst / times["staged"], 1),
            "persistent_solves_per_s": round(n_inst / times["persistent"], 1),
            "persistent_over_staged": round(times["staged"] / times["persistent"], 3),
            "identical_outputs": bool(same)}


def _lib_ws(n: int, n_layers: int, checkpoint: int) -> int:
    from negfpy_amd import _lib

    return _lib.load().negf_adjoint_workspace_bytes(n, n_layers, 1, checkpoint)


def adjoint_batch(steps: int, warmup: int) -> dict:
    """The adjoint workload (48-DOF mass-mismatch interface between two different
    leads, 64 layers): forward transmission solves and adjoint solves (T and
    dT/d(every device block)) per second, one k point; also with the gradients of both
    leads (their three blocks and the device-to-lead couplings)."""
    import torch

    from negfpy_amd.core import _engine
    from negfpy_amd.workloads import interface_problem

    p = interface_problem()
    ll, lr, dev = p.lead_left, p.lead_right, p.device
    n = ll.d00.shape[0]
    sets = _engine.TransportSets.from_arrays(
        np.stack([ll.d00, ll.d01, np.conj(ll.d01).T])[None],
        np.stack([lr.d00, lr.d01, np.conj(lr.d01).T])[None], np.conj(ll.d01).T[None],
        lr.d01[None], np.arange(n), np.arange(n), np.stack(dev.onsite_blocks)[None],
        np.stack(dev.coupling_blocks)[None], False)
    m = p.omegas.size
    set_of = torch.zeros(m, dtype=torch.int32, device="cuda")
    omega = torch.tensor(p.omegas, dtype=torch.float64, device="cuda")
    eta = torch.full((m,), p.eta, dtype=torch.float64, device="cuda")
    solver = _engine.TransmissionSolver(sets)
    times, res = {}, {}
    for name, fn in (("forward", lambda: solver.run(set_of, omega, eta)),
                     ("adjoint", lambda: _engine.adjoint(sets, set_of, omega, eta, eta)),
                     ("adjoint_leads", lambda: _engine.adjoint(sets, set_of, omega, eta, eta,
                                                               lead_grads=True)),
                     ("adjoint_ckpt", lambda: _engine.adjoint(sets, set_of, omega, eta, eta,
                                                              checkpoint=8))):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            res[name] = fn()
        torch.cuda.synchronize()
        times[name] = (time.perf_counter() - t0) / steps
    tf, ta = res["forward"].t.cpu().numpy(), res["adjoint"].t.cpu().numpy()
    ok = (res["forward"].status.cpu().numpy() == 0) & (res["adjoint"].status.cpu().numpy() == 0)
    return {"workload": f"mass-mismatch interface: n={n} DOF/layer, {dev.n_layers} layers, "
                        f"28 | 73 amu leads, last 4 layers perturbed, {m} omega, eta=1e-6",
            "instances": m, "forward_solves_per_s": round(m / times["forward"], 1),
            "adjoint_solves_per_s": round(m / times["adjoint"], 1),
            "adjoint_over_forward_time": round(times["adjoint"] / times["forward"], 3),
            "adjoint_with_lead_gradients_solves_per_s": round(m / times["adjoint_leads"], 1),
            "adjoint_checkpoint_every_8_layers_solves_per_s": round(m / times["adjoint_ckpt"], 1),
            "adjoint_workspace_bytes_per_instance": {
                "stored": int(_lib_ws(n, dev.n_layers, 0)),
                "checkpoint_every_8": int(_lib_ws(n, dev.n_layers, 8))},
            "t_adjoint_vs_forward_max_rel": float(np.max(np.abs(ta[ok] - tf[ok])
                                                         / np.maximum(np.abs(tf[ok]), 1e-3)))}


def sheet_kavg(steps: int, warmup: int) -> dict:
    """The 2-D sheet workload (12-DOF graphene-like cells, 16 layers, interface-contact
    leads; 128 ky x 2048 omega = 262,144 instances per step).  A step builds the blocks
    of every k point on the GPU, solves every (omega, k) in one batch, averages T over
    k and integrates kappa(300 K) on the GPU; FP64 and mixed precision."""
    import torch

    from negfpy_amd.core import _engine
    from negfpy_amd.core.negf import kspace_sets
    from negfpy_amd.workloads import OMEGA_UNIT_RAD_S, sheet_problem

    p = sheet_problem()
    nk, nw = len(p.kpoints), p.omegas.size
    n_inst = nk * nw
    set_of = torch.arange(nk, dtype=torch.int32, device="cuda").repeat(nw)
    omega = torch.tensor(np.repeat(p.omegas, nk), dtype=torch.float64, device="cuda")
    eta = torch.full((n_inst,), p.eta, dtype=torch.float64, device="cuda")
    w_rad = torch.tensor(p.omegas * OMEGA_UNIT_RAD_S, dtype=torch.float64, device="cuda")
    temps = torch.tensor([300.0], dtype=torch.float64, device="cuda")
    res = {}
    # the leads are mirrored (one decimation gives both surfaces); "two" decimates each;
    # "mixed_persistent": the multi-problem entry (staged mixed decimation, then one
    # persistent launch for the self-energies, sweep and trace)
    for name, prec, mirror in (("fp64", "fp64", True), ("mixed", "mixed", True),
                               ("two", "fp64", False), ("mixed_persistent", "mixed", True)):
        out = None

        def step():
            nonlocal out
            sets = kspace_sets(p.device, p.lead_left, p.lead_right, p.kpoints,
                               device_to_lead_left=p.v_left, device_to_lead_right=p.v_right,
                               precision=prec)
            assert sets.mirror_leads
            if not mirror:
                sets = dataclasses.replace(sets, mirror_leads=False)
            solver = _engine.TransmissionSolver(sets, precision=prec)
            if out is None:
                out = solver.alloc_result(n_inst)
            if name == "mixed_persistent":
                _engine.transmission_multi([(solver, set_of, omega, eta, None)], outs=[out])
            else:
                solver.run(set_of, omega, eta, out=out)
            t = out.t.view(nw, nk).mean(dim=1)
            return t, _engine.thermal_conductance_tensor(w_rad, t, temps)

        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            t, kap = step()
        torch.cuda.synchronize()
        res[name] = {"dt": (time.perf_counter() - t0) / steps, "t": t.cpu().numpy(),
                     "kappa": float(kap.cpu()[0]),
                     "failed": int((out.status != 0).sum().item())}
    t64, t32 = res["fp64"]["t"], res["mixed"]["t"]
    ok = np.isfinite(t64) & np.isfinite(t32)
    rel_t = np.abs(t32[ok] - t64[ok]) / np.maximum(np.abs(t64[ok]), 1e-3)
    rel_k = abs(res["mixed"]["kappa"] - res["fp64"]["kappa"]) / abs(res["fp64"]["kappa"])
    return {"workload": f"graphene-like sheet: n=12 DOF/layer, 16 layers, interface-contact "
                        f"leads, {nk} ky x {nw} omega, eta={p.eta:g}; k-point blocks built "
                        f"on the GPU, T(omega) k-averaged and kappa(300 K) in each step",
            "instances": n_inst,
            "fp64_solves_per_s": round(n_inst / res["fp64"]["dt"], 1),
            "fp64_ms_per_step": round(res["fp64"]["dt"] * 1e3, 3),
            "fp64_two_decimations_ms_per_step": round(res["two"]["dt"] * 1e3, 3),
            "one_over_two_decimations": round(res["two"]["dt"] / res["fp64"]["dt"], 3),
            "t_one_vs_two_decimations_max_rel": float(np.max(
                np.abs(res["two"]["t"] - t64) / np.maximum(np.abs(res["two"]["t"]), 1e-3))),
            "mixed_solves_per_s": round(n_inst / res["mixed"]["dt"], 1),
            "mixed_over_fp64": round(res["fp64"]["dt"] / res["mixed"]["dt"], 3),
            "kappa_300K_W_per_K": res["fp64"]["kappa"],
            "mixed_kappa_rel_err": rel_k,
            "mixed_t_kavg_rel_err_max": float(rel_t.max()) if rel_t.size else None,
            "mixed_meets_1e-6": bool(rel_t.size and rel_t.max() <= 1e-6 and rel_k <= 1e-6),
            "mixed_persistent_ms_per_step": round(res["mixed_persistent"]["dt"] * 1e3, 3),
            "mixed_persistent_t_kavg_rel_err_max": float(np.max(
                np.abs(res["mixed_persistent"]["t"] - t64) / np.maximum(np.abs(t64), 1e-3))),
            "failed_instances": {"fp64": res["fp64"]["failed"],
                                 "mixed": res["mixed"]["failed"],
                                 "mixed_persistent": res["mixed_persistent"]["failed"]}}


def kappa_gradient_3d() -> dict:
    """The 3-D gradient workload at full size: 48-DOF simple-cubic cells, 64 layers,
    28 | 73 amu leads, the last 4 layers perturbed, a 16 x 16 (ky, kz) mesh x 1024
    omega = 262,144 instances.  kappa(300 K) and d kappa / d(every device
    force-constant term, 64 x 9 same-layer and 63 x 9 next-layer 48 x 48 blocks)
    through the k-point blocks (torch) and transmission_sum_torch: kappa is
    sum_i c_i T_i with c_i = conductance weight / n_k, so one adjoint pass gives the
    value and the gradient.  Timed once."""
    import torch

    from negfpy_amd.core import _engine, transmission_sum_torch
    from negfpy_amd.workloads import OMEGA_UNIT_RAD_S, interface3d_blocks, interface3d_problem

    p = interface3d_problem()
    nk, nw = len(p.kpoints), p.omegas.size
    n_inst = nk * nw
    k_of = torch.arange(nk, dtype=torch.int32, device="cuda").repeat(nw)
    om = torch.tensor(np.repeat(p.omegas, nk), dtype=torch.float64, device="cuda")
    up = [torch.as_tensor(x, device="cuda") for x in (p.lead_left, p.lead_right, p.v_left,
                                                      p.v_right)]
    f00 = torch.tensor(p.fc00, dtype=torch.float64, device="cuda", requires_grad=True)
    f01 = torch.tensor(p.fc01, dtype=torch.float64, device="cuda", requires_grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    wts = _engine.conductance_weights(
        torch.as_tensor(p.omegas * OMEGA_UNIT_RAD_S, device="cuda"),
        torch.tensor([300.0], dtype=torch.float64, device="cuda"))[0]
    ons, cpl = interface3d_blocks(p, f00, f01)
    kap = transmission_sum_torch(wts.repeat_interleave(nk) / nk, om, k_of, ons, cpl, *up,
                                 eta=p.eta)
    kap.backward()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    g00, g01 = f00.grad, f01.grad
    n = int(p.fc00.shape[-1])
    return {"workload": f"3-D mass-mismatch interface: n={n} DOF/layer (4x2x2 simple-cubic "
                        f"cell), {p.fc00.shape[0]} layers, 28 | 73 amu, last 4 layers "
                        f"perturbed +-20%, {int(round(nk ** 0.5))}x{int(round(nk ** 0.5))} k "
                        f"x {nw} omega, eta={p.eta:g}, FP64",
            "instances": n_inst,
            "kappa_and_gradient_s": round(dt, 3),
            "adjoint_solves_per_s": round(n_inst / dt, 1),
            "kappa_300K_W_per_K": float(kap.detach()),
            "gradient_terms": int(g00.numel() + g01.numel()),
            "grad_norm_perturbed_layers": float(g00[-4:].norm()),
            "grad_norm_all": float((g00.norm() ** 2 + g01.norm() ** 2).sqrt())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-gradient", action="store_true",
                    help="skip the full-size d kappa / d(force constants) run (about 30 s)")
    ap.add_argument("--paths", default="fp64,mixed",
                    help="comma list of precision paths to run (profiling one path alone)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        src = torch.ones(args.calibrate_copy * (1 << 18), dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_mib": args.calibrate_copy}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X; "
              "the multi-GPU path (sharding k points, one all-reduce of T(omega)) is not run.",
              file=sys.stderr)
        sys.exit(2)

    from negfpy_amd.workloads import ribbon_problem

    problem = ribbon_problem(seed=0, n_atoms=8, n_layers=32, n_omega=4096, eta=1e-6)
    n = problem.lead.d00.shape[0]
    n_layers = problem.device.n_layers
    n_inst = problem.omegas.size

    cpu = None
    if not args.no_cpu:
        sel, cpu_vals, wall, workers = cpu_baseline(problem, args.cpu_sample)
        cpu = {"sel": sel, "vals": cpu_vals, "wall": wall, "workers": workers}
        print(f"cpu baseline: {sel.size} solves in {wall:.2f} s on {workers} processes",
              file=sys.stderr, flush=True)

    import torch

    from negfpy_amd.core import _engine
    from negfpy_amd.core.negf import resolve_k

    torch.cuda.set_device(0)
    kp = resolve_k(problem.device, problem.lead, problem.lead, None, eta=problem.eta,
                   eta_device=None, device_to_lead_left=None, device_to_lead_right=None,
                   contact_left_indices=None, contact_right_indices=None, omega_scale=None,
                   max_block=64)
    sets = _engine.TransportSets.from_arrays(kp.lead_l[None], kp.lead_r[None], kp.v_l[None],
                                             kp.v_r[None], kp.pos_l, kp.pos_r, kp.onsite[None],
                                             kp.coupling[None], True)
    set_of = torch.zeros(n_inst, dtype=torch.int32, device="cuda")
    omega = torch.tensor(problem.omegas, dtype=torch.float64, device="cuda")
    eta = torch.full((n_inst,), problem.eta, dtype=torch.float64, device="cuda")

    paths = [q for q in args.paths.split(",") if q]
    results = {}
    for prec in paths:
        solver = _engine.TransmissionSolver(sets, precision=prec)
        out = solver.alloc_result(n_inst)
        for _ in range(args.warmup):
            solver.run(set_of, omega, eta, out=out)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            solver.run(set_of, omega, eta, out=out)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        # per-kernel times: the same kernels launched stage by stage between HIP events
        ev: dict = {}
        stats: dict = {}
        for _ in range(max(3, min(args.steps, 10))):
            solver.run_staged(set_of, omega, eta, events=ev, stats=stats)
        torch.cuda.synchronize()
        stage_ms = {k: float(np.mean([a.elapsed_time(b) for a, b in v])) for k, v in ev.items()}
        status = out.status.cpu().numpy()
        results[prec] = {"t": out.t.cpu().numpy(), "iters": out.iters_l.cpu().numpy(),
                         "dos": out.dos_l.cpu().numpy(), "status": status, "s_per_step": dt,
                         "stage_ms": stage_ms, "stats": stats}
        print(f"{prec}: {dt * 1e3:.3f} ms/step, {n_inst / dt:.1f} solves/s, stages {stage_ms}",
              file=sys.stderr, flush=True)

    if set(paths) != {"fp64", "mixed"}:
        print(json.dumps({"paths": paths, "ms_per_step": {k: v["s_per_step"] * 1e3
                                                          for k, v in results.items()},
                          "stage_ms": {k: v["stage_ms"] for k, v in results.items()}}))
        return
    t64 = results["fp64"]["t"]
    t32 = results["mixed"]["t"]
    ok = (results["fp64"]["status"] == 0) & (results["mixed"]["status"] == 0)
    gold_path = os.path.join(ROOT, "tests", "golden", "ribbon24_oracle.npz")
    gold = np.load(gold_path) if os.path.exists(gold_path) else None
    if gold is not None and not np.array_equal(gold["omega"], problem.omegas):
        gold = None
    t_ref = gold["t"] if gold is not None else t64
    scale = np.maximum(np.abs(t_ref), 1e-3)
    rel = np.where(ok, np.abs(t32 - t_ref) / scale, np.inf)
    rel64 = np.where(ok, np.abs(t64 - t_ref) / scale, np.inf)
    rel_vs64 = np.where(ok, np.abs(t32 - t64) / np.maximum(np.abs(t64), 1e-3), np.inf)
    d64, d32 = results["fp64"]["dos"][ok], results["mixed"]["dos"][ok]
    rel_dos = np.abs(d32 - d64) / np.maximum(np.abs(d64), 1e-3 * np.abs(d64).max())
    from negfpy_amd.core import _engine as _eng
    from negfpy_amd.workloads import OMEGA_UNIT_RAD_S

    w_rad = problem.omegas * OMEGA_UNIT_RAD_S
    k_ref = _eng.thermal_conductance(w_rad, np.where(ok, t_ref, 0.0), [300.0])[0]
    k64 = _eng.thermal_conductance(w_rad, np.where(ok, t64, 0.0), [300.0])[0]
    k32 = _eng.thermal_conductance(w_rad, np.where(ok, t32, 0.0), [300.0])[0]
    rel_k = abs(k32 - k_ref) / abs(k_ref)

    mix, f64 = results["mixed"], results["fp64"]
    n3 = float(n) ** 3
    s64 = float(np.mean(f64["iters"]))
    cert = mix["stats"].get("fp32_certified_l", 0) / n_inst
    # algorithmic real flops per instance of each stage (complex multiply-add = 8
    # flops).  "survey": the hot-path table's formulas -- decimation (56 s + 8) n^3
    # per lead (one decimation serves both identical leads), sweep 8 n^3 + 40 (L - 1)
    # n^3, self-energy 16 n^3 per lead, trace 16 n^3; the "sweep" stage is one kernel
    # that forms both self-energies, runs the sweep and takes the trace.  "kernel": what
    # the kernels execute -- the sweep carries G_1N with one product per layer,
    # 8 n^3 + 32 (L - 1) n^3.
    fl_survey = {"sancho_rubio_l": (56.0 * s64 + 8.0) * n3,
                 "sweep": 8.0 * n3 + 40.0 * (n_layers - 1) * n3 + 3 * 16.0 * n3}
    fl_kernel = dict(fl_survey, sweep=8.0 * n3 + 32.0 * (n_layers - 1) * n3 + 3 * 16.0 * n3)
    fp64_fl_total = sum(fl_survey.values())

    def roofline(path, dom=None):
        st = results[path]["stage_ms"]
        cand = [k for k in ("sancho_rubio_l", "sweep") if k in st]
        dom = dom or max(cand, key=lambda k: st[k])
        fs, fk = fl_survey[dom], fl_kernel[dom]
        peak = PEAK_TFLOPS["fp64"]
        if path == "mixed" and dom == "sancho_rubio_l":  # the FP32 decimation pass
            s32 = float(np.mean(results[path]["iters"]))
            fs = fk = 56.0 * s32 * n3
            peak = PEAK_TFLOPS["mixed"]
        ach = fs * n_inst / (st[dom] * 1e-3) / 1e12
        ach_k = fk * n_inst / (st[dom] * 1e-3) / 1e12
        name = {"sancho_rubio_l": "sancho_rubio", "sweep": "sweep"}[dom]
        return {"bound": "mfma", "kernel": name, "achieved": round(ach, 3), "peak": peak,
                "unit": "TFLOP/s", "frac": round(ach / peak, 5), "traffic": TRAFFIC.get(name),
                "flops_per_instance": fs,
                "kernel_count": {"flops_per_instance": fk, "achieved": round(ach_k, 3),
                                 "frac": round(ach_k / peak, 5)},
                "stage_ms": round(st[dom], 4)}

    mixed_value = n_inst / mix["s_per_step"]
    fp64_value = n_inst / f64["s_per_step"]
    # the metric counts mixed-precision solves only at <= 1e-6 relative error in T
    # (against the oracle at every frequency), DOS and kappa(300 K)
    mixed_ok = bool(np.all(np.isfinite(rel)) and rel.max() <= 1e-6 and rel_dos.max() <= 1e-6
                    and rel_k <= 1e-6)
    head = "mixed" if mixed_ok else "fp64"
    value = mixed_value if mixed_ok else fp64_value
    above = (np.abs(t64 - t_ref) > 1e-10 * scale) & ok
    spread_ratio = None
    if gold is not None and above.any():
        spread_ratio = float(np.max(np.abs(t64 - t_ref)[above] / gold["spread"][above]))
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(results[head]["s_per_step"] * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("complex64 decimation + fp64 (Newton certificate, uncertified decimations, "
                  "sweep, trace)" if mixed_ok else "complex128"),
        "headline_path": head,
        "mixed_meets_1e-6": mixed_ok,
        "mixed_value": round(mixed_value, 1),
        "data": "synthetic",
        "config": {"workload": "multi-DOF ribbon: n=24 DOF/layer, 32 layers, 4096 omega, eta=1e-6",
                   "instances": n_inst, "block": n, "layers": n_layers},
        "fp64_value": round(fp64_value, 1),
        "fp64_ms_per_step": round(f64["s_per_step"] * 1e3, 4),
        "mixed_over_fp64": round(mixed_value / fp64_value, 3),
        "mixed_fp32_certified_fraction": round(cert, 4),
        "mixed_rel_err_vs_oracle": {"max": float(rel.max()), "median": float(np.median(rel)),
                                    "frac_above_1e-6": float(np.mean(rel > 1e-6)),
                                    "floor": "relative to max(|T|, 1e-3)",
                                    "reference": "oracle, every frequency" if gold is not None
                                    else "fp64 path"},
        "mixed_rel_err_vs_fp64": {"max": float(rel_vs64.max()),
                                  "median": float(np.median(rel_vs64))},
        "mixed_surface_dos_rel_err_vs_fp64": {"max": float(rel_dos.max()),
                                              "median": float(np.median(rel_dos)),
                                              "floor": "relative to max(|DOS|, 1e-3 max DOS)"},
        "mixed_kappa_300K_rel_err": rel_k,
        "fp64_rel_err_vs_oracle": {"max": float(rel64.max()),
                                   "points_above_1e-10": int(above.sum()),
                                   "max_dT_over_oracle_spread_there": spread_ratio},
        "fp64_kappa_300K_rel_err": abs(k64 - k_ref) / abs(k_ref),
        "mean_decimation_steps": s64,
        "fp64_flops_per_instance": fp64_fl_total,
        "roofline": roofline(head),
        "roofline_fp64_path": roofline("fp64") if head != "fp64" else None,
        # both FP64 kernels of the step (the decimation and the fused sweep), each
        # against the FP64 matrix peak
        "roofline_fp64_kernels": {k: roofline("fp64", s) for k, s in
                                  (("sancho_rubio", "sancho_rubio_l"), ("sweep", "sweep"))},
        "stage_ms": {"mixed": mix["stage_ms"], "fp64": f64["stage_ms"]},
    }
    line["uneven_batch"] = uneven_batch(max(1, min(args.steps, 5)), 1)
    print(f"uneven batch: {line['uneven_batch']}", file=sys.stderr, flush=True)
    line["uneven_batch_3d"] = uneven_batch_3d(max(1, min(args.steps, 2)), 1)
    print(f"uneven batch 3-D: {line['uneven_batch_3d']}", file=sys.stderr, flush=True)
    line["adjoint_batch"] = adjoint_batch(max(1, min(args.steps, 3)), 1)
    print(f"adjoint batch: {line['adjoint_batch']}", file=sys.stderr, flush=True)
    line["sheet_kavg"] = sheet_kavg(max(1, min(args.steps, 3)), 1)
    print(f"sheet k-average: {line['sheet_kavg']}", file=sys.stderr, flush=True)
    if not args.no_gradient:
        line["kappa_gradient_3d"] = kappa_gradient_3d()
        print(f"kappa gradient 3-D: {line['kappa_gradient_3d']}", file=sys.stderr, flush=True)
    if cpu is not None:
        sel = cpu["sel"]
        err64 = np.abs(t64[sel] - cpu["vals"]) / np.maximum(np.abs(cpu["vals"]), 1e-3)
        line["fp64_rel_err_vs_cpu_baseline_sample"] = float(err64.max())
        line["cpu_baseline"] = {
            "value": round(sel.size / cpu["wall"], 2), "unit": "solves/s",
            "cores": cpu["workers"], "kind": "port",
            "sample": f"{sel.size} of the {n_inst} instances (evenly spaced omega), FP64 NumPy/"
                      f"SciPy oracle: decimation + SuperLU, 1 BLAS thread per process, "
                      f"{cpu['workers']} processes",
            "host": host_description()}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
