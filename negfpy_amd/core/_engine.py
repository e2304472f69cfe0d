"""Batched GPU engine: packs block sets into device tensors and drives the HIP
library through its C ABI.

A *set* holds everything that depends on k_par only (lead blocks, device
blocks, device-to-lead couplings); an *instance* is one (omega, k_par, eta)
point that refers to its set.  All per-set blocks are uploaded once and reused
by every omega of that k_par.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from negfpy_amd import _lib

PRECISIONS = {"fp64": _lib.FP64, "mixed": _lib.MIXED}
DEFAULT_REFINE_STEPS = 2


def precision_code(precision: str) -> int:
    try:
        return PRECISIONS[precision]
    except KeyError:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}, got {precision!r}") from None


def device() -> torch.device:
    _lib.load()
    return torch.device("cuda", torch.cuda.current_device())


def to_dev_c128(x) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(device=device(), dtype=torch.complex128).contiguous()
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.complex128)).to(device())


def to_dev_f64(x) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(device=device(), dtype=torch.float64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(device())


def to_dev_i32(x) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(device=device(), dtype=torch.int32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.int32)).to(device())


# ---------------------------------------------------------------------------
# surface Green's functions
# ---------------------------------------------------------------------------

@dataclass
class SurfaceResult:
    g: torch.Tensor        # (n_inst, n, n) complex128
    dos: torch.Tensor      # (n_inst,) float64
    iters: torch.Tensor    # (n_inst,) int32
    status: torch.Tensor   # (n_inst,) int32


def surface_gf_batch(leads, set_of, omega, eta, *, tol=1e-12, maxiter=200, precision="fp64",
                     refine_steps=DEFAULT_REFINE_STEPS, route=0) -> SurfaceResult:
    """Decimate every instance's lead.  ``leads`` is (n_sets, 3, n, n) = (d00, d01, d10).
    With ``route`` = SURFACE_EIGEN, instances whose decimation fails take the
    eigenmode route (blocks up to 48)."""
    lib = _lib.load()
    leads = to_dev_c128(leads)
    if leads.ndim != 4 or leads.shape[1] != 3 or leads.shape[2] != leads.shape[3]:
        raise ValueError("leads must have shape (n_sets, 3, n, n)")
    n = int(leads.shape[2])
    set_of = to_dev_i32(set_of)
    omega = to_dev_f64(omega)
    eta = to_dev_f64(eta)
    n_inst = int(omega.numel())
    if set_of.numel() != n_inst or eta.numel() != n_inst:
        raise ValueError("set_of, omega and eta must have one entry per instance")
    dev = leads.device
    g = torch.empty((n_inst, n, n), dtype=torch.complex128, device=dev)
    dos = torch.empty((n_inst,), dtype=torch.float64, device=dev)
    iters = torch.zeros((n_inst,), dtype=torch.int32, device=dev)
    status = torch.zeros((n_inst,), dtype=torch.int32, device=dev)
    _lib.check(lib.negf_sancho_rubio_batched(
        _lib.stream(), precision_code(precision), n, n_inst, leads.data_ptr(), set_of.data_ptr(),
        omega.data_ptr(), eta.data_ptr(), float(tol), int(maxiter), int(refine_steps),
        g.data_ptr(), dos.data_ptr(), iters.data_ptr(), status.data_ptr()))
    if route == _lib.SURFACE_EIGEN:
        _lib.check(lib.negf_surface_gf_eigen_batched(
            _lib.stream(), n, n_inst, leads.data_ptr(), set_of.data_ptr(), omega.data_ptr(),
            eta.data_ptr(), g.data_ptr(), dos.data_ptr(), status.data_ptr()))
    return SurfaceResult(g=g, dos=dos, iters=iters, status=status)


# ---------------------------------------------------------------------------
# transmission
# ---------------------------------------------------------------------------

@dataclass
class TransportSets:
    """Device tensors of one transport problem, one entry per k_par set."""

    lead_l: torch.Tensor     # (S, 3, nl, nl)
    lead_r: torch.Tensor     # (S, 3, nr, nr)
    v_l: torch.Tensor        # (S, ncl, nl)
    v_r: torch.Tensor        # (S, ncr, nr)
    idx_l: torch.Tensor      # (ncl,) int32, positions in the first layer
    idx_r: torch.Tensor      # (ncr,) int32, positions in the last layer
    onsite: torch.Tensor     # (S, L, n, n)
    coupling: torch.Tensor   # (S, max(L-1,1), n, n)
    share_leads: bool
    # the right lead is the left one with d01 and d10 exchanged in every set (the
    # interface contacts of a bulk force-constant model): one decimation gives both
    mirror_leads: bool = False

    @property
    def n(self) -> int:
        return int(self.onsite.shape[2])

    @property
    def n_layers(self) -> int:
        return int(self.onsite.shape[1])

    @property
    def n_sets(self) -> int:
        return int(self.onsite.shape[0])

    @classmethod
    def from_arrays(cls, lead_l, lead_r, v_l, v_r, idx_l, idx_r, onsite, coupling,
                    share_leads: bool) -> "TransportSets":
        onsite = to_dev_c128(onsite)
        s, nl_, n, _ = onsite.shape
        if nl_ > 1:
            coupling = to_dev_c128(coupling)
        else:
            coupling = torch.zeros((s, 1, n, n), dtype=torch.complex128, device=onsite.device)
        lead_l = to_dev_c128(lead_l)
        lead_r = lead_l if share_leads else to_dev_c128(lead_r)
        return cls(lead_l=lead_l, lead_r=lead_r, v_l=to_dev_c128(v_l), v_r=to_dev_c128(v_r),
                   idx_l=to_dev_i32(idx_l), idx_r=to_dev_i32(idx_r), onsite=onsite,
                   coupling=coupling, share_leads=bool(share_leads),
                   mirror_leads=not share_leads and mirrored(lead_l, lead_r))


def mirrored(lead_l: torch.Tensor, lead_r: torch.Tensor) -> bool:
    """lead_r = (d00, d10, d01) of lead_l in every set, exactly."""
    return (lead_l.shape == lead_r.shape and torch.equal(lead_l[:, 0], lead_r[:, 0])
            and torch.equal(lead_l[:, 1], lead_r[:, 2]) and torch.equal(lead_l[:, 2], lead_r[:, 1]))


@dataclass
class LadderResult:
    rung: torch.Tensor     # (n_inst,) int32: accepted rung, or the rung count if none was
    t: torch.Tensor        # (n_inst, max_rungs) float64: T of every attempted rung
    status: torch.Tensor   # (n_inst, max_rungs) int32: status of every attempted rung


@dataclass
class TransmissionResult:
    t: torch.Tensor        # (n_inst,) float64
    status: torch.Tensor   # (n_inst,) int32
    iters_l: torch.Tensor
    iters_r: torch.Tensor
    dos_l: torch.Tensor
    dos_r: torch.Tensor
    ladder: LadderResult | None = None


class EtaLadder:
    """Eta ladders of the instances of one :func:`transmission_multi` job: instance i
    tries ``ladders[i]`` in order (eta_lead = the rung's eta; eta_device the same, or
    the job's ``eta_dev`` when ``eta_dev_fixed``) until a rung is accepted: status 0,
    finite T, T >= -nonneg_tol, T <= t_max[i] (None: no limit)."""

    def __init__(self, ladders, *, t_max=None, nonneg_tol=0.0, eta_dev_fixed=False):
        count = np.array([len(lad) for lad in ladders], dtype=np.int32)
        first = np.zeros(count.size, dtype=np.int32)
        if count.size > 1:
            first[1:] = np.cumsum(count)[:-1]
        etas = (np.concatenate([np.asarray(lad, dtype=float) for lad in ladders])
                if count.sum() else np.zeros(1))
        self.n_inst = int(count.size)
        self.max_rungs = max(int(count.max()) if count.size else 1, 1)
        self.etas = to_dev_f64(etas)
        self.first = to_dev_i32(first)
        self.count = to_dev_i32(count)
        self.t_max = None
        if t_max is not None:
            self.t_max = to_dev_f64([np.inf if v is None else float(v) for v in t_max])
        self.nonneg_tol = float(nonneg_tol)
        self.eta_dev_fixed = bool(eta_dev_fixed)


class TransmissionSolver:
    """Reusable batched solver: workspace is kept between calls of equal size, so the
    calls of one solver must stay on one stream (use one solver per stream or thread;
    the public functions make a solver per call)."""

    def __init__(self, sets: TransportSets, *, precision="fp64", tol=1e-12, maxiter=200,
                 refine_steps=DEFAULT_REFINE_STEPS, surface_route=0):
        self.sets = sets
        self.precision = precision_code(precision)
        self.surface_route = int(surface_route)
        self.tol = float(tol)
        self.maxiter = int(maxiter)
        self.refine_steps = int(refine_steps)
        self._ws = None
        self._lib = _lib.load()

    def _problem(self, n_inst, set_of, omega, eta_lead, eta_dev, out):
        s = self.sets
        p = _lib.TransmissionProblem()
        p.precision = self.precision
        p.n = s.n
        p.n_layers = s.n_layers
        p.n_inst = n_inst
        p.n_lead_l = int(s.lead_l.shape[2])
        p.n_lead_r = int(s.lead_r.shape[2])
        p.n_c_l = int(s.v_l.shape[1])
        p.n_c_r = int(s.v_r.shape[1])
        p.share_leads = (_lib.LEADS_SHARED if s.share_leads else
                         _lib.LEADS_MIRRORED if s.mirror_leads else _lib.LEADS_INDEPENDENT)
        p.surface_route = self.surface_route
        p.maxiter = self.maxiter
        p.refine_steps = self.refine_steps
        p.tol = self.tol
        p.lead_l = s.lead_l.data_ptr()
        p.lead_r = s.lead_r.data_ptr()
        p.v_l = s.v_l.data_ptr()
        p.v_r = s.v_r.data_ptr()
        p.idx_l = s.idx_l.data_ptr()
        p.idx_r = s.idx_r.data_ptr()
        p.onsite = s.onsite.data_ptr()
        p.coupling = s.coupling.data_ptr()
        p.set_of = set_of.data_ptr()
        p.omega = omega.data_ptr()
        p.eta_lead = eta_lead.data_ptr()
        p.eta_dev = eta_dev.data_ptr()
        p.t_out = out.t.data_ptr()
        p.dos_l = out.dos_l.data_ptr()
        p.dos_r = out.dos_r.data_ptr()
        p.iters_l = out.iters_l.data_ptr()
        p.iters_r = out.iters_r.data_ptr()
        p.status = out.status.data_ptr()
        return p

    def alloc_result(self, n_inst: int) -> TransmissionResult:
        dev = self.sets.onsite.device
        return TransmissionResult(
            t=torch.empty((n_inst,), dtype=torch.float64, device=dev),
            status=torch.zeros((n_inst,), dtype=torch.int32, device=dev),
            iters_l=torch.zeros((n_inst,), dtype=torch.int32, device=dev),
            iters_r=torch.zeros((n_inst,), dtype=torch.int32, device=dev),
            dos_l=torch.zeros((n_inst,), dtype=torch.float64, device=dev),
            dos_r=torch.zeros((n_inst,), dtype=torch.float64, device=dev))

    def run(self, set_of, omega, eta_lead, eta_dev=None, out: TransmissionResult | None = None
            ) -> TransmissionResult:
        set_of = to_dev_i32(set_of)
        omega = to_dev_f64(omega)
        eta_lead = to_dev_f64(eta_lead)
        eta_dev = eta_lead if eta_dev is None else to_dev_f64(eta_dev)
        n_inst = int(omega.numel())
        if not (set_of.numel() == eta_lead.numel() == eta_dev.numel() == n_inst):
            raise ValueError("set_of, omega, eta_lead, eta_dev need one entry per instance")
        if out is None:
            out = self.alloc_result(n_inst)
        if n_inst == 0:
            return out
        p = self._problem(n_inst, set_of, omega, eta_lead, eta_dev, out)
        need = int(self._lib.negf_transmission_workspace_bytes(ctypes.byref(p)))
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty((need,), dtype=torch.uint8, device=self.sets.onsite.device)
        _lib.check(self._lib.negf_transmission_batched(
            _lib.stream(), ctypes.byref(p), self._ws.data_ptr(), need))
        if self.sets.share_leads:
            out.iters_r.copy_(out.iters_l)
            out.dos_r.copy_(out.dos_l)
        return out


    def run_staged(self, set_of, omega, eta_lead, eta_dev=None, events=None, stats=None
                   ) -> TransmissionResult:
        """Same computation as :meth:`run`, one C-ABI call per stage.

        ``events`` (a dict) receives a pair of torch.cuda.Event per stage, recorded
        on the launch stream around the stage's single kernel, so callers can time
        every kernel separately.  ``stats`` (a dict), in mixed precision, receives
        per lead side the number of instances whose FP32 decimation the Newton
        completion certified (``fp32_certified_l``/``_r``) out of ``instances``.
        """
        lib = self._lib
        s = self.sets
        set_of = to_dev_i32(set_of)
        omega = to_dev_f64(omega)
        eta_lead = to_dev_f64(eta_lead)
        eta_dev = eta_lead if eta_dev is None else to_dev_f64(eta_dev)
        n_inst = int(omega.numel())
        out = self.alloc_result(n_inst)
        dev = s.onsite.device
        nl, nr, n = int(s.lead_l.shape[2]), int(s.lead_r.shape[2]), s.n
        c128 = torch.complex128
        g_l = torch.empty((n_inst, nl, nl), dtype=c128, device=dev)
        g_r = g_l if s.share_leads else torch.empty((n_inst, nr, nr), dtype=c128, device=dev)
        sig_l = torch.empty((n_inst, n, n), dtype=c128, device=dev)
        sig_r = torch.empty((n_inst, n, n), dtype=c128, device=dev)
        g1n = torch.empty((n_inst, n, n), dtype=c128, device=dev)
        st_l = torch.zeros((n_inst,), dtype=torch.int32, device=dev)
        st_r = st_l.clone() if not s.share_leads else torch.zeros_like(st_l)
        st = _lib.stream()

        def stage(name, fn):
            if events is not None:
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record()
                _lib.check(fn())
                b.record()
                events.setdefault(name, []).append((a, b))
            else:
                _lib.check(fn())

        mixed = self.precision == _lib.MIXED
        sides = [("l", s.lead_l, g_l, out.dos_l, out.iters_l, st_l)]
        if not s.share_leads:
            sides.append(("r", s.lead_r, g_r, out.dos_r, out.iters_r, st_r))
        for side, lead, g, dos, iters, stt in sides:
            nn = int(lead.shape[2])
            stage(f"sancho_rubio_{side}", lambda lead=lead, g=g, dos=dos, iters=iters, stt=stt, nn=nn:
                  lib.negf_sancho_rubio_batched(
                      st, self.precision, nn, n_inst, lead.data_ptr(), set_of.data_ptr(),
                      omega.data_ptr(), eta_lead.data_ptr(), self.tol, self.maxiter,
                      -1 if mixed else self.refine_steps, g.data_ptr(), dos.data_ptr(),
                      iters.data_ptr(), stt.data_ptr()))
            if mixed:
                stage(f"refine_{side}", lambda lead=lead, g=g, dos=dos, stt=stt, nn=nn:
                      lib.negf_surface_refine_batched(
                          st, nn, n_inst, lead.data_ptr(), set_of.data_ptr(), omega.data_ptr(),
                          eta_lead.data_ptr(), self.refine_steps, g.data_ptr(), dos.data_ptr(),
                          stt.data_ptr()))
                if stats is not None:
                    stats[f"fp32_certified_{side}"] = int((stt == 0).sum().item())
                    stats["instances"] = n_inst
                stage(f"redo_{side}", lambda lead=lead, g=g, dos=dos, iters=iters, stt=stt, nn=nn:
                      lib.negf_sancho_rubio_redo(
                          st, nn, n_inst, lead.data_ptr(), set_of.data_ptr(), omega.data_ptr(),
                          eta_lead.data_ptr(), self.tol, self.maxiter, g.data_ptr(),
                          dos.data_ptr(), iters.data_ptr(), stt.data_ptr()))
            if self.surface_route == _lib.SURFACE_EIGEN:
                stage(f"eigen_{side}", lambda lead=lead, g=g, dos=dos, stt=stt, nn=nn:
                      lib.negf_surface_gf_eigen_batched(
                          st, nn, n_inst, lead.data_ptr(), set_of.data_ptr(), omega.data_ptr(),
                          eta_lead.data_ptr(), g.data_ptr(), dos.data_ptr(), stt.data_ptr()))
        if s.share_leads:
            st_r.copy_(st_l)
        # the self-energies, the device sweep (FP64 in both precisions) and the trace:
        # one kernel where the block class has the fused form, as in
        # negf_transmission_batched
        stage("sweep", lambda: lib.negf_transmission_sweep_batched(
            st, n, s.n_layers, n_inst, s.onsite.data_ptr(), s.coupling.data_ptr(),
            set_of.data_ptr(), omega.data_ptr(), eta_dev.data_ptr(), nl, int(s.v_l.shape[1]),
            s.v_l.data_ptr(), g_l.data_ptr(), s.idx_l.data_ptr(), st_l.data_ptr(), nr,
            int(s.v_r.shape[1]), s.v_r.data_ptr(), g_r.data_ptr(), s.idx_r.data_ptr(),
            st_r.data_ptr(), sig_l.data_ptr(), sig_r.data_ptr(), g1n.data_ptr(),
            out.status.data_ptr(), out.t.data_ptr()))
        if s.share_leads:
            out.iters_r.copy_(out.iters_l)
            out.dos_r.copy_(out.dos_l)
        return out


def _ladder_struct(lad: EtaLadder, out: TransmissionResult):
    dev = out.t.device
    out.ladder = LadderResult(
        rung=torch.zeros((lad.n_inst,), dtype=torch.int32, device=dev),
        t=torch.full((lad.n_inst, lad.max_rungs), float("nan"), dtype=torch.float64, device=dev),
        status=torch.full((lad.n_inst, lad.max_rungs), -1, dtype=torch.int32, device=dev))
    c = _lib.EtaLadder()
    c.etas = lad.etas.data_ptr()
    c.first = lad.first.data_ptr()
    c.count = lad.count.data_ptr()
    c.t_max = None if lad.t_max is None else lad.t_max.data_ptr()
    c.rung_out = out.ladder.rung.data_ptr()
    c.t_rungs = out.ladder.t.data_ptr()
    c.status_rungs = out.ladder.status.data_ptr()
    c.nonneg_tol = lad.nonneg_tol
    c.max_rungs = lad.max_rungs
    c.eta_dev_fixed = 1 if lad.eta_dev_fixed else 0
    return c


def _multi_prepare(jobs, outs):
    prepared, keep = [], []
    for j, job in enumerate(jobs):
        solver, set_of, omega, eta_lead, eta_dev = job[:5]
        lad = job[5] if len(job) > 5 else None
        if solver.precision != _lib.FP64 and len(job) > 5 and job[5] is not None:
            raise ValueError("eta ladders in the persistent launch run FP64 only")
        set_of = to_dev_i32(set_of)
        omega = to_dev_f64(omega)
        n_inst = int(omega.numel())
        if lad is not None and eta_lead is None:
            eta_lead = torch.zeros((n_inst,), dtype=torch.float64, device=omega.device)
        eta_lead = to_dev_f64(eta_lead)
        eta_dev = eta_lead if eta_dev is None else to_dev_f64(eta_dev)
        if not (set_of.numel() == eta_lead.numel() == eta_dev.numel() == n_inst):
            raise ValueError("set_of, omega, eta_lead, eta_dev need one entry per instance")
        if lad is not None and lad.n_inst != n_inst:
            raise ValueError("the eta ladder needs one ladder per instance")
        out = outs[j] if outs is not None else solver.alloc_result(n_inst)
        prepared.append((solver, set_of, omega, eta_lead, eta_dev, out, lad))
    chunks = []
    for c0 in range(0, len(prepared), _lib.MULTI_MAX_PROBLEMS):
        chunk = prepared[c0:c0 + _lib.MULTI_MAX_PROBLEMS]
        arr = (_lib.TransmissionProblem * len(chunk))()
        for i, (solver, so, om, el, ed, out, lad) in enumerate(chunk):
            arr[i] = solver._problem(int(om.numel()), so, om, el, ed, out)
            if lad is not None:
                c = _ladder_struct(lad, out)
                keep.append(c)
                arr[i].ladder = ctypes.pointer(c)
        need = int(_lib.load().negf_transmission_multi_workspace_bytes(len(chunk), arr))
        chunks.append((arr, need))
    return prepared, (chunks, keep)


def _multi_launch(prepared, chunks, ws):
    lib = _lib.load()
    for arr, need in chunks[0]:
        _lib.check(lib.negf_transmission_multi_batched(_lib.stream(), len(arr), arr,
                                                       ws.data_ptr(), need))
    for solver, _, _, _, _, out, _ in prepared:
        if solver.sets.share_leads:
            out.iters_r.copy_(out.iters_l)
            out.dos_r.copy_(out.dos_l)
    return [p[5] for p in prepared]


def transmission_multi(jobs, outs=None) -> list:
    """Several transmission problems of any block sizes solved together.

    ``jobs``: sequence of ``(solver, set_of, omega, eta_lead, eta_dev[, ladder])`` with
    a :class:`TransmissionSolver` per problem (its own sets, block sizes and leads)
    and that problem's instances.  Every instance of every problem goes through one
    persistent kernel launch per ``MULTI_MAX_PROBLEMS`` problems
    (``negf_transmission_multi_batched``); results as ``solver.run`` on each.  A
    mixed-precision solver's surface GFs come from the staged mixed decimation before
    that launch, which then runs the FP64 stages of every problem.  With an
    :class:`EtaLadder` (``eta_lead`` may then be None) each instance retries at its
    next eta inside the launch; ``result.ladder`` holds the accepted rung and every
    attempted rung's T and status.  ``outs`` (optional) are preallocated results.
    """
    prepared, chunks = _multi_prepare(jobs, outs)
    need = max((c[1] for c in chunks[0]), default=0)
    # a workspace per call (the caching allocator makes it cheap): calls from several
    # threads or streams never share one
    ws = torch.empty((max(need, 1),), dtype=torch.uint8, device=device())
    return _multi_launch(prepared, chunks, ws)


class MultiTransmissionGraph:
    """:func:`transmission_multi` captured as a hipGraph: the setup kernel, the
    persistent launch and the result copies of every problem.  Inputs are owned by
    the graph (``omega[j]``, ``eta_lead[j]``, ... per job; update in place), and so
    is its workspace; ``replay()`` solves all problems again.  Problems must fit the
    LDS classes (blocks up to 48 DOF): the 96-DOF pool orders its users through an
    event, which stream capture does not allow."""

    def __init__(self, jobs):
        self.solvers = [j[0] for j in jobs]
        self.set_of = [to_dev_i32(j[1]).clone() for j in jobs]
        self.omega = [to_dev_f64(j[2]).clone() for j in jobs]
        self.eta_lead = [to_dev_f64(j[3]).clone() for j in jobs]
        self.eta_dev = [(el if j[4] is None else to_dev_f64(j[4])).clone()
                        for j, el in zip(jobs, self.eta_lead)]
        self.out = [s.alloc_result(int(w.numel())) for s, w in zip(self.solvers, self.omega)]
        prepared, chunks = _multi_prepare(self._jobs(), self.out)
        need = max((c[1] for c in chunks[0]), default=0)
        self.workspace = torch.empty((max(need, 1),), dtype=torch.uint8, device=device())
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            _multi_launch(prepared, chunks, self.workspace)
        torch.cuda.current_stream().wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            _multi_launch(prepared, chunks, self.workspace)

    def _jobs(self):
        return list(zip(self.solvers, self.set_of, self.omega, self.eta_lead, self.eta_dev))

    def replay(self) -> list:
        self.graph.replay()
        return self.out


class TransmissionGraph:
    """One batched transmission solve captured as a hipGraph (``torch.cuda.CUDAGraph``).

    The graph holds the whole launch sequence of ``TransmissionSolver.run`` (both
    decimations, the self-energies, the sweep, the trace; in mixed mode also the
    refinement and FP64 redo launches, which exit early per instance).  Its inputs
    ``set_of``, ``omega``, ``eta_lead``, ``eta_dev`` are device tensors owned by the
    graph: write new values into them in place (``g.omega.copy_(...)``) and
    ``replay()`` solves the batch again with no host work and no launch overhead.
    ``out`` is overwritten by every replay.  The captured solve allocates nothing and
    records no memset node.
    """

    def __init__(self, solver: "TransmissionSolver", set_of, omega, eta_lead, eta_dev=None):
        self.solver = solver
        self.set_of = to_dev_i32(set_of).clone()
        self.omega = to_dev_f64(omega).clone()
        self.eta_lead = to_dev_f64(eta_lead).clone()
        self.eta_dev = (self.eta_lead if eta_dev is None else to_dev_f64(eta_dev)).clone()
        n_inst = int(self.omega.numel())
        self.out = solver.alloc_result(n_inst)
        # one eager run on a side stream sizes the workspace before capture
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            solver.run(self.set_of, self.omega, self.eta_lead, self.eta_dev, out=self.out)
        torch.cuda.current_stream().wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            solver.run(self.set_of, self.omega, self.eta_lead, self.eta_dev, out=self.out)

    def replay(self) -> "TransmissionResult":
        self.graph.replay()
        return self.out


def lead_self_energies(sets: TransportSets, set_of, omega, eta_lead, *, tol=1e-12, maxiter=200,
                       route=0, return_g=False):
    """FP64 decimation of both leads and Sigma = V g V^H embedded in the end layers.

    Returns (Sigma_L, Sigma_R [n_inst, n, n], lead status [n_inst]) on the GPU, and with
    ``return_g`` also the surface GFs (g_L, g_R).
    """
    lib = _lib.load()
    n_inst = int(omega.numel())
    dev = sets.onsite.device
    n = sets.n
    c128 = torch.complex128
    st = _lib.stream()
    sig = []
    stats = []
    gs = []
    for lead, v, idx in ((sets.lead_l, sets.v_l, sets.idx_l), (sets.lead_r, sets.v_r, sets.idx_r)):
        nld = int(lead.shape[2])
        g = torch.empty((n_inst, nld, nld), dtype=c128, device=dev)
        dos = torch.empty((n_inst,), dtype=torch.float64, device=dev)
        it = torch.zeros((n_inst,), dtype=torch.int32, device=dev)
        stt = torch.zeros((n_inst,), dtype=torch.int32, device=dev)
        _lib.check(lib.negf_sancho_rubio_batched(
            st, _lib.FP64, nld, n_inst, lead.data_ptr(), set_of.data_ptr(), omega.data_ptr(),
            eta_lead.data_ptr(), float(tol), int(maxiter), 0, g.data_ptr(), dos.data_ptr(),
            it.data_ptr(), stt.data_ptr()))
        if route == _lib.SURFACE_EIGEN:
            _lib.check(lib.negf_surface_gf_eigen_batched(
                st, nld, n_inst, lead.data_ptr(), set_of.data_ptr(), omega.data_ptr(),
                eta_lead.data_ptr(), g.data_ptr(), dos.data_ptr(), stt.data_ptr()))
        sg = torch.zeros((n_inst, n, n), dtype=c128, device=dev)
        _lib.check(lib.negf_self_energy_batched(
            st, nld, int(v.shape[1]), n, n_inst, v.data_ptr(), set_of.data_ptr(), g.data_ptr(),
            idx.data_ptr(), stt.data_ptr(), sg.data_ptr()))
        sig.append(sg)
        stats.append(stt)
        gs.append(g)
    lead_status = torch.where(stats[0] != 0, stats[0], stats[1])
    if return_g:
        return sig[0], sig[1], lead_status, (gs[0], gs[1])
    return sig[0], sig[1], lead_status


def device_green(sets: TransportSets, set_of, omega, eta_lead, eta_dev, *, tol=1e-12,
                 maxiter=200, route=0):
    """Full device Green's function of every instance (FP64).

    Returns (G [n_inst, Ln, Ln], Sigma_L, Sigma_R [n_inst, n, n] end-layer blocks,
    lead status, device status), all torch tensors on the GPU.
    """
    lib = _lib.load()
    set_of = to_dev_i32(set_of)
    omega = to_dev_f64(omega)
    eta_lead = to_dev_f64(eta_lead)
    eta_dev = to_dev_f64(eta_dev)
    n_inst = int(omega.numel())
    n, nl = sets.n, sets.n_layers
    sig_l, sig_r, lead_status = lead_self_energies(sets, set_of, omega, eta_lead, tol=tol,
                                                   maxiter=maxiter, route=route)
    dev = sets.onsite.device
    gfull = torch.zeros((n_inst, n * nl, n * nl), dtype=torch.complex128, device=dev)
    dst = torch.zeros((n_inst,), dtype=torch.int32, device=dev)
    _lib.check(lib.negf_device_green_batched(
        _lib.stream(), n, nl, n_inst, sets.onsite.data_ptr(), sets.coupling.data_ptr(),
        set_of.data_ptr(), omega.data_ptr(), eta_dev.data_ptr(), sig_l.data_ptr(),
        sig_r.data_ptr(), gfull.data_ptr(), dst.data_ptr()))
    return gfull, sig_l, sig_r, lead_status, dst


@dataclass
class AdjointResult:
    t: torch.Tensor              # (n_inst,) float64
    grad_onsite: torch.Tensor    # (n_inst or n_sets, L, n, n) complex128, weight * dT/dD_ii
    grad_coupling: torch.Tensor  # (n_inst or n_sets, max(L-1, 1), n, n), weight * dT/dC_i
    status: torch.Tensor         # (n_inst,) int32
    # with lead_grads: weight * dT/d(d00, d01, d10) (n_inst, 3, nl, nl) and weight * dT/dV
    # (n_inst, n_c, nl) of each lead
    grad_lead_l: torch.Tensor | None = None
    grad_lead_r: torch.Tensor | None = None
    grad_v_l: torch.Tensor | None = None
    grad_v_r: torch.Tensor | None = None


# doublings of the lead adjoint's Stein solve: the increments shrink like |lambda|^(2^k),
# |lambda| the slowest lead mode, which at eta = 1e-8 still leaves ~30 doublings
LEAD_ADJOINT_MAXITER = 64


def adjoint(sets: TransportSets, set_of, omega, eta_lead, eta_dev, weight=None, *, tol=1e-12,
            maxiter=200, route=0, lead_grads=False, checkpoint=0,
            per_set=False) -> AdjointResult:
    """T and weight * dT/d(device blocks) of every instance (FP64 adjoint kernel); with
    ``lead_grads`` also weight * dT/d(lead blocks) and weight * dT/dV of both leads
    (negf_lead_adjoint_batched).  A lead gradient whose Stein solve fails raises.

    ``checkpoint`` K > 0 keeps the left-connected sweep only every K layers and
    recomputes the rest per segment (less workspace, one more left sweep);
    ``per_set`` sums the device-block gradients into their sets on the GPU
    (grad_onsite / grad_coupling then have n_sets rows)."""
    lib = _lib.load()
    set_of = to_dev_i32(set_of)
    omega = to_dev_f64(omega)
    eta_lead = to_dev_f64(eta_lead)
    eta_dev = to_dev_f64(eta_dev)
    n_inst = int(omega.numel())
    n, nl = sets.n, sets.n_layers
    out = lead_self_energies(sets, set_of, omega, eta_lead, tol=tol, maxiter=maxiter, route=route,
                             return_g=lead_grads)
    sig_l, sig_r, lead_status = out[:3]
    dev = sets.onsite.device
    c128 = torch.complex128
    w = None if weight is None else to_dev_f64(weight)
    g1n = torch.empty((n_inst, n, n), dtype=c128, device=dev) if lead_grads else None
    ends = torch.empty((n_inst, 2, n, n), dtype=c128, device=dev) if lead_grads else None
    ws_bytes = int(lib.negf_adjoint_workspace_bytes(n, nl, n_inst, int(checkpoint)))
    ws = torch.empty((max(ws_bytes, 1),), dtype=torch.uint8, device=dev)
    t = torch.empty((n_inst,), dtype=torch.float64, device=dev)
    rows = sets.n_sets if per_set else n_inst
    g_ons = torch.zeros((rows, nl, n, n), dtype=c128, device=dev)
    g_cpl = torch.zeros((rows, max(nl - 1, 1), n, n), dtype=c128, device=dev)
    status = torch.zeros((n_inst,), dtype=torch.int32, device=dev)
    _lib.check(lib.negf_adjoint_batched(
        _lib.stream(), n, nl, n_inst, sets.onsite.data_ptr(), sets.coupling.data_ptr(),
        set_of.data_ptr(), omega.data_ptr(), eta_dev.data_ptr(), sig_l.data_ptr(),
        sig_r.data_ptr(), lead_status.data_ptr(), None if w is None else w.data_ptr(),
        ws.data_ptr(), ws_bytes, t.data_ptr(), g_ons.data_ptr(), g_cpl.data_ptr(),
        status.data_ptr(), None if g1n is None else g1n.data_ptr(), int(checkpoint),
        1 if per_set else 0, None if ends is None else ends.data_ptr()))
    res = AdjointResult(t=t, grad_onsite=g_ons, grad_coupling=g_cpl, status=status)
    if not lead_grads:
        return res
    del ws
    grads = []
    for side, lead, v, idx, g in ((0, sets.lead_l, sets.v_l, sets.idx_l, out[3][0]),
                                  (1, sets.lead_r, sets.v_r, sets.idx_r, out[3][1])):
        nld, nc = int(lead.shape[2]), int(v.shape[1])
        gl = torch.empty((n_inst, 3, nld, nld), dtype=c128, device=dev)
        gv = torch.empty((n_inst, nc, nld), dtype=c128, device=dev)
        st = torch.empty((n_inst,), dtype=torch.int32, device=dev)
        _lib.check(lib.negf_lead_adjoint_batched(
            _lib.stream(), side, n, n_inst, nld, nc, lead.data_ptr(), v.data_ptr(),
            idx.data_ptr(), set_of.data_ptr(), g.data_ptr(), g1n.data_ptr(), sig_l.data_ptr(),
            sig_r.data_ptr(), ends.data_ptr(), None if w is None else w.data_ptr(),
            status.data_ptr(), LEAD_ADJOINT_MAXITER, gl.data_ptr(), gv.data_ptr(),
            st.data_ptr()))
        bad = (st != 0) & (status == 0)
        if bool(bad.any()):
            i = int(torch.nonzero(bad)[0])
            raise RuntimeError(
                f"lead adjoint ({'left' if side == 0 else 'right'} lead) failed at instance {i} "
                f"(status {int(st[i])}: the Stein solve for dT/dg did not converge in "
                f"{LEAD_ADJOINT_MAXITER} doublings or met non-finite values)")
        grads.append((gl, gv))
    res.grad_lead_l, res.grad_v_l = grads[0]
    res.grad_lead_r, res.grad_v_r = grads[1]
    return res


# ---------------------------------------------------------------------------
# omega integrals
# ---------------------------------------------------------------------------

def thermal_conductance(omega_rad_s, tvals, temps) -> np.ndarray:
    lib = _lib.load()
    w = to_dev_f64(omega_rad_s)
    t = to_dev_f64(tvals)
    tk = to_dev_f64(temps)
    out = torch.empty((int(tk.numel()),), dtype=torch.float64, device=w.device)
    _lib.check(lib.negf_thermal_conductance(_lib.stream(), int(w.numel()), w.data_ptr(),
                                            t.data_ptr(), int(tk.numel()), tk.data_ptr(),
                                            out.data_ptr()))
    return out.cpu().numpy()


def heat_current(omega_rad_s, tvals, temp_left, temp_right) -> float:
    lib = _lib.load()
    w = to_dev_f64(omega_rad_s)
    t = to_dev_f64(tvals)
    out = torch.empty((1,), dtype=torch.float64, device=w.device)
    _lib.check(lib.negf_heat_current(_lib.stream(), int(w.numel()), w.data_ptr(), t.data_ptr(),
                                     float(temp_left), float(temp_right), out.data_ptr()))
    return float(out.cpu().item())


def thermal_conductance_tensor(omega_rad_s: torch.Tensor, tvals: torch.Tensor,
                               temps: torch.Tensor) -> torch.Tensor:
    """kappa(T_t) of a GPU spectrum, as a GPU tensor (the k_conductance kernel)."""
    lib = _lib.load()
    w, t, tk = to_dev_f64(omega_rad_s), to_dev_f64(tvals), to_dev_f64(temps)
    out = torch.empty((int(tk.numel()),), dtype=torch.float64, device=w.device)
    _lib.check(lib.negf_thermal_conductance(_lib.stream(), int(w.numel()), w.data_ptr(),
                                            t.data_ptr(), int(tk.numel()), tk.data_ptr(),
                                            out.data_ptr()))
    return out


def conductance_weights(omega_rad_s: torch.Tensor, temps: torch.Tensor) -> torch.Tensor:
    """(n_t, n_w) weights with kappa(T_t) = sum_i weights[t, i] T(w_i)."""
    lib = _lib.load()
    w, tk = to_dev_f64(omega_rad_s), to_dev_f64(temps)
    out = torch.empty((int(tk.numel()), int(w.numel())), dtype=torch.float64, device=w.device)
    _lib.check(lib.negf_conductance_weights(_lib.stream(), int(w.numel()), w.data_ptr(),
                                            int(tk.numel()), tk.data_ptr(), out.data_ptr()))
    return out
