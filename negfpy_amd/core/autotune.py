"""Autotuner for the launch layout of each block size.

Every kernel exists in block-size classes (8, 16, 24, 32, 48 DOF, 64 in mixed
precision, 96 in FP64 with the matrices in a device-memory pool): the class fixes
the workgroup size, the rows each lane owns in the inversion, the product tile per
thread and the LDS per instance, and so the occupancy.  A block of n DOF
runs in any class >= n; which one is fastest depends on how the inversion's
per-step latency (rows per lane) trades against resident instances per CU.  Blocks of
up to 32 DOF can also run one instance per wavefront ("wave": 64-lane workgroups,
products on MFMA tiles, a register-resident wave inversion), which trades the
workgroup's barrier-bound step for more resident instances.  In FP64 the block
products of a class (8..48) have two tile shapes: v_mfma_f64_4x4x4f64 tiles on the
matrix cores (four per 8 x 8 tile-group of a wave) or a TM x TM register tile of FMA
loops per thread on the vector ALUs, the candidate ``(class, VECTOR)``.
``tune_block_class`` times every candidate -- the classes >= n in each product form and,
up to 32 DOF, the wave family -- on a synthetic batch of n-DOF instances (HIP events on
the solve's stream, best of ``repeats``), records the fastest in the library's
process-wide tables (``negf_set_block_class``, ``negf_set_kernel_family``,
``negf_set_product_form``) and returns the timings.  Results are cached per
(precision, n) for the process.
"""

from __future__ import annotations

import threading

import numpy as np
import torch

from negfpy_amd import _lib

from . import _engine
from .surface_gf import pack_lead
from .types import LeadBlocks

__all__ = ["CLASSES", "WAVE", "VECTOR", "candidate_classes", "tune_block_class", "block_class",
           "kernel_family", "product_form", "reset_block_class"]

CLASSES = (8, 16, 24, 32, 48, 64, 96)
WAVE = "wave"          # the one-wavefront family, a candidate for blocks up to 32 DOF
WAVE_MAX = 32
VECTOR = "vector"      # (class, VECTOR): the class with its products on the vector ALUs
FORM_MAX = 48          # FP64 classes with both product forms
_lock = threading.Lock()
_results: dict = {}


def candidate_classes(n: int, precision: str = "fp64", *, wave: bool = True,
                      vector: bool = True) -> list:
    """Workgroup classes that hold n-DOF blocks (matrix-core products in FP64), in FP64
    then ``(class, VECTOR)`` for the same classes with vector-ALU products, then ``WAVE``
    for n <= 32."""
    if precision == "fp64":
        # the pooled 96 class only for blocks the LDS classes cannot hold
        out = [c for c in (8, 16, 24, 32, 48) if n <= c] or ([96] if n <= 96 else [])
        if vector and 1 <= n <= FORM_MAX:
            out += [(c, VECTOR) for c in out]
    else:
        out = [c for c in (8, 16, 24, 32, 48, 64) if n <= c]
    if wave and 1 <= n <= WAVE_MAX and out:
        out.append(WAVE)
    return out


def kernel_family(n: int, precision: str = "fp64") -> str:
    """"wave" or "workgroup": the family n-DOF decimations use now."""
    f = int(_lib.load().negf_kernel_family(_engine.precision_code(precision), int(n)))
    return WAVE if f == _lib.FAMILY_WAVE else "workgroup"


def product_form(n: int) -> str:
    """"matrix_core" or "vector": the product form of FP64 n-DOF workgroup kernels now
    (n <= 48; larger blocks always use the matrix cores)."""
    if not 1 <= n <= FORM_MAX:
        return "matrix_core"
    f = int(_lib.load().negf_product_form(int(n)))
    return VECTOR if f == _lib.PRODUCTS_VECTOR else "matrix_core"


def block_class(n: int, precision: str = "fp64") -> int:
    """Class currently used for n-DOF blocks."""
    return int(_lib.load().negf_block_class(_engine.precision_code(precision), int(n)))


def reset_block_class(n: int, precision: str = "fp64") -> None:
    lib = _lib.load()
    code = _engine.precision_code(precision)
    _lib.check(lib.negf_set_block_class(code, int(n), 0))
    if 1 <= n <= WAVE_MAX:
        _lib.check(lib.negf_set_kernel_family(code, int(n), 0))
    if precision == "fp64" and 1 <= n <= FORM_MAX:
        _lib.check(lib.negf_set_product_form(int(n), 0))
    with _lock:
        _results.pop((precision, int(n)), None)


def _synthetic_sets(n: int, n_layers: int):
    from negfpy_amd.workloads import omega_max, random_lead

    d00, d01 = random_lead(1000 + n, n)
    rng = np.random.default_rng(n)
    onsite = np.stack([d00 + np.diag(rng.uniform(-0.05, 0.05, n)) for _ in range(n_layers)])
    coupling = np.stack([d01] * max(n_layers - 1, 1))
    lead = pack_lead(d00, d01)[None]
    sets = _engine.TransportSets.from_arrays(
        lead, lead, lead[:, 2], lead[:, 1], np.arange(n), np.arange(n), onsite[None],
        coupling[None] if n_layers > 1 else None, True)
    return sets, omega_max(LeadBlocks(d00=d00, d01=d01))


def tune_block_class(n: int, precision: str = "fp64", *, n_layers: int = 8, n_inst: int = 2048,
                     repeats: int = 3, candidates=None, apply: bool = True,
                     refresh: bool = False) -> dict:
    """Time each candidate (class, ``(class, VECTOR)`` or ``WAVE``) for n-DOF blocks and
    (if ``apply``) select the fastest.  A class candidate runs with the workgroup family
    (and, in FP64, matrix-core or vector-ALU products), ``WAVE`` with the one-wavefront
    family for the decimations and sweeps (the other stages keep the default class).

    Returns {"best": candidate, "timings_ms": {candidate: ms per batch},
    "n_inst": ..., "n_layers": ...}.
    """
    key = (precision, int(n))
    with _lock:
        if not refresh and key in _results:
            return _results[key]
    lib = _lib.load()
    code = _engine.precision_code(precision)
    cands = list(candidates) if candidates is not None else candidate_classes(n, precision)
    if not cands:
        raise ValueError(f"no block-size class holds {n}-DOF blocks in {precision}")
    sets, wmax = _synthetic_sets(int(n), int(n_layers))
    dev = sets.onsite.device
    omega = torch.as_tensor(np.linspace(0.05, 1.05, n_inst) * wmax, device=dev)
    eta = torch.full((n_inst,), 1e-6, dtype=torch.float64, device=dev)
    set_of = torch.zeros(n_inst, dtype=torch.int32, device=dev)
    solver = _engine.TransmissionSolver(sets, precision=precision)
    out = solver.alloc_result(n_inst)
    timings = {}
    previous = int(lib.negf_block_class(code, int(n)))
    fam_ok = 1 <= n <= WAVE_MAX
    # the families of both arithmetic tables: FP64 (decimation and sweep of FP64, the
    # FP64 stages of mixed) and FP32 (the mixed decimation)
    tables = (_lib.FP64,) if precision == "fp64" else (_lib.FP64, _lib.MIXED)
    prev_fam = {t: int(lib.negf_kernel_family(t, int(n))) for t in tables} if fam_ok else {}

    form_ok = 1 <= n <= FORM_MAX
    prev_form = int(lib.negf_product_form(int(n))) if form_ok else 0

    def family(f):
        if fam_ok:
            for t in tables:
                _lib.check(lib.negf_set_kernel_family(t, int(n), f))

    def select(c):
        if c == WAVE:
            _lib.check(lib.negf_set_block_class(code, int(n), 0))
            family(_lib.FAMILY_WAVE)
            form = 0
        else:
            vec = isinstance(c, tuple)
            _lib.check(lib.negf_set_block_class(code, int(n), int(c[0] if vec else c)))
            family(_lib.FAMILY_WORKGROUP)
            form = _lib.PRODUCTS_VECTOR if vec else _lib.PRODUCTS_MATRIX_CORE
        if form_ok:
            _lib.check(lib.negf_set_product_form(int(n), form))

    try:
        for c in cands:
            select(c)
            solver.run(set_of, omega, eta, out=out)
            best = float("inf")
            for _ in range(max(int(repeats), 1)):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                solver.run(set_of, omega, eta, out=out)
                e1.record()
                e1.synchronize()
                best = min(best, float(e0.elapsed_time(e1)))
            timings[c if (c == WAVE or isinstance(c, tuple)) else int(c)] = best
    finally:
        _lib.check(lib.negf_set_block_class(code, int(n), previous))
        for t, f in prev_fam.items():
            _lib.check(lib.negf_set_kernel_family(t, int(n), f))
        if form_ok:
            _lib.check(lib.negf_set_product_form(int(n), prev_form))
    choice = min(timings, key=timings.get)
    if apply:
        select(choice)
    res = {"best": choice, "timings_ms": timings, "n_inst": int(n_inst), "n_layers": int(n_layers)}
    with _lock:
        _results[key] = res
    return res
