"""ctypes binding of libnegf_hip.so (C ABI declared in include/negf_hip.h).

``import torch`` happens first so that the HIP runtime PyTorch ships is the one
the library binds to (both carry the soname libamdhip64.so.7); device pointers
and streams are then shared between torch tensors and the kernels.

There is no CPU fallback: :func:`load` raises if the library is not built or no
GPU is visible.
"""

from __future__ import annotations

import ctypes
import os
import threading

import torch  # noqa: F401  (must precede loading the HIP library)

HERE = os.path.dirname(os.path.abspath(__file__))
# NEGFPY_AMD_LIB selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("NEGFPY_AMD_LIB") or os.path.join(HERE, "libnegf_hip.so")
ABI_VERSION = 1

FP64 = 0
MIXED = 1
SURFACE_DECIMATION = 0
SURFACE_EIGEN = 1
FAMILY_WORKGROUP, FAMILY_WAVE = 1, 2   # NEGF_FAMILY_*
PRODUCTS_MATRIX_CORE, PRODUCTS_VECTOR = 1, 2  # NEGF_PRODUCTS_*
LEADS_INDEPENDENT = 0   # NEGF_LEADS_*: negf_transmission_problem.share_leads
LEADS_SHARED = 1
LEADS_MIRRORED = 2
MULTI_MAX_PROBLEMS = 8   # NEGF_MULTI_MAX_PROBLEMS
NOALLOC_SCRATCH_BYTES = 64 << 20   # NEGF_NOALLOC_SCRATCH_BYTES

OK = 0
ST_OK = 0
ST_NOCONV = 1
ST_NONFINITE = 2
ST_SINGULAR = 3
ST_EIGEN_ZERO_ETA = 4
ST_DEV_NONFINITE = 12
ST_DEV_SINGULAR = 13

_c_int = ctypes.c_int
_c_dbl = ctypes.c_double
_vp = ctypes.c_void_p


class EtaLadder(ctypes.Structure):
    """Mirror of ``negf_eta_ladder``."""

    _fields_ = [
        ("etas", _vp), ("first", _vp), ("count", _vp), ("t_max", _vp), ("rung_out", _vp),
        ("t_rungs", _vp), ("status_rungs", _vp), ("nonneg_tol", _c_dbl), ("max_rungs", _c_int),
        ("eta_dev_fixed", _c_int),
    ]


class TransmissionProblem(ctypes.Structure):
    """Mirror of ``negf_transmission_problem``."""

    _fields_ = [
        ("precision", _c_int), ("n", _c_int), ("n_layers", _c_int), ("n_inst", _c_int),
        ("n_lead_l", _c_int), ("n_lead_r", _c_int), ("n_c_l", _c_int), ("n_c_r", _c_int),
        ("share_leads", _c_int), ("maxiter", _c_int), ("refine_steps", _c_int),
        ("tol", _c_dbl),
        ("lead_l", _vp), ("lead_r", _vp), ("v_l", _vp), ("v_r", _vp),
        ("idx_l", _vp), ("idx_r", _vp), ("onsite", _vp), ("coupling", _vp),
        ("set_of", _vp), ("omega", _vp), ("eta_lead", _vp), ("eta_dev", _vp),
        ("t_out", _vp), ("dos_l", _vp), ("dos_r", _vp),
        ("iters_l", _vp), ("iters_r", _vp), ("status", _vp), ("surface_route", _c_int),
        ("ladder", ctypes.POINTER(EtaLadder)),
    ]


_SIGS = {
    "negf_version": ([], _c_int),
    "negf_last_error": ([], ctypes.c_char_p),
    "negf_max_block": ([_c_int], _c_int),
    "negf_max_block_green": ([], _c_int),
    "negf_set_block_class": ([_c_int, _c_int, _c_int], _c_int),
    "negf_block_class": ([_c_int, _c_int], _c_int),
    "negf_set_wave_kernels": ([_c_int], _c_int),
    "negf_set_matrix_core_products": ([_c_int], _c_int),
    "negf_set_kernel_family": ([_c_int, _c_int, _c_int], _c_int),
    "negf_kernel_family": ([_c_int, _c_int], _c_int),
    "negf_set_product_form": ([_c_int, _c_int], _c_int),
    "negf_product_form": ([_c_int], _c_int),
    "negf_sancho_rubio_batched": (
        [_vp, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _c_dbl, _c_int, _c_int, _vp, _vp, _vp, _vp],
        _c_int),
    "negf_surface_gf_eigen_batched": (
        [_vp, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _c_int),
    "negf_surface_refine_batched": (
        [_vp, _c_int, _c_int, _vp, _vp, _vp, _vp, _c_int, _vp, _vp, _vp], _c_int),
    "negf_sancho_rubio_redo": (
        [_vp, _c_int, _c_int, _vp, _vp, _vp, _vp, _c_dbl, _c_int, _vp, _vp, _vp, _vp], _c_int),
    "negf_self_energy_batched": (
        [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp], _c_int),
    "negf_rgf_g1n_batched": (
        [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
        _c_int),
    "negf_device_green_batched": (
        [_vp, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _c_int),
    "negf_adjoint_workspace_bytes": ([_c_int, _c_int, _c_int, _c_int], ctypes.c_size_t),
    "negf_adjoint_batched": (
        [_vp, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
         ctypes.c_size_t, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _vp], _c_int),
    "negf_lead_adjoint_batched": (
        [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
         _vp, _vp, _vp, _c_int, _vp, _vp, _vp], _c_int),
    "negf_transmission_trace_batched": ([_vp, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp], _c_int),
    "negf_merge_status": ([_vp, _c_int, _vp, _vp], _c_int),
    "negf_transmission_sweep_batched": (
        [_vp, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _vp, _vp, _vp, _vp,
         _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _c_int),
    "negf_transmission_workspace_bytes": ([ctypes.POINTER(TransmissionProblem)], ctypes.c_size_t),
    "negf_transmission_multi_workspace_bytes": ([_c_int, ctypes.POINTER(TransmissionProblem)],
                                                ctypes.c_size_t),
    "negf_transmission_multi_batched": (
        [_vp, _c_int, ctypes.POINTER(TransmissionProblem), _vp, ctypes.c_size_t], _c_int),
    "negf_transmission_batched": (
        [_vp, ctypes.POINTER(TransmissionProblem), _vp, ctypes.c_size_t], _c_int),
    "negf_transmission_batched_noalloc": ([_vp, ctypes.POINTER(TransmissionProblem)], _c_int),
    "negf_thermal_conductance": ([_vp, _c_int, _vp, _vp, _c_int, _vp, _vp], _c_int),
    "negf_heat_current": ([_vp, _c_int, _vp, _vp, _c_dbl, _c_dbl, _vp], _c_int),
    "negf_conductance_weights": ([_vp, _c_int, _vp, _c_int, _vp, _vp], _c_int),
    "negf_fcs_cumulants": ([_vp, _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp], _c_int),
    "negf_kspace_blocks": (
        [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _c_dbl,
         _vp, _vp, _vp, _vp, _vp], _c_int),
}

EXPORTS = tuple(_SIGS)

_lock = threading.Lock()
_lib = None


class NegfLibraryError(RuntimeError):
    """Raised when the HIP library fails (bad arguments or a HIP runtime error)."""


def load_symbols_only(path: str = LIB_PATH) -> ctypes.CDLL:
    """Open the library and bind every exported symbol without touching a GPU."""
    if not os.path.exists(path):
        raise NegfLibraryError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(path)
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if lib.negf_version() != ABI_VERSION:
        raise NegfLibraryError(f"ABI version mismatch: {lib.negf_version()} != {ABI_VERSION}")
    return lib


def load() -> ctypes.CDLL:
    """Return the bound library; requires a visible ROCm GPU."""
    global _lib
    with _lock:
        if _lib is None:
            if not torch.cuda.is_available():
                raise NegfLibraryError(
                    "negfpy_amd runs on an AMD Instinct GPU (gfx950) and none is visible; "
                    "there is no CPU fallback")
            _lib = load_symbols_only()
    return _lib


def check(rc: int) -> None:
    if rc != OK:
        msg = _lib.negf_last_error().decode() if _lib is not None else "library not loaded"
        raise NegfLibraryError(f"libnegf_hip error {rc}: {msg}")


def ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def stream() -> int:
    return torch.cuda.current_stream().cuda_stream
