"""The C-ABI library loads without a GPU and exports every function declared in
include/negf_hip.h; the ctypes binding covers exactly those symbols."""

import os
import re

from negfpy_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "negf_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(negf_[a-z0-9_]+)\s*\(", text))


def test_header_and_binding_agree():
    assert declared_functions() == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = _lib.load_symbols_only()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.negf_version() == _lib.ABI_VERSION
    assert lib.negf_max_block(_lib.FP64) == 96
    assert lib.negf_max_block(_lib.MIXED) == 48
    assert lib.negf_max_block_green() == 96


def test_problem_struct_layout():
    import ctypes

    # 11 ints, a double (8-aligned at offset 48), 18 pointers, surface_route (padded),
    # the ladder pointer
    assert ctypes.sizeof(_lib.TransmissionProblem) == 48 + 8 + 18 * 8 + 8 + 8
    assert _lib.TransmissionProblem.surface_route.offset == 48 + 8 + 18 * 8
    assert _lib.TransmissionProblem.ladder.offset == 48 + 8 + 18 * 8 + 8
    # 7 pointers, a double, two ints
    assert ctypes.sizeof(_lib.EtaLadder) == 7 * 8 + 8 + 8
    assert _lib.EtaLadder.max_rungs.offset == 64


def test_multi_problem_entry_rejects_bad_arguments():
    """Argument checks of negf_transmission_multi_batched (and the lead-sharing mode of
    negf_transmission_batched), which return before any HIP call (so they run
    without a GPU)."""
    import ctypes

    lib = _lib.load_symbols_only()
    arr = (_lib.TransmissionProblem * 9)()
    for p in arr:
        p.precision, p.n, p.n_layers, p.n_inst = _lib.FP64, 4, 2, 1
        p.n_lead_l = p.n_lead_r = p.n_c_l = p.n_c_r = 4
        p.maxiter = 10
    assert lib.negf_transmission_multi_batched(None, 0, arr, None, 0) == _lib.OK
    assert lib.negf_transmission_multi_batched(None, 9, arr, None, 0) == 1
    assert b"at most 8" in lib.negf_last_error()
    arr[0].precision = 5
    assert lib.negf_transmission_multi_batched(None, 2, arr, None, 0) == 1
    assert b"problem 0" in lib.negf_last_error() and b"precision" in lib.negf_last_error()
    arr[0].precision = _lib.MIXED
    arr[0].refine_steps = -1
    assert lib.negf_transmission_multi_batched(None, 2, arr, None, 0) == 1
    assert b"refine_steps" in lib.negf_last_error()
    arr[0].refine_steps = 2
    arr[0].n = arr[0].n_lead_l = arr[0].n_lead_r = 60
    assert lib.negf_transmission_multi_batched(None, 1, arr, None, 0) == 1
    assert b"mixed-precision limit" in lib.negf_last_error()
    arr[0].n = arr[0].n_lead_l = arr[0].n_lead_r = 4
    lad = _lib.EtaLadder()
    arr[0].ladder = ctypes.pointer(lad)
    assert lib.negf_transmission_multi_batched(None, 1, arr, None, 0) == 1
    assert b"eta ladders run in FP64" in lib.negf_last_error()
    arr[0].ladder = None
    arr[0].precision = _lib.FP64
    arr[0].n = 97
    assert lib.negf_transmission_multi_batched(None, 1, arr, None, 0) == 1
    assert b"exceeds 96" in lib.negf_last_error()
    arr[0].n = 4
    arr[0].n_c_l = 5
    assert lib.negf_transmission_multi_batched(None, 1, arr, None, 0) == 1
    assert b"invalid sizes" in lib.negf_last_error()
    arr[0].n_c_l = 4
    assert lib.negf_transmission_multi_batched(None, 1, arr, None, 0) == 1  # null pointers
    assert b"null pointer" in lib.negf_last_error()
    arr[0].share_leads = 3  # not NEGF_LEADS_INDEPENDENT, _SHARED or _MIRRORED
    assert lib.negf_transmission_multi_batched(None, 1, arr, None, 0) == 1
    assert b"share_leads 3" in lib.negf_last_error()
    need = lib.negf_transmission_workspace_bytes(ctypes.byref(arr[0]))
    # the staged entry checks it before any HIP call too (the workspace is not touched)
    assert lib.negf_transmission_batched(None, ctypes.byref(arr[0]), ctypes.c_void_p(256),
                                         need) == 1
    assert b"share_leads 3" in lib.negf_last_error()
    arr[0].share_leads = _lib.LEADS_INDEPENDENT
    assert lib.negf_transmission_multi_workspace_bytes(0, arr) == 0
    assert lib.negf_transmission_multi_workspace_bytes(1, arr) > \
        lib.negf_transmission_workspace_bytes(ctypes.byref(arr[0]))


def test_wave_kernel_mode_switch():
    """negf_set_wave_kernels takes modes 0, 1, 2 and returns the previous one; other
    values are rejected with -1 (host-only, no GPU work)."""
    lib = _lib.load_symbols_only()
    prev = lib.negf_set_wave_kernels(2)
    assert prev in (0, 1, 2)
    assert lib.negf_set_wave_kernels(0) == 2
    assert lib.negf_set_wave_kernels(7) == -1
    assert b"mode" in lib.negf_last_error()
    assert lib.negf_set_wave_kernels(prev) == 0


def test_matrix_core_product_switch():
    """negf_set_matrix_core_products takes modes 0 and 1 and returns the previous one;
    other values are rejected with -1 (host-only, no GPU work)."""
    lib = _lib.load_symbols_only()
    prev = lib.negf_set_matrix_core_products(0)
    assert prev in (0, 1)
    assert lib.negf_set_matrix_core_products(1) == 0
    assert lib.negf_set_matrix_core_products(2) == -1
    assert b"matrix-core" in lib.negf_last_error()
    assert lib.negf_set_matrix_core_products(prev) == 1


def test_product_form_table():
    """negf_set_product_form stores a per-size product form (0 default, 1 matrix cores,
    2 vector ALUs) for FP64 blocks of 1..48 DOF; negf_product_form reports the one in
    force (0 outside that range); matrix-core mode 0 overrides the table (host-only)."""
    lib = _lib.load_symbols_only()
    prev = lib.negf_set_matrix_core_products(1)
    try:
        for n in (1, 6, 24, 48):
            assert lib.negf_product_form(n) == _lib.PRODUCTS_MATRIX_CORE
        assert lib.negf_product_form(0) == lib.negf_product_form(60) == 0
        assert lib.negf_set_product_form(24, _lib.PRODUCTS_VECTOR) == 0
        assert lib.negf_product_form(24) == _lib.PRODUCTS_VECTOR
        assert lib.negf_product_form(23) == _lib.PRODUCTS_MATRIX_CORE
        assert lib.negf_set_product_form(24, _lib.PRODUCTS_MATRIX_CORE) == 0
        assert lib.negf_product_form(24) == _lib.PRODUCTS_MATRIX_CORE
        lib.negf_set_matrix_core_products(0)
        assert lib.negf_product_form(24) == _lib.PRODUCTS_VECTOR
        for args, word in (((49, 1), b"1..48"), ((0, 1), b"1..48"), ((24, 3), b"form")):
            assert lib.negf_set_product_form(*args) == 1
            assert word in lib.negf_last_error()
    finally:
        lib.negf_set_product_form(24, 0)
        lib.negf_set_matrix_core_products(prev)


def test_kernel_family_table():
    """negf_set_kernel_family stores a per-size family (0 default, 1 workgroup, 2 wave)
    for blocks of 1..32 DOF and negf_kernel_family reports the one in force in mode 1;
    bad arguments are rejected (host-only, no GPU work)."""
    lib = _lib.load_symbols_only()
    prev = lib.negf_set_wave_kernels(1)
    try:
        for prec in (_lib.FP64, _lib.MIXED):                                  # default rule
            for n in (1, 8, 12, 20, 24, 32):
                assert lib.negf_kernel_family(prec, n) == _lib.FAMILY_WORKGROUP
        assert lib.negf_set_kernel_family(_lib.FP64, 24, _lib.FAMILY_WAVE) == 0
        assert lib.negf_kernel_family(_lib.FP64, 24) == _lib.FAMILY_WAVE
        assert lib.negf_set_kernel_family(_lib.FP64, 12, _lib.FAMILY_WORKGROUP) == 0
        assert lib.negf_kernel_family(_lib.FP64, 12) == _lib.FAMILY_WORKGROUP
        assert lib.negf_set_kernel_family(_lib.MIXED, 20, _lib.FAMILY_WAVE) == 0  # FP32 side
        assert lib.negf_kernel_family(_lib.MIXED, 20) == _lib.FAMILY_WAVE
        assert lib.negf_kernel_family(_lib.FP64, 20) == _lib.FAMILY_WORKGROUP
        assert lib.negf_set_wave_kernels(2) == 1  # mode 2: always, up to 32 DOF
        assert lib.negf_kernel_family(_lib.FP64, 32) == _lib.FAMILY_WAVE
        assert lib.negf_kernel_family(_lib.FP64, 33) == _lib.FAMILY_WORKGROUP
        lib.negf_set_wave_kernels(0)  # mode 0: never, whatever the table says
        assert lib.negf_kernel_family(_lib.FP64, 24) == _lib.FAMILY_WORKGROUP
        for args, word in (((_lib.FP64, 33, 1), b"1..32"), ((_lib.FP64, 0, 1), b"1..32"),
                           ((_lib.FP64, 12, 3), b"family"), ((5, 12, 1), b"precision")):
            assert lib.negf_set_kernel_family(*args) == 1
            assert word in lib.negf_last_error()
    finally:
        for n in (12, 24):
            lib.negf_set_kernel_family(_lib.FP64, n, 0)
        lib.negf_set_kernel_family(_lib.MIXED, 20, 0)
        lib.negf_set_wave_kernels(prev)


def test_noalloc_entry_rejects_without_touching_a_gpu():
    """The allocation-free entry point validates its problem before any HIP call."""
    import ctypes

    lib = _lib.load_symbols_only()
    assert lib.negf_transmission_batched_noalloc(None, None) == 1
    p = _lib.TransmissionProblem()
    p.n = p.n_lead_l = p.n_lead_r = p.n_c_l = p.n_c_r = 24
    p.n_layers = 32
    p.n_inst = 1 << 20          # far beyond the 64 MiB scratch
    assert lib.negf_transmission_batched_noalloc(None, ctypes.byref(p)) == 1
    assert b"scratch" in lib.negf_last_error()
    p.n_inst = 4
    p.n = p.n_lead_l = p.n_lead_r = p.n_c_l = p.n_c_r = 60   # pooled 96 class
    assert lib.negf_transmission_batched_noalloc(None, ctypes.byref(p)) == 1
    assert b"pool" in lib.negf_last_error()
