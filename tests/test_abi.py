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

    # 11 ints, a double (8-aligned at offset 48), 18 pointers, then surface_route (padded)
    assert ctypes.sizeof(_lib.TransmissionProblem) == 48 + 8 + 18 * 8 + 8
    assert _lib.TransmissionProblem.surface_route.offset == 48 + 8 + 18 * 8
