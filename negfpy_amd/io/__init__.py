"""Force-constant file readers and their registry (reference ``src/negfpy/io``)."""

from __future__ import annotations

from collections.abc import Callable
from typing import Any

from .phonopy import read_phonopy_ifc
from .qe import read_qe_q2r_ifc

__all__ = ["get_reader", "list_readers", "read_ifc", "read_phonopy_ifc", "read_qe_q2r_ifc",
           "register_reader"]

_READERS: dict[str, Callable[[Any], Any]] = {}


def register_reader(name: str, reader) -> None:
    key = name.strip().lower()
    if not key:
        raise ValueError("Reader name must be non-empty.")
    _READERS[key] = reader


def get_reader(name: str):
    key = name.strip().lower()
    try:
        return _READERS[key]
    except KeyError as exc:
        available = ", ".join(sorted(_READERS)) or "<none>"
        raise KeyError(f"Unknown IFC reader '{name}'. Available readers: {available}") from exc


def list_readers() -> tuple[str, ...]:
    return tuple(sorted(_READERS))


def read_ifc(source: Any, reader: str = "qe"):
    return get_reader(reader)(source)


register_reader("qe", read_qe_q2r_ifc)
register_reader("phonopy", read_phonopy_ifc)
