"""Force-constant file readers (reference ``src/negfpy/io``)."""

from .qe import read_qe_q2r_ifc

__all__ = ["read_ifc", "read_qe_q2r_ifc"]


def read_ifc(source, reader: str = "qe"):
    """Reader dispatch (reference ``io/registry.py:35``); QE q2r text files only."""
    name = str(reader).lower()
    if name in ("qe", "q2r", "qe_q2r"):
        return read_qe_q2r_ifc(source)
    raise NotImplementedError(f"IFC reader '{reader}' is not available (QE q2r text files only).")
