"""Left/right leads and device-to-lead couplings of a bulk IFC transport model
(reference ``src/negfpy/workflows/ifc_bulk.py:683-742``).

With Phi10 given (or Phi10(dy, dz) = Phi01(-dy, -dz)^H), the right lead's
inter-layer block is K_R01(dy, dz) = Phi10(-dy, -dz)^T and the left lead's
K_L01 = Phi10; both leads share the same-layer block Phi00 (+ pinning) and the
cell masses.  The right device-to-lead coupling is the right lead's d01 and the
left one its Hermitian conjugate.

The builders carry tags that let the batched k-point solvers build every k point
of these leads and couplings on the GPU (``negfpy_amd.core.kspace``): each lead is
the force-constant lead of derived terms (Phi00, K_01) whose d10 is d01^H, and
each coupling names the right lead's d01 (or its conjugate transpose).
"""

from __future__ import annotations

import numpy as np

from negfpy_amd.core.types import LeadKSpace
from negfpy_amd.models import MaterialKspaceParams


def _kyz(kpar):
    if kpar is None or len(kpar) == 0:
        return 0.0, 0.0
    if len(kpar) > 2:
        raise ValueError("kpar must contain at most two transverse components: (ky,) or (ky, kz).")
    if len(kpar) == 1:
        return float(kpar[0]), 0.0
    return float(kpar[0]), float(kpar[1])


def _fourier(terms, ky, kz, nd):
    out = np.zeros((nd, nd), dtype=np.complex128)
    for (dy, dz), blk in terms.items():
        out += np.asarray(blk, dtype=np.complex128) * np.exp(1j * (ky * dy + kz * dz))
    return out


def interface_contacts(params):
    """(lead_left, lead_right, device_to_lead_left, device_to_lead_right) of a bulk
    force-constant model (``_build_interface_contacts``, ifc_bulk.py:690-742)."""
    f00 = {k: np.asarray(v, dtype=np.complex128) for k, v in params.fc00_terms.items()}
    if params.fc10_terms is None:
        f10 = {(-dy, -dz): np.asarray(v, dtype=np.complex128).conj().T
               for (dy, dz), v in params.fc01_terms.items()}
    else:
        f10 = {k: np.asarray(v, dtype=np.complex128) for k, v in params.fc10_terms.items()}
    kr01 = {(-dy, -dz): np.asarray(v, dtype=np.complex128).T for (dy, dz), v in f10.items()}
    kl01 = {k: np.asarray(v, dtype=np.complex128) for k, v in f10.items()}
    s = 1.0 / np.sqrt(np.repeat(np.asarray(params.masses, dtype=float), int(params.dof_per_atom)))
    nd = int(params.ndof)
    pin = float(params.onsite_pinning)

    def weigh(phi):
        return s[:, None] * phi * s[None, :]

    def d00(ky, kz):
        phi = _fourier(f00, ky, kz, nd)
        if pin != 0.0:
            phi = phi + pin * np.eye(nd, dtype=np.complex128)
        return weigh(phi)

    def right(kpar):
        ky, kz = _kyz(kpar)
        d01 = weigh(_fourier(kr01, ky, kz, nd))
        return d00(ky, kz), d01, d01.conj().T

    def left(kpar):
        ky, kz = _kyz(kpar)
        d01 = weigh(_fourier(kl01, ky, kz, nd))
        return d00(ky, kz), d01, d01.conj().T

    def v_right(kpar):
        ky, kz = _kyz(kpar)
        return weigh(_fourier(kr01, ky, kz, nd))

    def v_left(kpar):
        return v_right(kpar).conj().T

    def derived(k01):
        return MaterialKspaceParams(masses=np.asarray(params.masses, dtype=float), fc00_terms=f00,
                                    fc01_terms=k01, fc10_terms=None,
                                    dof_per_atom=int(params.dof_per_atom), onsite_pinning=pin)

    left.material_kspace, right.material_kspace = derived(kl01), derived(kr01)
    left.d10_is_d01_h = right.d10_is_d01_h = True
    v_right.material_kspace_contact = ("d01", right.material_kspace)
    v_left.material_kspace_contact = ("d01_h", right.material_kspace)
    return LeadKSpace(blocks_builder=left), LeadKSpace(blocks_builder=right), v_left, v_right


def build_interface_contacts(interface, ndof: int):
    """Validated (left_contact, right_contact) couplings of an ``InterfaceData``, each
    (ndof, ndof) complex or None (reference ``modeling/builders/interface_builder.py:13``)."""
    def check(mat, side):
        if mat is None:
            return None
        arr = np.asarray(mat, dtype=np.complex128)
        if arr.shape != (ndof, ndof):
            raise ValueError(f"{side}_contact must have shape (ndof, ndof).")
        return arr

    return check(interface.left_contact, "left"), check(interface.right_contact, "right")
