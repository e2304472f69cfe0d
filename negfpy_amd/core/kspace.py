"""k-space blocks of force-constant models for many k points at once, on the GPU.

The reference evaluates lead and device blocks by a host callback per k point
(``src/negfpy/models/material_kspace.py:131-193``), re-summing every force-constant
term and re-weighting every layer for each k.  ``kspace_blocks`` builds all k
points' blocks in one launch of ``k_kspace_blocks`` (C entry ``negf_kspace_blocks``):
the phases e^{i (ky dy + kz dz)} and the 1/sqrt(m) weights are formed on the host
exactly as the reference forms them, the term sums and mass weighting run on the
GPU with separately rounded products and sums, so the blocks are the reference's
to the last bit.  The result feeds ``_engine.TransportSets`` without a host round
trip.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from negfpy_amd import _lib

from . import _engine

__all__ = ["KspaceBlocks", "kspace_blocks"]


@dataclass
class KspaceBlocks:
    lead: torch.Tensor | None       # (S, 3, n, n): d00, d01, d10 per k point
    onsite: torch.Tensor | None     # (S, L, n, n)
    coupling: torch.Tensor | None   # (S, L - 1, n, n)


def _kyz(kpar) -> tuple[float, float]:
    """The reference's (ky, kz) parsing, material_kspace.py:79-88."""
    if kpar is None or len(kpar) == 0:
        return 0.0, 0.0
    if len(kpar) > 2:
        raise ValueError("kpar must contain at most two transverse components: (ky,) or (ky, kz).")
    if len(kpar) == 1:
        return float(kpar[0]), 0.0
    return float(kpar[0]), float(kpar[1])


def _phases(terms: dict, kyz) -> np.ndarray:
    keys = list(terms.keys())
    ph = np.empty((len(kyz), len(keys)), dtype=np.complex128)
    for a, (ky, kz) in enumerate(kyz):
        for b, (dy, dz) in enumerate(keys):
            ph[a, b] = np.exp(1j * (ky * dy + kz * dz))
    return ph


def _stack_terms(terms: dict) -> np.ndarray:
    return np.stack([np.asarray(blk, dtype=np.complex128) for blk in terms.values()])


def _inv_sqrt_masses(masses, dof: int) -> np.ndarray:
    return 1.0 / np.sqrt(np.repeat(np.asarray(masses, dtype=float), dof))


def kspace_blocks(params, kpoints, *, lead: bool = True, n_layers: int | None = None,
                  layer_masses=None) -> KspaceBlocks:
    """Lead blocks (if ``lead``) and device blocks (if ``n_layers``) of a
    ``MaterialKspaceParams`` model at every k point of ``kpoints``, as GPU tensors.

    ``layer_masses`` (n_layers, n_atoms) sets per-layer device masses as in
    ``material_kspace_device``; by default every layer has the cell masses.
    """
    lib = _lib.load()
    kyz = [_kyz(k) for k in kpoints]
    nk = len(kyz)
    n = int(params.ndof)
    dof = int(params.dof_per_atom)
    if n_layers is not None:
        if n_layers <= 0:
            raise ValueError("n_layers must be positive.")
        if layer_masses is None:
            lm = np.repeat(np.asarray(params.masses, dtype=float)[None, :], n_layers, axis=0)
        else:
            lm = np.asarray(layer_masses, dtype=float)
            if lm.shape != (n_layers, params.n_atoms):
                raise ValueError("layer_masses shape must be (n_layers, n_atoms).")
            if np.any(lm <= 0.0):
                raise ValueError("All layer_masses entries must be positive.")
    if not lead and n_layers is None:
        raise ValueError("nothing to build: lead=False and no n_layers")
    dev = _engine.device()
    c128 = torch.complex128

    def up(x):
        return torch.from_numpy(np.ascontiguousarray(x)).to(dev)

    phi00 = up(_stack_terms(params.fc00_terms))
    phi01 = up(_stack_terms(params.fc01_terms))
    has10 = params.fc10_terms is not None
    phi10 = up(_stack_terms(params.fc10_terms)) if has10 else None
    ph00 = up(_phases(params.fc00_terms, kyz))
    ph01 = up(_phases(params.fc01_terms, kyz))
    ph10 = up(_phases(params.fc10_terms, kyz)) if has10 else None
    s_lead = up(_inv_sqrt_masses(params.masses, dof)) if lead else None
    nl = int(n_layers) if n_layers is not None else 0
    s_dev = up(np.stack([_inv_sqrt_masses(lm[i], dof) for i in range(nl)])) if nl else None
    lead_out = torch.empty((nk, 3, n, n), dtype=c128, device=dev) if lead else None
    ons_out = torch.empty((nk, nl, n, n), dtype=c128, device=dev) if nl else None
    cpl_out = torch.empty((nk, max(nl - 1, 0), n, n), dtype=c128, device=dev) if nl else None

    def ptr(t):
        return None if t is None or t.numel() == 0 else t.data_ptr()

    if nk:
        _lib.check(lib.negf_kspace_blocks(
            _lib.stream(), n, nl, nk, int(phi00.shape[0]), int(phi01.shape[0]),
            int(phi10.shape[0]) if has10 else 0, ptr(phi00), ptr(phi01), ptr(phi10), ptr(ph00),
            ptr(ph01), ptr(ph10), float(params.onsite_pinning), ptr(s_lead), ptr(s_dev),
            ptr(lead_out), ptr(ons_out), ptr(cpl_out)))
    return KspaceBlocks(lead=lead_out, onsite=ons_out, coupling=cpl_out)
