"""The dense autograd reference used by the GPU adjoint tests, pinned on the CPU.

Its T equals the oracle's sparse-LU transmission, and its gradient equals central
finite differences of the oracle along random directions.  The reference package
has no gradients, so finite differences of the oracle (itself pinned against the
reference's own vectors) are what fixes the adjoint.
"""

import numpy as np
import pytest
import torch

from adjoint_util import dense_grads, dense_transmission
from oracle import negf_oracle as o
from test_reference_vectors import REF


def _md():
    return (REF["md_onsite"], REF["md_coupling"], (REF["md_d00"], REF["md_d01"], None))


CASES = {
    "plain": dict(eta=1e-6),
    "scaled": dict(eta=1e-6, omega_scale=1.7),
    "contacts": dict(eta=1e-6, v_left=REF["md_vl"], v_right=REF["md_vr"],
                     idx_left=[0, 2, 4], idx_right=[25, 27, 29]),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_dense_reference_matches_oracle_and_fd(case):
    ons, cps, lead = _md()
    kw = CASES[case]
    w = 0.45 * float(REF["md_omegas"][-1] / 1.15)
    t, g_ons, g_cpl = dense_grads(w, ons, cps, lead, lead, **kw)
    t_or = o.transmission_sparse(w, list(ons), list(cps), lead, lead, **kw)
    assert abs(t - t_or) <= 1e-11 * max(abs(t_or), 1e-3)
    rng = np.random.default_rng(5)
    for _ in range(3):
        d_ons = rng.standard_normal(ons.shape) + 1j * rng.standard_normal(ons.shape)
        d_cpl = rng.standard_normal(cps.shape) + 1j * rng.standard_normal(cps.shape)
        h = 1e-6 * np.max(np.abs(ons))
        tp = o.transmission_sparse(w, list(ons + h * d_ons), list(cps + h * d_cpl), lead, lead, **kw)
        tm = o.transmission_sparse(w, list(ons - h * d_ons), list(cps - h * d_cpl), lead, lead, **kw)
        fd = (tp - tm) / (2 * h)
        an = np.sum(np.real(np.conj(g_ons) * d_ons)) + np.sum(np.real(np.conj(g_cpl) * d_cpl))
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-3), (fd, an)


def test_dense_reference_single_layer():
    ons, _, lead = _md()
    w = 0.3 * float(REF["md_omegas"][-1] / 1.15)
    t, g_ons, g_cpl = dense_grads(w, ons[:1], np.zeros((0, 6, 6)), lead, lead, eta=1e-6)
    assert abs(t - o.transmission_sparse(w, [ons[0]], [], lead, lead, eta=1e-6)) < 1e-11
    assert g_cpl.shape == (0, 6, 6) and np.all(np.isfinite(g_ons))


def test_dense_reference_detects_a_dropped_conjugate_coupling():
    """Planted mistake: a device matrix without the C^H lower blocks changes T, so the
    reference is sensitive to how coupling enters D."""
    ons, cps, lead = _md()
    w = 0.45 * float(REF["md_omegas"][-1] / 1.15)
    t_ok = dense_transmission(w, torch.as_tensor(ons), torch.as_tensor(cps), lead, lead, eta=1e-6)
    t_bad = o.transmission_sparse(w, list(ons), [c.T for c in cps], lead, lead, eta=1e-6)
    assert abs(float(t_ok) - t_bad) > 1e-6
