"""Batched adaptive-eta spectrum driver on the GPU path.

Each eta rung is one batched solve over every pending (omega, k) instance of the
spectrum; omega by omega the result must be what the per-omega drivers give (which
the reference-made vectors pin), exactly, since every instance is solved by its own
workgroup whatever the batch around it.
"""

import numpy as np
import pytest

from negfpy_amd.core import (
    transmission_kavg_adaptive,
    transmission_kavg_adaptive_global_eta,
    transmission_kavg_adaptive_spectrum,
)
from negfpy_amd.models import MaterialKspaceParams, material_kspace_device, material_kspace_lead
from test_reference_vectors import (ADAPTIVE, MK_K, REF, TOL, _sq_problem, assert_info_matches,
                                    material_params)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scheme", ["adaptive", "adaptive_global"])
def test_spectrum_equals_per_omega_drivers(gpu, scheme):
    dev, lead, kpts = _sq_problem()
    ws = [0.9, 1.1, 0.0, 1.6, -0.3, 0.4, 1.25]
    kw = dict(eta_values=(1e-8, 1e-7), nonnegative_tolerance=1e-6, max_channel_factor=0.5,
              min_success_fraction=0.2, collect_rejected=True)
    vals, recs = transmission_kavg_adaptive_spectrum(ws, dev, lead, lead, kpts,
                                                     eta_scheme=scheme, **kw)
    single = (transmission_kavg_adaptive if scheme == "adaptive"
              else transmission_kavg_adaptive_global_eta)
    for w, v, rec in zip(ws, vals, recs):
        if w <= 0.0:
            assert v == 0.0 and rec is None
            continue
        try:
            t, info = single(w, dev, lead, lead, kpoints=kpts, **kw)
        except RuntimeError as exc:
            assert np.isnan(v) and rec == {"error": str(exc)}
            continue
        assert v == t
        assert rec["info"] == info
    if scheme == "adaptive_global":
        ref = ADAPTIVE["sq_global_eta"]
        assert abs(vals[1] - ref["t"]) <= TOL * max(abs(ref["t"]), 1e-3)
        assert_info_matches(recs[1]["info"], ref["info"])


def test_spectrum_reference_default_case(gpu):
    dev, lead, kpts = _sq_problem()
    ref = ADAPTIVE["sq_default"]
    vals, recs = transmission_kavg_adaptive_spectrum(
        [ref["omega"], 0.5], dev, lead, lead, kpts, eta_values=(1e-8, 1e-6),
        nonnegative_tolerance=1e-6, collect_rejected=True)
    assert abs(vals[0] - ref["t"]) <= TOL * max(abs(ref["t"]), 1e-3)
    assert_info_matches(recs[0]["info"], ref["info"])


def test_spectrum_on_gpu_built_material_blocks(gpu):
    """Force-constant model: blocks built on the GPU once for all k, every rung batched
    over all omegas; per-omega ladders."""
    mp = material_params(MaterialKspaceParams)
    ml = material_kspace_lead(mp)
    md = material_kspace_device(4, mp, layer_masses=REF["mk_layer_masses"])
    ws = list(REF["mk_omegas"]) + [0.25]
    ladders = [(1e-6,), (1e-8, 1e-6), (1e-6, 1e-5), (1e-7,)]
    vals, recs = transmission_kavg_adaptive_spectrum(ws, md, ml, ml, MK_K, eta_values=ladders,
                                                     max_channel_factor=1.5)
    for w, lad, v, rec in zip(ws, ladders, vals, recs):
        t, info = transmission_kavg_adaptive(w, md, ml, ml, kpoints=MK_K, eta_values=lad,
                                             max_channel_factor=1.5)
        assert v == t and rec["info"] == info
    # the first omega's ladder is the reference's single eta = 1e-6 kavg
    assert abs(vals[0] - REF["mk_t_kavg"][0]) <= 1e-9 * max(abs(REF["mk_t_kavg"][0]), 1e-3)
