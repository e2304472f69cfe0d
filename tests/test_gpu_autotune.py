"""Autotuner of the block-size class and kernel family: timings, selection, and results
that do not depend on the class or family a block runs in."""

import numpy as np
import pytest

from negfpy_amd.core import LeadBlocks, transmission_spectrum
from negfpy_amd.core.autotune import (
    VECTOR,
    WAVE,
    block_class,
    candidate_classes,
    kernel_family,
    product_form,
    reset_block_class,
    tune_block_class,
)
from negfpy_amd.core.types import Device1D
from negfpy_amd.workloads import omega_max, random_lead
from oracle import negf_oracle as o
from parity_util import assert_parity, oracle_t

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore_classes():
    yield
    for prec, ns in (("fp64", (12, 20)), ("mixed", (12, 20))):
        for n in ns:
            reset_block_class(n, prec)


def _applied(n, prec, best):
    if best == WAVE:
        return kernel_family(n, prec) == WAVE
    if isinstance(best, tuple):
        return (kernel_family(n, prec) == "workgroup" and block_class(n, prec) == best[0]
                and product_form(n) == VECTOR)
    return (kernel_family(n, prec) == "workgroup" and block_class(n, prec) == best
            and (prec != "fp64" or product_form(n) == "matrix_core"))


def test_tuner_times_every_candidate_and_applies_the_fastest(gpu, restore_classes):
    res = tune_block_class(12, "fp64", n_layers=4, n_inst=512, repeats=2)
    assert set(res["timings_ms"]) == set(candidate_classes(12, "fp64")) == {
        16, 24, 32, 48, (16, VECTOR), (24, VECTOR), (32, VECTOR), (48, VECTOR), WAVE}
    assert all(t > 0 for t in res["timings_ms"].values())
    assert res["best"] == min(res["timings_ms"], key=res["timings_ms"].get)
    assert _applied(12, "fp64", res["best"])
    assert tune_block_class(12, "fp64") is res  # cached
    res_m = tune_block_class(20, "mixed", n_layers=4, n_inst=512, repeats=2)
    assert set(res_m["timings_ms"]) == {24, 32, 48, 64, WAVE}
    assert _applied(20, "mixed", res_m["best"])
    print(f"tuned: fp64 n=12 {res['timings_ms']} -> {res['best']}; mixed n=20 "
          f"{res_m['timings_ms']} -> {res_m['best']}", flush=True)
    reset_block_class(12, "fp64")
    assert block_class(12, "fp64") == 16 and kernel_family(12, "fp64") == "workgroup"  # default
    assert product_form(12) == "matrix_core"
    assert 96 not in candidate_classes(40, "fp64") and WAVE not in candidate_classes(40, "fp64")


def test_results_do_not_depend_on_the_class(gpu, restore_classes):
    from negfpy_amd import _lib

    lib = _lib.load()
    d00, d01 = random_lead(77, 12)
    rng = np.random.default_rng(3)
    ons = [d00 + np.diag(rng.uniform(-0.1, 0.1, 12)) for _ in range(5)]
    cps = [d01] * 4
    lead = LeadBlocks(d00=d00, d01=d01)
    dev = Device1D(onsite_blocks=ons, coupling_blocks=cps)
    w = np.linspace(0.05, 1.05, 24) * omega_max(lead)
    refs = [oracle_t(x, ons, cps, (d00, d01, None), (d00, d01, None), eta=1e-6) for x in w]
    got = {}
    assert lib.negf_set_kernel_family(0, 12, _lib.FAMILY_WORKGROUP) == 0  # the classes proper
    for c in (16, 24, 32, 48):
        assert lib.negf_set_block_class(0, 12, c) == 0
        got[c] = transmission_spectrum(w, dev, lead, lead, eta=1e-6)
        assert_parity(got[c], refs)
    for c in (24, 32, 48):
        np.testing.assert_allclose(got[c], got[16], rtol=1e-11, atol=1e-14)
    # the two product forms of each class
    for c in (16, 24, 32, 48):
        assert lib.negf_set_block_class(0, 12, c) == 0
        assert lib.negf_set_product_form(12, _lib.PRODUCTS_VECTOR) == 0
        t_vec = transmission_spectrum(w, dev, lead, lead, eta=1e-6)
        assert lib.negf_set_product_form(12, 0) == 0
        assert_parity(t_vec, refs)
        # other rounding (another product order), same T to the oracle's bar
        np.testing.assert_allclose(t_vec, got[c], rtol=1e-10, atol=1e-14)
    # the two kernel families
    assert lib.negf_set_block_class(0, 12, 0) == 0
    for fam in (_lib.FAMILY_WORKGROUP, _lib.FAMILY_WAVE):
        assert lib.negf_set_kernel_family(0, 12, fam) == 0
        assert_parity(transmission_spectrum(w, dev, lead, lead, eta=1e-6), refs)
    assert lib.negf_set_kernel_family(0, 12, 0) == 0
