"""GPU full-counting-statistics cumulants (k_fcs_cumulants) and batched k-moments,
against the reference's own numbers (study runs, reference-made vectors) and the
oracle restatement of core/fcs.py."""

import numpy as np
import pytest

from negfpy_amd.core import (
    HeatCurrentCumulants,
    heat_current_cumulants_from_k_moments,
    heat_current_cumulants_from_spectrum,
    heat_current_cumulants_scan,
    heat_current_uncertainty,
    transmission_k_moments_spectrum,
)
from oracle import negf_oracle as o
from test_fcs_oracle import study
from test_reference_vectors import REF, _moments_by_reference_loop, _sq_problem

pytestmark = pytest.mark.gpu


def _close(a, b, rtol):
    if np.isnan(b):
        return np.isnan(a)
    return abs(a - b) <= rtol * abs(b) + 1e-300


@pytest.mark.parametrize("form", ["spectrum", "moments"])
def test_cumulants_match_reference_vectors(gpu, form):
    for (tl, tr), (c1_ref, c2_ref) in zip(REF["fcs_pairs"], REF[f"fcs_{form}"]):
        if form == "spectrum":
            c = heat_current_cumulants_from_spectrum(REF["fcs_omega"], REF["fcs_t1"], tl, tr)
        else:
            c = heat_current_cumulants_from_k_moments(REF["fcs_omega"], REF["fcs_t1"],
                                                      REF["fcs_t2"], tl, tr)
        assert isinstance(c, HeatCurrentCumulants)
        assert _close(c.c1_j_per_s, c1_ref, 1e-12) or abs(c.c1_j_per_s - c1_ref) < 1e-30
        assert _close(c.c2_j2_per_s, c2_ref, 1e-12)


@pytest.mark.parametrize("name", ["graphene_bias", "tan_cond"])
def test_scan_reproduces_study_runs(gpu, name):
    """A whole study bias/temperature scan in one launch: the reference's printed C1, C2
    (11 digits), including its NaN C2 where expm1(x)^2 overflows."""
    d, rows = study(name)
    w = o.qe_omega_to_rad_s(d["omega_qe"])
    c = heat_current_cumulants_scan(w, d["t_mean"], d["t2_mean"], rows[:, 1], rows[:, 2])
    for (c1, c2), r in zip(c, rows):
        assert _close(c1, r[7], 1e-9), (r[1], r[2], c1, r[7])
        assert _close(c2, r[8], 1e-9), (r[1], r[2], c2, r[8])


def test_uncertainty_and_errors(gpu):
    c = HeatCurrentCumulants(c1_j_per_s=REF["fcs_moments"][0, 0], c2_j2_per_s=REF["fcs_moments"][0, 1])
    u = heat_current_uncertainty(c, measurement_time_s=1e-9)
    np.testing.assert_allclose([u.measurement_time_s, u.std_current_j_per_s, u.variance_energy_j2,
                                u.std_energy_j], REF["fcs_uncertainty"], rtol=1e-15)
    with pytest.raises(ValueError, match="Temperature must be positive"):
        heat_current_cumulants_from_spectrum([1.0, 2.0], [1.0, 1.0], 0.0, 10.0)
    with pytest.raises(ValueError, match="strictly increasing"):
        heat_current_cumulants_from_k_moments([2.0, 1.0], [1.0, 1.0], [1.0, 1.0], 5.0, 1.0)
    with pytest.raises(ValueError, match="measurement_time_s"):
        heat_current_uncertainty(c, measurement_time_s=0.0)


def test_k_moments_spectrum_matches_reference_loop(gpu):
    dev, lead, kpts = _sq_problem()
    ws = [0.9, 1.1, 0.0, 1.6, 0.4, 1.9]
    res = transmission_k_moments_spectrum(ws, dev, lead, lead, kpts, eta_values=(1e-8, 1e-7),
                                          min_success_fraction=0.7, nonnegative_tolerance=1e-8)
    ref = _moments_by_reference_loop(ws, dev, lead, kpts, (1e-8, 1e-7), 0.7, 1e-8)
    for i, (t1, t2, eta, nk) in enumerate(ref):
        if np.isnan(t1):
            assert np.isnan(res["t_mean"][i]) and res["n_k_success"][i] == 0
            continue
        assert abs(res["t_mean"][i] - t1) <= 1e-10 * max(t1, 1e-3)
        assert abs(res["t2_mean"][i] - t2) <= 1e-10 * max(t2, 1e-3)
        assert res["n_k_success"][i] == nk
    assert len(res["omega_failures"]) == sum(1 for r in ref if np.isnan(r[0]))
