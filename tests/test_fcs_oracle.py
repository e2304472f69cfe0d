"""The oracle's full-counting-statistics restatement (core/fcs.py) against the
reference's own numbers: its study runs (k-moments <T>, <T^2> -> C1, C2 over a
bias or temperature scan, printed to 11 digits) and vectors it wrote at full
precision (both cumulant forms, an overflow-prone low temperature, equal
temperatures, the finite-time uncertainty)."""

import os

import numpy as np
import pytest

from oracle import negf_oracle as o
from test_reference_vectors import REF

HERE = os.path.dirname(os.path.abspath(__file__))


def study(name):
    d = np.load(os.path.join(HERE, "golden", "study_fcs", f"{name}_fcs_moments.npz"))
    rows = np.loadtxt(os.path.join(HERE, "golden", "study_fcs", f"{name}_fcs.tsv"))
    return d, rows


def _close(a, b, rtol):
    if np.isnan(b):
        return np.isnan(a)
    return abs(a - b) <= rtol * abs(b) + 1e-300


@pytest.mark.parametrize("name", ["graphene_bias", "tan_cond"])
def test_oracle_reproduces_study_cumulants(name):
    d, rows = study(name)
    w = o.qe_omega_to_rad_s(d["omega_qe"])
    n_checked = 0
    for r in rows:
        tl, tr, c1_ref, c2_ref = r[1], r[2], r[7], r[8]
        with np.errstate(over="ignore", invalid="ignore"):
            c1, c2 = o.fcs_cumulants(w, d["t_mean"], d["t2_mean"], tl, tr)
        assert _close(c1, c1_ref, 1e-9), (tl, tr, c1, c1_ref)
        assert _close(c2, c2_ref, 1e-9), (tl, tr, c2, c2_ref)
        n_checked += 1
    assert n_checked == rows.shape[0] > 50


@pytest.mark.parametrize("form", ["spectrum", "moments"])
def test_oracle_matches_reference_cumulants(form):
    t2 = None if form == "spectrum" else REF["fcs_t2"]
    for (tl, tr), (c1_ref, c2_ref) in zip(REF["fcs_pairs"], REF[f"fcs_{form}"]):
        with np.errstate(over="ignore", invalid="ignore"):
            c1, c2 = o.fcs_cumulants(REF["fcs_omega"], REF["fcs_t1"], t2, tl, tr)
        assert _close(c1, c1_ref, 1e-13) or abs(c1 - c1_ref) < 1e-30
        assert _close(c2, c2_ref, 1e-13)


def test_phi_series_join_the_closed_forms():
    """The small-x series and x/expm1(x), x^2 e^x/expm1(x)^2 agree at the switch."""
    x = np.array([0.999e-6, 1.001e-6])
    assert abs(o.fcs_phi1(x)[0] - o.fcs_phi1(x)[1]) < 1e-9
    assert abs(o.fcs_phi2(x)[0] - o.fcs_phi2(x)[1]) < 1e-11
    assert o.fcs_phi1(np.array([700.0]))[0] == 0.0 and o.fcs_phi2(np.array([701.0]))[0] == 0.0


def test_oracle_fcs_argument_errors():
    with pytest.raises(ValueError, match="strictly increasing"):
        o.fcs_cumulants([2.0, 1.0], [1.0, 1.0], None, 300.0, 290.0)
    with pytest.raises(ValueError, match="non-negative"):
        o.fcs_cumulants([1.0, 2.0], [1.0, -1.0], None, 300.0, 290.0)
    with pytest.raises(ValueError, match="Temperature must be positive"):
        o.fcs_cumulants([1.0, 2.0], [1.0, 1.0], [1.0, 1.0], 0.0, 290.0)
