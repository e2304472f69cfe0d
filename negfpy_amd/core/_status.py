"""Map per-instance kernel status codes to the exceptions the reference raises.

Lead decimation (src/negfpy/core/surface_gf.py:52-79): non-convergence and
non-finite iterates raise RuntimeError; an exactly singular (z - eps) makes
numpy.linalg.inv raise LinAlgError.  Device solve (src/negfpy/core/negf.py:353):
SuperLU raises RuntimeError on an exactly singular matrix and propagates NaN
otherwise.
"""

from __future__ import annotations

import numpy as np

from negfpy_amd import _lib

MSG_NOCONV = "Sancho-Rubio decimation did not converge within maxiter."
MSG_NONFINITE = ("Sancho-Rubio decimation produced non-finite values. "
                 "Try a larger eta or a different surface_gf method.")
MSG_EIGEN_ZERO_ETA = (
    "eigenmode surface GF at eta = 0 inside a band is not supported on the GPU: the "
    "propagating modes lie on the unit circle and the reference keeps n of them in the "
    "order its dense eigensolver returns them (surface_gf.py:118-124); use eta > 0")


def raise_for_lead_status(code: int) -> None:
    if code == _lib.ST_OK:
        return
    if code == _lib.ST_NOCONV:
        raise RuntimeError(MSG_NOCONV)
    if code == _lib.ST_NONFINITE:
        raise RuntimeError(MSG_NONFINITE)
    if code == _lib.ST_SINGULAR:
        raise np.linalg.LinAlgError("Singular matrix")
    if code == _lib.ST_EIGEN_ZERO_ETA:
        raise NotImplementedError(MSG_EIGEN_ZERO_ETA)
    raise RuntimeError(f"unexpected lead status {code}")


def transmission_value(code: int, t: float) -> float:
    """Return T for one instance or raise the reference's exception for its failure."""
    if code in (_lib.ST_NOCONV, _lib.ST_NONFINITE, _lib.ST_SINGULAR, _lib.ST_EIGEN_ZERO_ETA):
        raise_for_lead_status(code)
    if code == _lib.ST_DEV_SINGULAR:
        raise RuntimeError("Factor is exactly singular")
    if code == _lib.ST_DEV_NONFINITE:
        return float("nan")
    if code != _lib.ST_OK:
        raise RuntimeError(f"unexpected transmission status {code}")
    return float(t)
