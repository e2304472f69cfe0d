"""Surface Green's functions of periodic leads on the GPU.

API of the reference ``src/negfpy/core/surface_gf.py`` (:12-311): the same
functions, arguments, defaults, method aliases and exceptions.  Every method
runs the batched decimation kernel (``negf_sancho_rubio_batched``).

The eigenmode routes (``generalized_eigen``, ``generalized_eigen_svd``,
``legacy_eigen_svd``; surface_gf.py:84-251) select the n decaying Bloch modes
of the lead's quadratic eigenvalue problem and return the retarded surface
Green's function they span.  For eta > 0 the unit circle separates those modes
from the growing ones and that Green's function is the unique retarded
solution of g = (z - d00 - d01 g d10)^-1, which is the fixed point decimation
converges to; those entry points therefore keep their argument checks, take the
decimation result, and where decimation fails take the eigenmode kernel's
(``negf_surface_gf_eigen_batched``: the subspace of the n smallest-|lambda| modes
from a spectral projector).  Not reproduced, and reported as such: eta = 0
inside the band, where modes of modulus 1 sit on both sides of the n-th and the
reference keeps some of them in the order LAPACK's rounding gives them
(NotImplementedError, kernel status 4); legacy_eigen_svd where a decaying mode
lies within unit_circle_tol of the unit circle, which the reference then drops
(NotImplementedError); and the
SVD cutoff of rank-deficient mode sets (an exactly singular pivot instead).
"""

from __future__ import annotations

import numpy as np

from . import _engine
from ._status import raise_for_lead_status

Array = np.ndarray

_SR = {"sancho_rubio", "sancho-rubio", "sr"}
_GE = {"generalized_eigen", "generalized-eigen", "eig", "gep"}
_GESVD = {"generalized_eigen_svd", "generalized-eigen-svd", "ge_svd", "gep_svd"}
_LEGACY = {"legacy_eigen_svd", "legacy-eigen-svd", "legacy", "old_svd"}

_UNKNOWN = ("Unknown surface GF method. Use one of: "
            "'sancho_rubio', 'generalized_eigen', 'generalized_eigen_svd', 'legacy_eigen_svd'.")

_default_precision = "fp64"


def set_default_precision(precision: str) -> None:
    """Select "fp64" (default) or "mixed" for calls that pass no ``precision``."""
    global _default_precision
    _engine.precision_code(precision)
    _default_precision = precision


def get_default_precision() -> str:
    return _default_precision


def _prec(precision):
    return _default_precision if precision is None else precision


def canonical_method(method: str) -> str:
    """Map a method alias to its canonical name; ValueError if unknown (surface_gf.py:269-311)."""
    m = method.strip().lower()
    if m in _SR:
        return "sancho_rubio"
    if m in _GE:
        return "generalized_eigen"
    if m in _GESVD:
        return "generalized_eigen_svd"
    if m in _LEGACY:
        return "legacy_eigen_svd"
    raise ValueError(_UNKNOWN)


def surface_route(method: str) -> int:
    """Library route for a method: decimation for "sancho_rubio"; for the eigenmode
    methods, decimation with the eigenmode fallback where it fails."""
    from negfpy_amd import _lib

    return _lib.SURFACE_DECIMATION if canonical_method(method) == "sancho_rubio" else _lib.SURFACE_EIGEN


_EIGEN = 1   # the library's eigenmode route code (SURFACE_EIGEN)


def resolve_d10(d01: Array, d10: Array | None) -> Array:
    """d10 defaults to d01^H (surface_gf.py:12-15)."""
    if d10 is None:
        return np.asarray(d01).conj().T.astype(np.complex128, copy=True)
    return np.asarray(d10).astype(np.complex128, copy=True)


def pack_lead(d00, d01, d10=None) -> Array:
    """(3, n, n) complex128 stack (d00, d01, d10 resolved)."""
    return np.stack([np.asarray(d00, dtype=np.complex128), np.asarray(d01, dtype=np.complex128),
                     resolve_d10(np.asarray(d01), d10)])


def _single(omega, d00, d01, d10, eta, tol, maxiter, precision, route=0):
    res = _engine.surface_gf_batch(pack_lead(d00, d01, d10)[None], [0], [float(omega)],
                                   [float(eta)], tol=tol, maxiter=maxiter, precision=_prec(precision),
                                   route=route)
    raise_for_lead_status(int(res.status[0].item()))
    return res.g[0].cpu().numpy()


def surface_gf_sancho_rubio(omega: float, d00: Array, d01: Array, d10: Array | None = None,
                            eta: float = 1e-8, tol: float = 1e-12, maxiter: int = 200, *,
                            precision: str | None = None) -> Array:
    """Retarded surface GF by Lopez-Sancho decimation (surface_gf.py:18-81)."""
    return _single(omega, d00, d01, d10, eta, tol, maxiter, precision)


def _check_square(d00, d01):
    n = int(d00.shape[0])
    if d00.shape != (n, n) or d01.shape != (n, n):
        raise ValueError("d00 and d01 must be square matrices of the same shape.")


def surface_gf_generalized_eigen(omega: float, d00: Array, d01: Array, d10: Array | None = None,
                                 eta: float = 1e-8, unit_circle_tol: float = 1e-10, *,
                                 precision: str | None = None) -> Array:
    """Eigenmode route (surface_gf.py:84-133); see the module docstring.

    ``unit_circle_tol``: the reference keeps the modes with |lambda| < 1 - tol and, when
    fewer than n are left, the n of smallest |lambda| (:118-124).  For eta > 0 exactly
    n of the 2n modes lie inside the unit circle and they are the n smallest, so the
    selection -- and the result, the decaying (retarded) solution computed here -- is the
    same for every tol.  At eta = 0 inside a band the modes on the unit circle are chosen
    in the order the dense eigensolver returns them; that case raises
    NotImplementedError."""
    _check_square(d00, d01)
    return _single(omega, d00, d01, d10, eta, 1e-12, 200, precision, _EIGEN)


def surface_gf_generalized_eigen_svd(omega: float, d00: Array, d01: Array,
                                     d10: Array | None = None, eta: float = 1e-8,
                                     unit_circle_tol: float = 1e-10, svd_rcond: float = 1e-12, *,
                                     precision: str | None = None) -> Array:
    """Eigenmode route with SVD pseudo-inverse (surface_gf.py:136-185).

    Mode selection as :func:`surface_gf_generalized_eigen`.  ``svd_rcond`` sets the
    reference's pseudo-inverse cutoff (:182); for a mode matrix whose condition number
    is below 1 / svd_rcond (every lead of the tests and studies) the pseudo-inverse is
    the inverse and the result the decaying solution computed here.  A mode matrix
    beyond that (rank-deficient modes) ends as an exactly singular pivot on the GPU
    (numpy.linalg.LinAlgError) rather than the reference's truncated pseudo-inverse."""
    _check_square(d00, d01)
    if svd_rcond <= 0.0:
        raise ValueError("svd_rcond must be positive.")
    return _single(omega, d00, d01, d10, eta, 1e-12, 200, precision, _EIGEN)


def surface_gf_legacy_eigen_svd(omega: float, d00: Array, d01: Array, d10: Array | None = None,
                                eta: float = 1e-8, unit_circle_tol: float = 0.0, *,
                                precision: str | None = None) -> Array:
    """Legacy Fortran-style eigenmode route (surface_gf.py:201-251).

    The reference keeps every mode with |lambda| < 1 - unit_circle_tol (:218-223).  When
    those are the n decaying modes -- for eta > 0 unless a decaying mode lies within
    unit_circle_tol of the unit circle (1 - |lambda| is of order eta / group velocity) --
    its result is the decaying solution computed here.  The decaying modes' lambda are
    the eigenvalues of g d10 (the transfer matrix); if one of them is within the tolerance
    the reference would drop it and work with fewer than n modes and a pseudo-inverse,
    which is not reproduced: NotImplementedError."""
    _check_square(d00, d01)
    g = _single(omega, d00, d01, d10, eta, 1e-12, 200, precision, _EIGEN)
    if unit_circle_tol > 0.0:
        lam = np.linalg.eigvals(g @ resolve_d10(np.asarray(d01), d10))
        if np.any(1.0 - np.abs(lam) < unit_circle_tol):
            raise NotImplementedError(
                "legacy_eigen_svd: a decaying mode lies within unit_circle_tol of the unit "
                "circle, so the reference keeps fewer than n modes (pseudo-inverse); the GPU "
                "route computes the solution with all n decaying modes")
    return g


def surface_gf(omega: float, d00: Array, d01: Array, d10: Array | None = None, eta: float = 1e-8,
               *, method: str = "sancho_rubio", tol: float = 1e-12, maxiter: int = 200,
               unit_circle_tol: float = 1e-10, svd_rcond: float = 1e-12,
               precision: str | None = None) -> Array:
    """Method dispatcher (surface_gf.py:254-311)."""
    m = canonical_method(method)
    if m == "sancho_rubio":
        return surface_gf_sancho_rubio(omega, d00, d01, d10, eta, tol, maxiter, precision=precision)
    if m == "generalized_eigen":
        return surface_gf_generalized_eigen(omega, d00, d01, d10, eta, unit_circle_tol,
                                            precision=precision)
    if m == "generalized_eigen_svd":
        return surface_gf_generalized_eigen_svd(omega, d00, d01, d10, eta, unit_circle_tol,
                                                svd_rcond, precision=precision)
    return surface_gf_legacy_eigen_svd(omega, d00, d01, d10, eta, unit_circle_tol,
                                       precision=precision)
