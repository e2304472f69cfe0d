"""FP64 CPU oracle for the phonon-NEGF transmission path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``negfpy_amd``) never imports or calls it and has
no CPU fallback.

It restates, in plain NumPy/SciPy complex128, the algorithms of the reference
package ``negfpy`` (files cited per function, paths relative to the reference
root):

* surface Green's functions  -- ``src/negfpy/core/surface_gf.py``
* device solve, transmission -- ``src/negfpy/core/negf.py``
* lead surface DOS           -- ``src/negfpy/core/dos.py:31-61``
* Landauer heat current      -- ``src/negfpy/core/thermal_current.py:30-67``
* thermal conductance k(T)   -- ``src/negfpy/workflows/ifc_bulk.py:408-438``
* QE unit conversion         -- ``src/negfpy/modeling/units.py:8-26``

Besides the reference's own sparse-LU device solve it carries two independent
device solvers (a dense inverse and a block recursive Green's function sweep)
so the tests can pin each against the others and against closed forms.

Pinning status (what checks each function; see ``tests/test_oracle_*.py``):

* ``sancho_rubio``            -- chain closed form; Riccati fixed-point residual
                                 on random multi-DOF leads; agreement with the
                                 eigen routes.
* ``generalized_eigen[_svd]`` -- chain closed form; agreement with SR (<=1e-10).
* ``legacy_eigen_svd``        -- chain closed form; agreement with SR.
* ``transmission_sparse``     -- dense inverse and RGF agree to 1e-12; T = n
                                 plateau for identity-block leads; T = 1/0 in
                                 and out of band for the chain; omega_scale
                                 invariance.
* ``surface_dos``             -- chain closed form (w^2 sqrt(4 - w^2)/pi).
* ``heat_current_from_spectrum``, ``thermal_conductance`` -- reference study
                                 T(w) -> k(T) TSV pairs; step-function closed
                                 form.

Every function above is also checked against vectors the reference package
itself wrote on seeded inputs (``tests/golden/reference_vectors.npz``, made by
``tests/golden/make_reference_vectors.py``; ``tests/test_reference_vectors.py``):
random multi-DOF transmission with every surface-GF method, omega_scale,
contacts and two leads, surface GFs (including the eigen routes at eta = 0),
DOS, the device Green's function and the heat current.
"""

from __future__ import annotations

import numpy as np
from scipy.linalg import eig
from scipy.sparse import csc_matrix, identity, lil_matrix
from scipy.sparse.linalg import splu

# SI constants used by the reference (thermal_current.py:15-16, ifc_bulk.py:55-56).
HBAR = 1.054_571_817e-34
KB = 1.380_649e-23
# CODATA constants of modeling/units.py:9-12.
RYDBERG_J = 2.1798723611035e-18
BOHR_M = 5.29177210903e-11
ELECTRON_MASS_KG = 9.1093837015e-31


# --------------------------------------------------------------------------
# surface Green's functions (surface_gf.py)
# --------------------------------------------------------------------------

def _d10_of(d01, d10):
    """surface_gf.py:12-15 -- d10 defaults to d01^H; always a complex copy."""
    if d10 is None:
        return np.array(np.conj(d01).T, dtype=np.complex128)
    return np.array(d10, dtype=np.complex128)


def sancho_rubio(omega, d00, d01, d10=None, eta=1e-8, tol=1e-12, maxiter=200,
                 return_iters=False):
    """Lopez-Sancho decimation, surface_gf.py:18-81.

    Iterates  g = (z - e)^-1,  e_s += a g b,  e += a g b + b g a,
    a <- a g a,  b <- b g b  with z = (w + i eta)^2 until max(|a|,|b|) < tol,
    then returns (z - e_s)^-1.  Raises RuntimeError on FP traps, non-finite
    iterates or exhaustion of ``maxiter`` exactly where the reference does.
    """
    z = (omega + 1j * eta) ** 2
    n = d00.shape[0]
    zi = z * np.eye(n, dtype=np.complex128)
    a = np.array(d01, dtype=np.complex128)
    b = _d10_of(d01, d10)
    e = np.array(d00, dtype=np.complex128)
    es = e.copy()
    it = 0
    converged = False
    while it < maxiter:
        it += 1
        try:
            with np.errstate(over="raise", invalid="raise", divide="raise"):
                g = np.linalg.inv(zi - e)
                ag = a @ g
                bg = b @ g
                agb = ag @ b
                bga = bg @ a
                es_n = es + agb
                e_n = e + agb + bga
                a_n = ag @ a
                b_n = bg @ b
        except FloatingPointError as exc:
            raise RuntimeError("Sancho-Rubio decimation became numerically unstable.") from exc
        if not (np.isfinite(es_n).all() and np.isfinite(e_n).all()
                and np.isfinite(a_n).all() and np.isfinite(b_n).all()):
            raise RuntimeError("Sancho-Rubio decimation produced non-finite values.")
        err = max(float(np.max(np.abs(a_n))), float(np.max(np.abs(b_n))))
        es, e, a, b = es_n, e_n, a_n, b_n
        if err < tol:
            converged = True
            break
    if not converged:
        raise RuntimeError("Sancho-Rubio decimation did not converge within maxiter.")
    g_s = np.linalg.inv(zi - es)
    return (g_s, it) if return_iters else g_s


def _qep_pencil(omega, d00, d01, d10, eta):
    """Companion linearisation of (c l^2 + a l + b) phi = 0, surface_gf.py:102-111."""
    n = d00.shape[0]
    z = (omega + 1j * eta) ** 2
    eye = np.eye(n, dtype=np.complex128)
    zero = np.zeros((n, n), dtype=np.complex128)
    a = np.asarray(d00, dtype=np.complex128) - z * eye
    b = _d10_of(d01, d10)
    c = np.asarray(d01, dtype=np.complex128)
    pa = np.block([[-a, -b], [eye, zero]])
    pb = np.block([[c, zero], [zero, eye]])
    return z, pa, pb, c


def _pick_decaying(evals, evecs, n, unit_circle_tol, what):
    keep = np.isfinite(evals)
    evals = evals[keep]
    evecs = evecs[:, keep]
    if evals.size < n:
        raise RuntimeError(f"{what} surface GF failed: insufficient finite eigenmodes.")
    mag = np.abs(evals)
    inside = np.where(mag < 1.0 - float(unit_circle_tol))[0]
    if inside.size >= n:
        sel = inside[np.argsort(mag[inside])[:n]]
    else:
        sel = np.argsort(mag)[:n]
    return evals[sel], evecs[n:, sel]


def generalized_eigen(omega, d00, d01, d10=None, eta=1e-8, unit_circle_tol=1e-10):
    """QEP route with an explicit mode-matrix inverse, surface_gf.py:84-133."""
    n = int(d00.shape[0])
    if d00.shape != (n, n) or d01.shape != (n, n):
        raise ValueError("d00 and d01 must be square matrices of the same shape.")
    z, pa, pb, c = _qep_pencil(omega, d00, d01, d10, eta)
    w, v = eig(pa, pb, right=True)
    lam, phi = _pick_decaying(w, v, n, unit_circle_tol, "Generalized-eigen")
    if np.linalg.matrix_rank(phi) < n:
        raise RuntimeError("Generalized-eigen surface GF failed: selected eigenvectors are rank-deficient.")
    transfer = phi @ np.diag(lam) @ np.linalg.inv(phi)
    return np.linalg.inv(z * np.eye(n) - np.asarray(d00, dtype=np.complex128) - c @ transfer)


def generalized_eigen_svd(omega, d00, d01, d10=None, eta=1e-8, unit_circle_tol=1e-10,
                          svd_rcond=1e-12):
    """QEP route with an SVD pseudo-inverse, surface_gf.py:136-185."""
    n = int(d00.shape[0])
    if d00.shape != (n, n) or d01.shape != (n, n):
        raise ValueError("d00 and d01 must be square matrices of the same shape.")
    if svd_rcond <= 0.0:
        raise ValueError("svd_rcond must be positive.")
    z, pa, pb, c = _qep_pencil(omega, d00, d01, d10, eta)
    w, v = eig(pa, pb, right=True)
    lam, phi = _pick_decaying(w, v, n, unit_circle_tol, "Generalized-eigen-SVD")
    transfer = phi @ np.diag(lam) @ np.linalg.pinv(phi, rcond=float(svd_rcond))
    return np.linalg.inv(z * np.eye(n) - np.asarray(d00, dtype=np.complex128) - c @ transfer)


def _pinv_machine_cutoff(m):
    """SVD pseudo-inverse with cutoff max(tiny, s0*eps), surface_gf.py:188-198."""
    u, s, vh = np.linalg.svd(m, full_matrices=True)
    if s.size == 0:
        return np.zeros((m.shape[1], m.shape[0]), dtype=np.complex128)
    cut = max(np.finfo(float).tiny, float(s[0]) * np.finfo(float).eps)
    sinv = np.where(s >= cut, 1.0 / np.where(s >= cut, s, 1.0), 0.0)
    return (vh.conj().T[:, : s.size] * sinv) @ u[:, : s.size].conj().T


def legacy_eigen_svd(omega, d00, d01, d10=None, eta=1e-8, unit_circle_tol=0.0):
    """Fortran sfg_solve-style route, surface_gf.py:201-251."""
    n = int(d00.shape[0])
    if d00.shape != (n, n) or d01.shape != (n, n):
        raise ValueError("d00 and d01 must be square matrices of the same shape.")
    z = (omega + 1j * eta) ** 2
    eye = np.eye(n, dtype=np.complex128)
    zero = np.zeros((n, n), dtype=np.complex128)
    h01 = np.asarray(d01, dtype=np.complex128)
    h10 = _d10_of(d01, d10)
    th = z * eye - np.asarray(d00, dtype=np.complex128)
    pa = np.block([[th, -eye], [h10, zero]])
    pb = np.block([[h01, zero], [zero, eye]])
    w, v = eig(pa, pb, right=True)
    keep = np.isfinite(w)
    w, v = w[keep], v[:, keep]
    if w.size == 0:
        raise RuntimeError("Legacy-eigen-SVD surface GF failed: no finite eigenmodes.")
    dec = np.where(np.abs(w) < 1.0 - float(unit_circle_tol))[0]
    if dec.size == 0:
        raise RuntimeError("Legacy-eigen-SVD surface GF failed: no decaying eigenmodes.")
    e_up = v[:n, dec]
    g11 = np.linalg.inv(th - h01 @ e_up @ np.diag(w[dec]) @ _pinv_machine_cutoff(e_up))
    return np.linalg.inv(th - h01 @ g11 @ h10)


_METHODS = {
    "sancho_rubio": "sr", "sancho-rubio": "sr", "sr": "sr",
    "generalized_eigen": "ge", "generalized-eigen": "ge", "eig": "ge", "gep": "ge",
    "generalized_eigen_svd": "gesvd", "generalized-eigen-svd": "gesvd", "ge_svd": "gesvd",
    "gep_svd": "gesvd",
    "legacy_eigen_svd": "legacy", "legacy-eigen-svd": "legacy", "legacy": "legacy",
    "old_svd": "legacy",
}


def surface_gf(omega, d00, d01, d10=None, eta=1e-8, *, method="sancho_rubio", tol=1e-12,
               maxiter=200, unit_circle_tol=1e-10, svd_rcond=1e-12):
    """Method dispatcher, surface_gf.py:254-311."""
    key = _METHODS.get(method.strip().lower())
    if key == "sr":
        return sancho_rubio(omega, d00, d01, d10, eta, tol, maxiter)
    if key == "ge":
        return generalized_eigen(omega, d00, d01, d10, eta, unit_circle_tol)
    if key == "gesvd":
        return generalized_eigen_svd(omega, d00, d01, d10, eta, unit_circle_tol, svd_rcond)
    if key == "legacy":
        return legacy_eigen_svd(omega, d00, d01, d10, eta, unit_circle_tol)
    raise ValueError("Unknown surface GF method.")


def riccati_residual(omega, d00, d01, d10, eta, g):
    """|| g - (z - d00 - d01 g d10)^-1 || / ||g||: zero for any surface GF."""
    z = (omega + 1j * eta) ** 2
    n = d00.shape[0]
    d10 = _d10_of(d01, d10)
    rhs = np.linalg.inv(z * np.eye(n) - d00 - d01 @ g @ d10)
    return float(np.linalg.norm(g - rhs) / np.linalg.norm(g))


# --------------------------------------------------------------------------
# device solve (negf.py)
# --------------------------------------------------------------------------

def contact_indices(indices, dim, default):
    """negf.py:119-138."""
    if indices is None:
        idx = np.asarray(default, dtype=int).ravel()
    elif isinstance(indices, slice):
        idx = np.arange(dim, dtype=int)[indices]
    else:
        idx = np.asarray(indices, dtype=int).ravel()
    if idx.size == 0:
        raise ValueError("Contact indices must be non-empty.")
    if np.any(idx < 0) or np.any(idx >= dim):
        raise ValueError("Contact indices out of device range.")
    if np.unique(idx).size != idx.size:
        raise ValueError("Contact indices must be unique.")
    return idx


def assemble_device(onsite, coupling):
    """Sparse block-tridiagonal D with lower blocks = coupling^H, negf.py:53-69."""
    nl = len(onsite)
    n = onsite[0].shape[0]
    dim = nl * n
    d = lil_matrix((dim, dim), dtype=np.complex128)
    for i, blk in enumerate(onsite):
        d[i * n:(i + 1) * n, i * n:(i + 1) * n] = blk
    for i, blk in enumerate(coupling):
        r = slice(i * n, (i + 1) * n)
        c = slice((i + 1) * n, (i + 2) * n)
        d[r, c] = blk
        d[c, r] = np.conj(blk).T
    return d.tocsc()


def system(omega, onsite, coupling, lead_l, lead_r, eta=1e-8, eta_device=None,
           v_left=None, v_right=None, idx_left=None, idx_right=None,
           method="sancho_rubio", omega_scale=None):
    """A = (w + i eta_dev)^2 - D - Sigma_L - Sigma_R with contact blocks, negf.py:141-250.

    ``lead_l``/``lead_r`` are (d00, d01, d10-or-None) tuples.  Returns
    (A csc, sigma_l, sigma_r, idx_l, idx_r).
    """
    n = onsite[0].shape[0]
    nl = len(onsite)
    dim = nl * n
    eta_dev = float(eta if eta_device is None else eta_device)
    if eta_dev < 0.0:
        raise ValueError("eta_device must be non-negative.")
    eta_lead = float(eta)
    w = float(omega)
    onsite = [np.asarray(b, dtype=np.complex128) for b in onsite]
    coupling = [np.asarray(b, dtype=np.complex128) for b in coupling]
    lead_l = tuple(None if b is None else np.asarray(b, dtype=np.complex128) for b in lead_l)
    lead_r = tuple(None if b is None else np.asarray(b, dtype=np.complex128) for b in lead_r)
    if omega_scale is not None:
        s = float(omega_scale)
        if s <= 0.0:
            raise ValueError("omega_scale must be positive when provided.")
        s2 = s * s
        w, eta_lead, eta_dev = w / s, eta_lead / s, eta_dev / s
        onsite = [b / s2 for b in onsite]
        coupling = [b / s2 for b in coupling]
        lead_l = tuple(None if b is None else b / s2 for b in lead_l)
        lead_r = tuple(None if b is None else b / s2 for b in lead_r)
    if v_left is None:
        vl = lead_l[2] if lead_l[2] is not None else np.conj(lead_l[1]).T
    else:
        vl = np.asarray(v_left, dtype=np.complex128)
    if v_right is None:
        vr = lead_r[1]
    else:
        vr = np.asarray(v_right, dtype=np.complex128)
    if omega_scale is not None:
        vl = vl / (float(omega_scale) ** 2)
        vr = vr / (float(omega_scale) ** 2)
    ncl, ncr = vl.shape[0], vr.shape[0]
    il = contact_indices(idx_left, dim, np.arange(ncl))
    ir = contact_indices(idx_right, dim, np.arange(dim - ncr, dim))
    if il.size != ncl:
        raise ValueError("Left contact index count must match device_to_lead_left rows.")
    if ir.size != ncr:
        raise ValueError("Right contact index count must match device_to_lead_right rows.")
    dmat = assemble_device(onsite, coupling)

    def sigma(lead, v):
        g = surface_gf(w, lead[0], lead[1], lead[2], eta_lead, method=method)
        v = np.asarray(v, dtype=np.complex128)
        if v.ndim != 2:
            raise ValueError("device_to_lead coupling must be a 2D array.")
        if v.shape[1] != lead[0].shape[0]:
            raise ValueError("device_to_lead coupling has incompatible lead dimension.")
        return v @ g @ np.conj(v).T

    sl = sigma(lead_l, vl)
    sr = sigma(lead_r, vr)
    z = (w + 1j * eta_dev) ** 2
    a = (z * identity(dim, dtype=np.complex128, format="csc") - dmat).tolil()
    a[np.ix_(il, il)] = a[np.ix_(il, il)] - sl
    a[np.ix_(ir, ir)] = a[np.ix_(ir, ir)] - sr
    return a.tocsc(), sl, sr, il, ir


def broadening(sigma):
    """Gamma = i (Sigma - Sigma^H), negf.py:115-116."""
    return 1j * (sigma - np.conj(sigma).T)


def transmission_sparse(omega, onsite, coupling, lead_l, lead_r, **kw):
    """T = Tr[G_L G_LR G_R G_LR^H] from a sparse LU solve for the right-contact columns,
    negf.py:306-357 (the reference algorithm)."""
    a, sl, sr, il, ir = system(omega, onsite, coupling, lead_l, lead_r, **kw)
    dim = a.shape[0]
    rhs = np.zeros((dim, ir.size), dtype=np.complex128)
    rhs[ir, :] = np.eye(ir.size)
    glr = splu(csc_matrix(a)).solve(rhs)[il, :]
    t = np.trace(broadening(sl) @ glr @ broadening(sr) @ np.conj(glr).T)
    return float(np.real_if_close(t).real)


def transmission_dense(omega, onsite, coupling, lead_l, lead_r, **kw):
    """Same T from a brute-force dense inverse of A (pinning check only)."""
    a, sl, sr, il, ir = system(omega, onsite, coupling, lead_l, lead_r, **kw)
    g = np.linalg.inv(a.toarray())
    glr = g[np.ix_(il, ir)]
    return float(np.trace(broadening(sl) @ glr @ broadening(sr) @ np.conj(glr).T).real)


def device_green_dense(omega, onsite, coupling, lead_l, lead_r, **kw):
    """(G, Sigma_L, Sigma_R) embedded in the full device, negf.py:253-303."""
    a, sl, sr, il, ir = system(omega, onsite, coupling, lead_l, lead_r, **kw)
    dim = a.shape[0]
    g = splu(csc_matrix(a)).solve(np.eye(dim, dtype=np.complex128))
    big_l = np.zeros((dim, dim), dtype=np.complex128)
    big_r = np.zeros((dim, dim), dtype=np.complex128)
    big_l[np.ix_(il, il)] = sl
    big_r[np.ix_(ir, ir)] = sr
    return g, big_l, big_r


def rgf_g1n(z, onsite, coupling, sigma_l_blk, sigma_r_blk):
    """Block recursive Green's function for G_{1N} of the block-tridiagonal A.

    A_ii = z - D_i (- Sigma_L on layer 1, - Sigma_R on layer N), A_{i,i+1} = -H_i,
    A_{i+1,i} = -H_i^H.  Left-connected sweep  g_i = (A_ii - H_{i-1}^H Y_{i-1})^-1,
    Y_i = g_i H_i,  Q_i = Q_{i-1} Y_i,  G_{1N} = Q_{N-1} g_N.  This is the device
    algorithm the GPU kernel runs; it equals G[layer 1, layer N] of the sparse
    solve above because both invert the same matrix.
    """
    nl = len(onsite)
    n = onsite[0].shape[0]
    eye = np.eye(n, dtype=np.complex128)
    s = np.zeros((n, n), dtype=np.complex128)
    q = None
    for i in range(nl):
        a = z * eye - onsite[i] - s
        if i == 0:
            a = a - sigma_l_blk
        if i == nl - 1:
            a = a - sigma_r_blk
        g = np.linalg.inv(a)
        if i == nl - 1:
            return g if q is None else q @ g
        y = g @ coupling[i]
        s = np.conj(coupling[i]).T @ y
        q = y if q is None else q @ y
    raise AssertionError("unreachable")


def transmission_rgf(omega, onsite, coupling, lead, eta=1e-8, eta_device=None):
    """T via SR + RGF for default full-layer contacts of one lead on both sides."""
    d00, d01, d10 = lead
    d10 = _d10_of(d01, d10)
    g = sancho_rubio(omega, d00, d01, d10, eta)
    sl = d10 @ g @ np.conj(d10).T
    sr = d01 @ g @ np.conj(d01).T
    eta_dev = eta if eta_device is None else eta_device
    z = (omega + 1j * eta_dev) ** 2
    g1n = rgf_g1n(z, [np.asarray(b, np.complex128) for b in onsite],
                  [np.asarray(b, np.complex128) for b in coupling], sl, sr)
    return float(np.trace(broadening(sl) @ g1n @ broadening(sr) @ np.conj(g1n).T).real)


# --------------------------------------------------------------------------
# DOS, heat current, thermal conductance
# --------------------------------------------------------------------------

def surface_dos(omega, g, normalize_per_mode=False):
    """(w / pi) Re Tr[i (g - g^H)], dos.py:57-61."""
    d = float((omega / np.pi) * np.trace(1j * (g - np.conj(g).T)).real)
    return d / g.shape[0] if normalize_per_mode else d


def qe_omega_to_rad_s(omega):
    """units.py:18-26."""
    return np.asarray(omega) * np.sqrt(RYDBERG_J / (2.0 * ELECTRON_MASS_KG * BOHR_M * BOHR_M))


def bose_weight(omega_rad_s, temp_left, temp_right):
    """k_B T_L phi(x_L) - k_B T_R phi(x_R), phi(x) = x/(e^x - 1), thermal_current.py:30-47."""
    if temp_left <= 0.0 or temp_right <= 0.0:
        raise ValueError("Temperatures must be positive in Kelvin.")
    w = np.asarray(omega_rad_s, dtype=float)

    def phi(x):
        out = np.zeros_like(x)
        lo = x < 1e-6
        mid = (x >= 1e-6) & (x < 700.0)
        out[lo] = 1.0 - 0.5 * x[lo] + x[lo] ** 2 / 12.0
        out[mid] = x[mid] / np.expm1(x[mid])
        return out

    return (KB * temp_left * phi(HBAR * w / (KB * temp_left))
            - KB * temp_right * phi(HBAR * w / (KB * temp_right)))


def trapezoid(y, x):
    """np.trapz(y, x) restated: sum of (x[i+1]-x[i]) (y[i+1]+y[i]) / 2."""
    y = np.asarray(y, dtype=float)
    x = np.asarray(x, dtype=float)
    return float(np.sum(np.diff(x) * (y[1:] + y[:-1]) / 2.0))


def heat_current_from_spectrum(omegas, tvals, temp_left, temp_right):
    """J = (1/2pi) int [k T_L phi_L - k T_R phi_R] T(w) dw, thermal_current.py:50-67."""
    w = np.asarray(list(omegas), dtype=float)
    t = np.asarray(list(tvals), dtype=float)
    if w.ndim != 1:
        raise ValueError("omegas must be 1D.")
    if w.size < 2:
        raise ValueError("omegas must contain at least two points.")
    if np.any(np.diff(w) <= 0.0):
        raise ValueError("omegas must be strictly increasing.")
    if t.shape != w.shape:
        raise ValueError("transmission_vals must have the same length as omegas.")
    return float(trapezoid(bose_weight(w, temp_left, temp_right) * t, w) / (2.0 * np.pi))


def thermal_conductance(omega_qe, tvals, temps):
    """G(T) = (1/2pi) int k_B T(w) x^2 e^-x / (1 - e^-x)^2 dw, ifc_bulk.py:408-438."""
    omega_qe = np.asarray(omega_qe, dtype=float)
    tvals = np.asarray(tvals, dtype=float)
    temps = np.asarray(temps, dtype=float)
    if omega_qe.ndim != 1 or tvals.ndim != 1 or omega_qe.size != tvals.size:
        raise ValueError("omega_qe and transmission_vals must be 1D arrays of equal length.")
    if np.any(np.diff(omega_qe) <= 0.0):
        raise ValueError("omega_qe must be strictly increasing.")
    w = np.asarray(qe_omega_to_rad_s(omega_qe), dtype=float)
    out = np.zeros_like(temps)
    for i, tk in enumerate(temps):
        if tk <= 0.0:
            continue
        x = HBAR * w / (KB * float(tk))
        ker = np.zeros_like(x)
        lo = x < 1e-6
        mid = (x >= 1e-6) & (x < 700.0)
        ker[lo] = 1.0 - x[lo] ** 2 / 12.0
        y = np.exp(-x[mid])
        ker[mid] = x[mid] ** 2 * y / (1.0 - y) ** 2
        out[i] = trapezoid(KB * tvals * ker, w) / (2.0 * np.pi)
    return out


# --------------------------------------------------------------------------
# full counting statistics (core/fcs.py)
# --------------------------------------------------------------------------

def _fcs_grid(values, name):
    """fcs.py:36-44."""
    arr = np.asarray(list(values), dtype=float)
    if arr.ndim != 1:
        raise ValueError(f"{name} must be 1D.")
    if arr.size < 2:
        raise ValueError(f"{name} must contain at least two points.")
    if np.any(np.diff(arr) <= 0.0):
        raise ValueError(f"{name} must be strictly increasing.")
    return arr


def fcs_phi1(x):
    """x / expm1(x) with its series below |x| = 1e-6 and 0 from 700, fcs.py:47-56."""
    x = np.asarray(x, dtype=float)
    out = np.zeros_like(x)
    ax = np.abs(x)
    small = ax < 1e-6
    mid = (ax >= 1e-6) & (x < 700.0)
    xs = x[small]
    out[small] = 1.0 - 0.5 * xs + (xs * xs) / 12.0 - (xs ** 4) / 720.0
    out[mid] = x[mid] / np.expm1(x[mid])
    return out


def fcs_phi2(x):
    """x^2 e^x / expm1(x)^2 with its series below |x| = 1e-6 and 0 from 700, fcs.py:59-69."""
    x = np.asarray(x, dtype=float)
    out = np.zeros_like(x)
    ax = np.abs(x)
    small = ax < 1e-6
    mid = (ax >= 1e-6) & (x < 700.0)
    xs = x[small]
    out[small] = 1.0 - (xs * xs) / 12.0 + (xs ** 4) / 240.0
    em1 = np.expm1(x[mid])
    out[mid] = (x[mid] * x[mid]) * (em1 + 1.0) / (em1 * em1)
    return out


def _fcs_ab(omega_rad_s, temp):
    """A = k_B T phi1(x), B = (k_B T)^2 phi2(x), x = hbar w / k_B T, fcs.py:72-86."""
    if temp <= 0.0:
        raise ValueError("Temperature must be positive in Kelvin.")
    x = (HBAR * omega_rad_s) / (KB * temp)
    s = KB * temp
    return s * fcs_phi1(x), (s * s) * fcs_phi2(x)


def fcs_cumulants(omega_rad_s, t_mean, t2_mean, temp_left, temp_right):
    """(C1, C2): C1 = (1/2pi) int <T> (A_L - A_R), C2 = (1/2pi) int <T>(B_L + B_R)
    + (<T> + <T^2>)(A_L - A_R)^2 (fcs.py:125-156); with t2_mean=None the single-spectrum
    form T(B_L + B_R) + T(1 + T)(A_L - A_R)^2 (fcs.py:89-122)."""
    omega = _fcs_grid(omega_rad_s, "omega_rad_s")
    t1 = np.asarray(list(t_mean), dtype=float)
    if t2_mean is None:
        if t1.shape != omega.shape:
            raise ValueError("transmission_vals must have the same length as omega_rad_s.")
        if np.any(t1 < 0.0):
            raise ValueError("transmission_vals must be non-negative.")
    else:
        t2 = np.asarray(list(t2_mean), dtype=float)
        if t1.shape != omega.shape or t2.shape != omega.shape:
            raise ValueError("t_mean_vs_omega and t2_mean_vs_omega must match omega_rad_s shape.")
        if np.any(t1 < 0.0) or np.any(t2 < 0.0):
            raise ValueError("Transmission moments must be non-negative.")
    a_l, b_l = _fcs_ab(omega, temp_left)
    a_r, b_r = _fcs_ab(omega, temp_right)
    ad = a_l - a_r
    pref = 1.0 / (2.0 * np.pi)
    c1 = pref * trapezoid(t1 * ad, omega)
    if t2_mean is None:
        c2 = pref * trapezoid(t1 * (b_l + b_r) + t1 * (1.0 + t1) * (ad * ad), omega)
    else:
        c2 = pref * trapezoid(t1 * (b_l + b_r) + (t1 + t2) * (ad * ad), omega)
    return float(c1), float(c2)
