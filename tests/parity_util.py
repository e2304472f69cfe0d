"""Shared parity bar for GPU-vs-oracle and GPU-vs-reference-vector tests.

Bar for FP64 mode: |T_gpu - T_ref| <= max(1e-10 max(|T|, 1e-3), 10 s), where
s is the oracle's own rounding-level spread of T at that point: the largest change
of T_oracle when every input block is perturbed by about one ulp (three seeded
draws).  At most points s ~ 1e-14..1e-12 and the 1e-10 bar applies; near band
features decimation is conditioned like 1/eta and s reaches 1e-8, which no FP64
evaluation order can beat.
"""

import numpy as np

from oracle import negf_oracle as o

T_TOL = 1e-10


def rel(a, b, floor=1e-3):
    return np.max(np.abs(np.asarray(a) - np.asarray(b)) / np.maximum(np.abs(b), floor))


def _ulp_perturb(x, rng):
    if x is None:
        return None
    x = np.asarray(x, dtype=np.complex128)
    return x * (1.0 + 2.2e-16 * rng.standard_normal(x.shape))


def oracle_spread(w, ons, cps, ll, lr, t, **kw):
    rng = np.random.default_rng(1234)
    spread = 0.0
    for _ in range(3):
        pl = tuple(_ulp_perturb(b, rng) for b in ll)
        pr = tuple(_ulp_perturb(b, rng) for b in lr)
        po = [_ulp_perturb(b, rng) for b in ons]
        pc = [_ulp_perturb(b, rng) for b in cps]
        spread = max(spread, abs(o.transmission_sparse(w, po, pc, pl, pr, **kw) - t))
    return spread


def oracle_t(w, ons, cps, ll, lr, **kw):
    """(T, spread): oracle transmission and its rounding-level spread."""
    t = o.transmission_sparse(w, ons, cps, ll, lr, **kw)
    return t, oracle_spread(w, ons, cps, ll, lr, t, **kw)


def assert_parity(t_gpu, refs):
    for tg, (tr, sp) in zip(np.atleast_1d(t_gpu), refs):
        bound = max(T_TOL * max(abs(tr), 1e-3), 10.0 * sp)
        assert abs(tg - tr) <= bound, (tg, tr, sp)
