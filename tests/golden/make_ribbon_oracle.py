"""Write tests/golden/ribbon24_oracle.npz: the FP64 oracle's T(omega) and its
rounding-level spread at every one of the 4096 frequencies of the ribbon benchmark
workload (24 DOF per layer, 32 layers, eta = 1e-6, seed 0).

The oracle (decimation + sparse LU, oracle/negf_oracle.py) is the checker; the spread
is the largest change of T under three seeded ~1-ulp perturbations of every input
block (tests/parity_util.py), the accuracy no FP64 evaluation order can beat at that
point.  Run from the repository root:  python tests/golden/make_ribbon_oracle.py
"""

import hashlib
import multiprocessing as mp
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from negfpy_amd.workloads import ribbon_problem  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "ribbon24_oracle.npz")
_P = None


def _work(idx):
    from threadpoolctl import threadpool_limits

    from parity_util import oracle_t

    p = _P
    lead = (p.lead.d00, p.lead.d01, None)
    out = []
    with threadpool_limits(limits=1):
        for i in idx:
            out.append(oracle_t(float(p.omegas[i]), p.device.onsite_blocks,
                                p.device.coupling_blocks, lead, lead, eta=p.eta))
    return idx, out


def input_digest(p):
    h = hashlib.sha256()
    for a in (p.lead.d00, p.lead.d01, *p.device.onsite_blocks, *p.device.coupling_blocks, p.omegas):
        h.update(np.ascontiguousarray(a, dtype=np.complex128).tobytes())
    return h.hexdigest()


def main():
    global _P
    _P = ribbon_problem(seed=0, n_atoms=8, n_layers=32, n_omega=4096, eta=1e-6)
    n = _P.omegas.size
    workers = os.cpu_count() or 1
    chunks = [np.arange(k, n, workers) for k in range(workers)]
    t = np.empty(n)
    s = np.empty(n)
    with mp.get_context("fork").Pool(workers) as pool:
        for idx, vals in pool.imap_unordered(_work, chunks):
            for i, (tv, sv) in zip(idx, vals):
                t[i], s[i] = tv, sv
    np.savez_compressed(OUT, omega=_P.omegas, t=t, spread=s, digest=np.array(input_digest(_P)))
    print(f"wrote {OUT}: {n} points, max spread/max(|T|,1e-3) "
          f"{np.max(s / np.maximum(np.abs(t), 1e-3)):.2e}")


if __name__ == "__main__":
    main()
