"""One persistent launch over problems of different block sizes
(negf_transmission_multi_batched, _engine.transmission_multi).

Every output of every problem (T, status, decimation step counts, surface DOS)
must equal what the staged launch sequence (negf_transmission_batched) gives for
that problem alone, bit for bit: the persistent kernel runs the same per-stage
arithmetic, and each instance of a ticket stops decimating at its own step.  The
problems cover every block-size class (8, 16, 24, 32, 48 in LDS and 96 in the
pool), tickets cut short at the end of a problem, frequencies clustered at band
edges (so the instances of one ticket need different step counts), failures
(iteration limit, eigenmode fallback), two different leads with explicit
couplings on contact subsets, more problems than one launch takes, hipGraph
replays and mixed precision (surface GFs by the staged mixed decimation, the rest in
the persistent launch).  A few points per problem are also checked against the oracle.
"""

import numpy as np
import pytest
import torch

from negfpy_amd import _lib
from negfpy_amd.core import _engine
from negfpy_amd.workloads import omega_max, random_lead
from oracle import negf_oracle as o
from parity_util import assert_parity, oracle_t

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def classic_kernels(gpu):
    """The persistent launch runs the workgroup-per-instance stage bodies, so the staged
    launches it is compared with bit for bit are that family too (the one-wave MFMA
    kernels agree with them to rounding only; tests/test_gpu_wave_kernels.py)."""
    lib = _lib.load()
    prev = lib.negf_set_wave_kernels(0)
    yield
    lib.negf_set_wave_kernels(prev)


def _device(seed, n, n_layers):
    d00, d01 = random_lead(seed, n)
    rng = np.random.default_rng(seed + 7)
    ons = [d00 + np.diag(rng.uniform(-0.05, 0.05, n)) for _ in range(n_layers)]
    cps = [d01 * (1.0 + 0.1 * rng.standard_normal()) for _ in range(n_layers - 1)]
    return d00, d01, ons, cps


def _problem(seed, n, n_layers, n_w, *, maxiter=200, route=0, edges=True):
    """A shared-lead problem with n_w frequencies, half of them packed at the band top."""
    d00, d01, ons, cps = _device(seed, n, n_layers)
    lb = np.stack([d00, d01, d01.conj().T])[None]
    sets = _engine.TransportSets.from_arrays(
        lb, lb, d01.conj().T[None], d01[None], np.arange(n), np.arange(n), np.stack(ons)[None],
        np.stack(cps)[None] if n_layers > 1 else None, True)
    wm = omega_max(_engine_lead(d00, d01))
    w = np.linspace(0.03, 1.05, n_w) * wm
    if edges:
        w[1::2] = wm * (1.0 + np.linspace(-0.01, 0.01, w[1::2].size))
    solver = _engine.TransmissionSolver(sets, maxiter=maxiter, surface_route=route)
    return solver, w, (d00, d01, ons, cps)


def _engine_lead(d00, d01):
    from negfpy_amd.core import LeadBlocks

    return LeadBlocks(d00=d00, d01=d01)


def _job(solver, w, eta=1e-6):
    m = w.size
    return (solver, torch.zeros(m, dtype=torch.int32, device="cuda"),
            torch.as_tensor(w, device="cuda"), torch.full((m,), eta, dtype=torch.float64,
                                                          device="cuda"), None)


def _host(res):
    return {k: getattr(res, k).cpu().numpy().copy()
            for k in ("t", "status", "iters_l", "iters_r", "dos_l", "dos_r")}


def _assert_same(multi, staged, tag):
    for k in staged:
        np.testing.assert_array_equal(multi[k], staged[k], err_msg=f"{tag}: {k}")


# n -> class: 6 -> 8, 12 -> 16, 20 -> 24, 30 -> 32, 40 -> 48, 60 and 96 -> pool
SIZES = [(6, 4, 30), (12, 3, 23), (20, 5, 9), (30, 2, 7), (40, 3, 5), (60, 2, 3), (96, 2, 2)]


def test_mixed_classes_equal_staged(gpu):
    probs = [_problem(1000 + n, n, nl, nw) for n, nl, nw in SIZES]
    jobs = [_job(s, w) for s, w, _ in probs]
    staged = [_host(s.run(*j[1:])) for (s, _, _), j in zip(probs, jobs)]
    multi = [_host(r) for r in _engine.transmission_multi(jobs)]
    for (n, nl, nw), a, b in zip(SIZES, multi, staged):
        print(f"n={n} L={nl}: steps {b['iters_l'].min()}..{b['iters_l'].max()}, "
              f"T[0]={b['t'][0]:.6e}", flush=True)
        _assert_same(a, b, f"n={n}")
    # uneven step counts inside one ticket (the 24 and 16 classes hold 4 and 11)
    assert np.ptp(staged[2]["iters_l"][:4]) > 0 and np.ptp(staged[1]["iters_l"][:11]) > 0
    # and the oracle at a few points of each problem
    for (s, w, (d00, d01, ons, cps)), a in zip(probs, multi):
        tup = (d00, d01, None)
        for k in (0, w.size // 2):
            ref = oracle_t(w[k], ons, cps, tup, tup, eta=1e-6)
            assert_parity([a["t"][k]], [ref])


def test_failures_and_eigen_fallback(gpu):
    """Iteration limit 4: most decimations fail (status 1, T = NaN) in both paths; with
    the eigenmode route those instances are recovered the same way in both."""
    for route in (_lib.SURFACE_DECIMATION, _lib.SURFACE_EIGEN):
        probs = [_problem(2000 + n, n, 3, 13, maxiter=4, route=route) for n in (8, 24, 40)]
        jobs = [_job(s, w) for s, w, _ in probs]
        staged = [_host(s.run(*j[1:])) for (s, _, _), j in zip(probs, jobs)]
        multi = [_host(r) for r in _engine.transmission_multi(jobs)]
        for a, b in zip(multi, staged):
            _assert_same(a, b, f"route={route}")
        nfail = sum(int(np.sum(b["status"] != 0)) for b in staged)
        print(f"route={route}: {nfail} failed instances", flush=True)
        if route == _lib.SURFACE_DECIMATION:
            assert nfail > 0 and all(np.isnan(b["t"][b["status"] != 0]).all() for b in staged)


def test_two_leads_contact_subsets(gpu):
    """Different left/right leads (two decimations per instance), explicit couplings
    into contact subsets of the end layers."""
    jobs, probs = [], []
    for n, nl, seed in ((16, 3, 31), (24, 4, 32), (72, 2, 33)):
        d00, d01, ons, cps = _device(seed, n, nl)
        e00, e01 = random_lead(seed + 50, n)
        rng = np.random.default_rng(seed)
        nc = n - 2
        vl = rng.standard_normal((nc, n)) * 0.2 + 1j * rng.standard_normal((nc, n)) * 0.01
        vr = rng.standard_normal((nc, n)) * 0.2
        sets = _engine.TransportSets.from_arrays(
            np.stack([d00, d01, d01.conj().T])[None], np.stack([e00, e01, e01.conj().T])[None],
            vl[None], vr[None], np.arange(1, n - 1), np.arange(0, nc), np.stack(ons)[None],
            np.stack(cps)[None], False)
        s = _engine.TransmissionSolver(sets)
        w = np.linspace(0.05, 0.95, 17) * min(omega_max(_engine_lead(d00, d01)),
                                              omega_max(_engine_lead(e00, e01)))
        probs.append(s)
        jobs.append(_job(s, w))
    staged = [_host(s.run(*j[1:])) for s, j in zip(probs, jobs)]
    multi = [_host(r) for r in _engine.transmission_multi(jobs)]
    for a, b in zip(multi, staged):
        _assert_same(a, b, "two leads")
        assert np.all(b["status"] == 0)


def test_more_problems_than_one_launch(gpu):
    """11 problems: two persistent launches; also problems with 0 and 1 instances."""
    probs = [_problem(3000 + i, 4 + 3 * i, 2, 1 + 3 * i, edges=False) for i in range(11)]
    jobs = [_job(s, w) for s, w, _ in probs]
    jobs[4] = _job(probs[4][0], np.zeros(0))
    staged = [_host(s.run(*j[1:])) for (s, _, _), j in zip(probs, jobs)]
    multi = [_host(r) for r in _engine.transmission_multi(jobs)]
    for a, b in zip(multi, staged):
        _assert_same(a, b, "11 problems")
    assert multi[4]["t"].size == 0 and multi[0]["t"].size == 1


def test_graph_replays(gpu):
    probs = [_problem(4000 + n, n, 3, 21) for n in (6, 24, 48)]
    jobs = [_job(s, w) for s, w, _ in probs]
    eager = [_host(r) for r in _engine.transmission_multi(jobs)]
    g = _engine.MultiTransmissionGraph(jobs)
    for r in range(4):
        other = _engine.transmission_multi([_job(s, 0.5 * w) for s, w, _ in probs])
        torch.cuda.synchronize()
        _ = [(x.t * 2.0).cpu() for x in other]
        outs = [_host(x) for x in g.replay()]
        torch.cuda.synchronize()
        for a, b in zip(outs, eager):
            _assert_same(a, b, f"replay {r}")
    w2 = [0.7 * w for _, w, _ in probs]
    for gw, w in zip(g.omega, w2):
        gw.copy_(torch.as_tensor(w, device="cuda"))
    outs = [_host(x) for x in g.replay()]
    ref = [_host(r) for r in _engine.transmission_multi([_job(s, w) for (s, _, _), w in
                                                          zip(probs, w2)])]
    for a, b in zip(outs, ref):
        _assert_same(a, b, "new inputs")


def _mirrored(seed, n, n_layers, n_w):
    """Interface-style contacts: the right lead is the left one mirrored."""
    d00, d01, ons, cps = _device(seed, n, n_layers)
    ll = np.stack([d00, d01, d01.conj().T])[None]
    lr = np.stack([d00, d01.conj().T, d01])[None]
    sets = _engine.TransportSets.from_arrays(
        ll, lr, d01.conj().T[None], d01[None], np.arange(n), np.arange(n), np.stack(ons)[None],
        np.stack(cps)[None], False)
    assert sets.mirror_leads
    w = np.linspace(0.03, 1.05, n_w) * omega_max(_engine_lead(d00, d01))
    return sets, w


def test_mixed_precision_equals_staged(gpu):
    """Mixed-precision problems (shared and mirrored leads, classes 8-48) next to an FP64
    one: every output equals the problem's own staged solve bit for bit (the staged
    mixed decimation feeds the persistent FP64 stages), eta = 1e-6 and 1e-2 omega (where
    the FP32 decimation is certified); T within 1e-6 of the oracle."""
    for rel_eta in (None, 1e-2):
        jobs, solvers, refs = [], [], []
        for k, (n, nl, nw) in enumerate(((6, 3, 17), (12, 3, 13), (20, 3, 9), (30, 2, 7),
                                         (40, 2, 5))):
            s, w, blocks = _problem(6000 + n, n, nl, nw, edges=False)
            sm = _engine.TransmissionSolver(s.sets, precision="fp64" if k == 2 else "mixed")
            eta = 1e-6 if rel_eta is None else rel_eta * w
            job = list(_job(sm, w))
            job[3] = torch.as_tensor(np.broadcast_to(eta, w.shape).copy(), device="cuda")
            jobs.append(tuple(job))
            solvers.append(sm)
            refs.append((w, eta, blocks))
        sets, w = _mirrored(6100, 16, 3, 11)
        sm = _engine.TransmissionSolver(sets, precision="mixed")
        eta = 1e-6 if rel_eta is None else rel_eta * w
        job = list(_job(sm, w))
        job[3] = torch.as_tensor(np.broadcast_to(eta, w.shape).copy(), device="cuda")
        jobs.append(tuple(job))
        staged = [_host(j[0].run(*j[1:])) for j in jobs]
        multi = [_host(r) for r in _engine.transmission_multi(jobs)]
        for a, b in zip(multi, staged):
            _assert_same(a, b, f"mixed, eta/omega={rel_eta}")
            assert np.all(b["status"] == 0)
        for (w, eta, (d00, d01, ons, cps)), a in zip(refs, multi):
            tup = (d00, d01, None)
            for k in (0, w.size // 2, w.size - 1):
                e = eta if np.ndim(eta) == 0 else eta[k]
                ref = o.transmission_sparse(w[k], ons, cps, tup, tup, eta=e)
                assert abs(a["t"][k] - ref) <= 1e-6 * max(abs(ref), 1e-3)


def test_mixed_precision_with_ladder_rejected(gpu):
    s, w, _ = _problem(5000, 8, 2, 4)
    sm = _engine.TransmissionSolver(s.sets, precision="mixed")
    lad = _engine.EtaLadder([(1e-6, 1e-4)] * w.size)
    job = _job(sm, w)
    with pytest.raises(ValueError, match="FP64 only"):
        _engine.transmission_multi([job[:5] + (lad,)])


def test_transmission_spectra_matches_single_spectra(gpu):
    """Public entry: three systems of different block sizes (a 1-DOF chain, a 24-DOF
    ribbon, a k-resolved square lattice) in one call equal their own
    transmission_spectrum calls bit for bit, and the chain its closed form."""
    from negfpy_amd.core import transmission_spectra, transmission_spectrum
    from negfpy_amd.models import (ChainParams, SquareLatticeParams, device_perfect_chain,
                                   lead_blocks, square_lattice_device, square_lattice_lead)
    from negfpy_amd.workloads import ribbon_problem

    cp = ChainParams(mass=1.0, spring=1.0)
    chain = (np.linspace(0.05, 2.3, 40), device_perfect_chain(6, cp), lead_blocks(cp),
             lead_blocks(cp))
    rp = ribbon_problem(0, n_atoms=8, n_layers=6, n_omega=64)
    ribbon = (rp.omegas, rp.device, rp.lead, rp.lead)
    sp = SquareLatticeParams(mass=1.0, spring_x=1.0, spring_y=0.5)
    ks = [(ky,) for ky in np.linspace(-np.pi, np.pi, 5, endpoint=False)]
    square = (np.linspace(0.1, 2.5, 25), square_lattice_device(4, sp), square_lattice_lead(sp),
              square_lattice_lead(sp), ks)
    got = transmission_spectra([chain, ribbon, square], eta=1e-6)
    for g, prob in zip(got, (chain, ribbon, square)):
        kw = {"kpoints": prob[4]} if len(prob) == 5 else {}
        ref = transmission_spectrum(prob[0], *prob[1:4], eta=1e-6, **kw)
        np.testing.assert_array_equal(g, ref)
    inside = chain[0] < 1.99
    np.testing.assert_allclose(got[0][inside], 1.0, atol=1e-4)
    assert np.all(np.abs(got[0][chain[0] > 2.01]) < 1e-4)


def _spectrum_both_ways(monkeypatch, *args, **kw):
    """The adaptive spectrum with the eta ladders inside the persistent launch, and
    with the host's rung-by-rung loop (the default)."""
    from negfpy_amd.core import negf as nm

    host = nm.transmission_kavg_adaptive_spectrum(*args, **kw)
    with monkeypatch.context() as m:
        m.setattr(nm._KSolver, "in_kernel_ladders", True)
        calls = []
        orig = nm._KSolver.solve_ladder
        m.setattr(nm._KSolver, "solve_ladder",
                  lambda self, *a, **k: calls.append(1) or orig(self, *a, **k))
        in_kernel = nm.transmission_kavg_adaptive_spectrum(*args, **kw)
        assert calls, "the in-launch ladder was not used"
    return in_kernel, host


def test_eta_ladder_in_kernel_equals_rung_loop(gpu, monkeypatch):
    """Per-k adaptive eta: the ladder walked inside the persistent launch gives the
    rung-by-rung driver's values, statistics and rejection records exactly.  The
    ladder starts at eta = 0, where decimation inside the band does not converge
    (exception, retried), a channel limit below one channel rejects the in-band
    points of the 1-DOF lattice (over_channel_limit at every rung), and k points are
    resolved on the host (callbacks), with and without omega_scale and a fixed
    device eta."""
    from negfpy_amd.models import SquareLatticeParams, square_lattice_device, square_lattice_lead
    from negfpy_amd.workloads import ribbon_problem

    sp = SquareLatticeParams(mass=1.0, spring_x=1.0, spring_y=0.5)
    ks = [(ky,) for ky in np.linspace(-np.pi, np.pi, 6, endpoint=False)]
    lead, dev = square_lattice_lead(sp), square_lattice_device(5, sp)
    w = np.linspace(0.1, 2.6, 23)
    cases = [
        dict(eta_values=(0.0, 1e-9, 1e-6), collect_rejected=True),
        dict(eta_values=(0.0, 1e-7), collect_rejected=True, max_channel_factor=0.5,
             min_success_fraction=0.3),
        dict(eta_values=(1e-8, 1e-6), eta_device=1e-9, omega_scale=1.7, collect_rejected=True),
    ]
    for kw in cases:
        a, b = _spectrum_both_ways(monkeypatch, w, dev, lead, lead, ks, **kw)
        _same_spectrum(a, b, kw)
    # a multi-DOF spectrum, one k point (kpar = None)
    rp = ribbon_problem(0, n_atoms=4, n_layers=6, n_omega=40)
    a, b = _spectrum_both_ways(monkeypatch, rp.omegas, rp.device, rp.lead, rp.lead, [None],
                               eta_values=(0.0, 1e-8, 1e-6), collect_rejected=True)
    _same_spectrum(a, b, "ribbon")


def _same_spectrum(a, b, tag):
    import json

    np.testing.assert_array_equal(a[0], b[0], err_msg=str(tag))
    dump = lambda r: json.dumps(r, sort_keys=True, default=str)  # NaN-safe comparison
    assert dump(a[1]) == dump(b[1]), tag
    n_rej = sum(len(r.get("info", {}).get("rejected", [])) for r in a[1] if isinstance(r, dict))
    print(f"{tag}: {n_rej} rejected attempts", flush=True)


def test_eta_ladder_every_class_equals_staged_rungs(gpu):
    """Engine-level ladders on 20-, 40- and 60-DOF problems (classes 24, 48 and the
    pool) with rungs that fail (eta = 0 inside the band: no convergence within the
    iteration limit) and a channel limit: each instance's accepted rung and the T and
    status of every attempted rung equal the staged solves of those rungs, bit for
    bit."""
    ladders_of = lambda m: [((0.0, 1e-8, 1e-6) if i % 3 else (1e-7, 0.0, 1e-6)) for i in range(m)]
    jobs, probs = [], []
    for n, nl, nw, seed in ((20, 3, 9, 61), (40, 2, 5, 62), (60, 2, 3, 63)):
        s, w, _ = _problem(seed, n, nl, nw, maxiter=60)
        lads = ladders_of(nw)
        t_max = [(2.5 if i % 2 else None) for i in range(nw)]
        lad = _engine.EtaLadder(lads, t_max=t_max, nonneg_tol=0.0)
        jobs.append((s, torch.zeros(nw, dtype=torch.int32, device="cuda"),
                     torch.as_tensor(w, device="cuda"), None, None, lad))
        probs.append((s, w, lads, t_max))
    res = _engine.transmission_multi(jobs)
    for (s, w, lads, t_max), r in zip(probs, res):
        rung = r.ladder.rung.cpu().numpy()
        t_r = r.ladder.t.cpu().numpy()
        st_r = r.ladder.status.cpu().numpy()
        # the staged solve of every rung for every instance
        stage = []
        for q in range(3):
            eta = np.array([lad[q] for lad in lads])
            o = s.run(torch.zeros(w.size, dtype=torch.int32, device="cuda"),
                      torch.as_tensor(w, device="cuda"), torch.as_tensor(eta, device="cuda"))
            stage.append((o.t.cpu().numpy(), o.status.cpu().numpy()))
        n_retried = 0
        for i in range(w.size):
            want = 3
            for q in range(3):
                t, st = stage[q][0][i], stage[q][1][i]
                assert st_r[i, q] == st, (i, q)
                np.testing.assert_array_equal(t_r[i, q], t)
                lim = np.inf if t_max[i] is None else t_max[i]
                if st == 0 and np.isfinite(t) and t >= 0.0 and t <= lim:
                    want = q
                    break
            assert rung[i] == want, (i, rung[i], want)
            n_retried += int(want > 0)
            assert np.all(st_r[i, min(want, 2) + 1:] == -1)   # rungs after the accepted one
        print(f"n={s.sets.n}: {n_retried} of {w.size} instances retried", flush=True)
        assert n_retried > 0


def test_uneven_batch_equals_staged_repeatedly(gpu):
    """The uneven batch (4096 x 24, 1024 x 48, 256 x 96 DOF; long 96-DOF tickets run
    beside short ones) equals the staged launches bit for bit on every one of several
    calls: consecutive pipeline stages of a ticket reuse the same LDS matrices, and a
    stage may clear them only after the previous stage's last reads (a race that showed
    as a few wrong T values per call under this load)."""
    from negfpy_amd.workloads import uneven_problems

    jobs = []
    for p in uneven_problems():
        lead, dev = p.lead, p.device
        lb = np.stack([lead.d00, lead.d01, np.conj(lead.d01).T])[None]
        n = lead.d00.shape[0]
        sets = _engine.TransportSets.from_arrays(
            lb, lb, np.conj(lead.d01).T[None], lead.d01[None], np.arange(n), np.arange(n),
            np.stack(dev.onsite_blocks)[None], np.stack(dev.coupling_blocks)[None], True)
        jobs.append(_job(_engine.TransmissionSolver(sets), p.omegas, p.eta))
    staged = [_host(j[0].run(*j[1:])) for j in jobs]
    for rep in range(4):
        multi = [_host(r) for r in _engine.transmission_multi(jobs)]
        for k, (a, b) in enumerate(zip(multi, staged)):
            _assert_same(a, b, f"call {rep}, problem {k}")
