"""hipGraph capture of the batched transmission solve (_engine.TransmissionGraph).

Every replay is checked against an eager run of the same inputs, with other eager
work and synchronisations between the replays; inputs updated in place between
replays give the eager result for the new inputs.
"""

import numpy as np
import pytest
import torch

from negfpy_amd.core import _engine
from negfpy_amd.models import MultiDofCellParams, multidof_device, multidof_lead_blocks
from negfpy_amd.workloads import omega_max, ribbon_params
from oracle import negf_oracle as o
from parity_util import T_TOL

pytestmark = pytest.mark.gpu


def _sets(n_atoms=2, nl=5):
    p = ribbon_params(7, n_atoms)
    lead = multidof_lead_blocks(p)
    rng = np.random.default_rng(8)
    dev = multidof_device(nl, p, layer_masses=rng.uniform(10.0, 30.0, (nl, n_atoms)))
    lb = np.stack([lead.d00, lead.d01, np.conj(lead.d01).T])[None]
    n = lead.d00.shape[0]
    sets = _engine.TransportSets.from_arrays(
        lb, lb, np.conj(lead.d01).T[None], lead.d01[None], np.arange(n), np.arange(n),
        np.stack(dev.onsite_blocks)[None], np.stack(dev.coupling_blocks)[None], True)
    return sets, lead, dev


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replays_match_eager(gpu, precision):
    sets, lead, dev = _sets()
    solver = _engine.TransmissionSolver(sets, precision=precision)
    m = 96
    w = np.linspace(0.02, 1.1, m) * omega_max(lead)
    set_of = torch.zeros(m, dtype=torch.int32, device="cuda")
    eta = torch.full((m,), 1e-6, dtype=torch.float64, device="cuda")
    omega = torch.as_tensor(w, device="cuda")
    eager = solver.run(set_of, omega, eta)
    t_eager = eager.t.cpu().numpy().copy()
    st_eager = eager.status.cpu().numpy().copy()
    g = _engine.TransmissionGraph(solver, set_of, omega, eta)
    for r in range(4):
        # unrelated eager work between replays, including another solve
        other = solver.run(set_of, omega * 0.5, eta)
        torch.cuda.synchronize()
        _ = (other.t * 2.0).cpu()
        out = g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(out.status.cpu().numpy(), st_eager), r
        np.testing.assert_array_equal(out.t.cpu().numpy(), t_eager)
    # new inputs written in place
    w2 = np.linspace(0.05, 0.9, m) * omega_max(lead)
    g.omega.copy_(torch.as_tensor(w2, device="cuda"))
    out = g.replay()
    t2 = solver.run(set_of, torch.as_tensor(w2, device="cuda"), eta).t.cpu().numpy()
    np.testing.assert_array_equal(out.t.cpu().numpy(), t2)
    if precision == "fp64":
        tup = (lead.d00, lead.d01, None)
        for k in (3, 40, 77):
            tr = o.transmission_sparse(w2[k], dev.onsite_blocks, dev.coupling_blocks, tup, tup,
                                       eta=1e-6)
            assert abs(t2[k] - tr) <= max(T_TOL * max(abs(tr), 1e-3), 1e-8 * max(abs(tr), 1e-3))


def test_graph_pooled_class(gpu):
    """Blocks of 49..96 DOF (the class whose matrices sit in a device-memory pool) in a
    captured solve: the eager run before capture creates the pool, and every replay
    equals the eager result, with eager solves of the same class on the current stream
    between the replays."""
    sets, lead, dev = _sets(n_atoms=20, nl=3)
    assert sets.n == 60
    solver = _engine.TransmissionSolver(sets)
    m = 48
    w = np.linspace(0.05, 1.05, m) * omega_max(lead)
    set_of = torch.zeros(m, dtype=torch.int32, device="cuda")
    eta = torch.full((m,), 1e-6, dtype=torch.float64, device="cuda")
    omega = torch.as_tensor(w, device="cuda")
    eager = solver.run(set_of, omega, eta)
    t_eager = eager.t.cpu().numpy().copy()
    st_eager = eager.status.cpu().numpy().copy()
    g = _engine.TransmissionGraph(solver, set_of, omega, eta)
    for r in range(3):
        other = solver.run(set_of, omega * 0.7, eta)
        _ = (other.t * 2.0).cpu()
        out = g.replay()
        torch.cuda.synchronize()
        print(f"replay {r}: T[:3] {out.t[:3].tolist()}", flush=True)
        assert np.array_equal(out.status.cpu().numpy(), st_eager), r
        np.testing.assert_array_equal(out.t.cpu().numpy(), t_eager)
    tup = (lead.d00, lead.d01, None)
    for k in (5, 30):
        tr = o.transmission_sparse(w[k], dev.onsite_blocks, dev.coupling_blocks, tup, tup,
                                   eta=1e-6)
        assert abs(t_eager[k] - tr) <= max(T_TOL * max(abs(tr), 1e-3), 1e-8 * max(abs(tr), 1e-3))
