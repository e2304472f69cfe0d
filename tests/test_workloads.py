"""The synthetic crystals of the benchmark workloads: the silicon-like diamond cells of
the 3-D uneven batch against what the structure fixes (coordination, acoustic sum
rule, three acoustic modes at Gamma, a Hermitian positive semi-definite dynamical
matrix at every k)."""

import numpy as np
import pytest

from negfpy_amd import workloads as w


@pytest.mark.parametrize("cells,n", [((1, 1, 1), 24), ((2, 1, 1), 48), ((2, 2, 1), 96)])
def test_diamond_cells(cells, n):
    phi = w._diamond_phi(cells, seed=1, jitter=0.05)
    assert phi[(0, 0, 0)].shape == (n, n)
    assert {s[0] for s in phi} == {-1, 0, 1}  # only adjacent layers couple
    total = sum(phi.values())
    for a in range(n // 3):  # acoustic sum rule on every atom
        assert np.abs(total[3 * a:3 * a + 3].reshape(3, -1, 3).sum(axis=1)).max() < 1e-12
    for s, blk in phi.items():  # Phi(-R) = Phi(R)^T
        np.testing.assert_allclose(phi[tuple(-np.array(s))], blk.T, atol=1e-13)
    rng = np.random.default_rng(0)
    for q in [np.zeros(3)] + [rng.uniform(-np.pi, np.pi, 3) for _ in range(3)]:
        dq = sum(v * np.exp(1j * np.dot(q, s)) for s, v in phi.items())
        np.testing.assert_allclose(dq, dq.conj().T, atol=1e-12)
        ev = np.linalg.eigvalsh(dq)
        assert ev.min() > -1e-12
        if not q.any():
            assert int(np.sum(np.abs(ev) < 1e-10)) == 3


def test_diamond_coordination():
    pos = np.array([(f[0] + o, f[1] + o, f[2] + o) for o in (0.0, 0.25) for f in w._FCC])
    for a in range(8):
        d = [np.linalg.norm(pos[b] + np.array(s) - 1 - pos[a])
             for b in range(8) for s in np.ndindex(3, 3, 3)]
        assert sum(abs(x - np.sqrt(3) / 4) < 1e-12 for x in d) == 4
        assert sum(abs(x - np.sqrt(2) / 2) < 1e-12 for x in d) == 12


def test_uneven3d_parts_are_diamond():
    ps = w.uneven3d_problems(n_layers=2, n_omega=8)
    assert [p.params.ndof for p in ps] == [24, 48, 96]
    assert [len(p.kpoints) for p in ps] == [1024, 64, 16]
    assert all(np.all(p.params.masses == 28.0) for p in ps)
