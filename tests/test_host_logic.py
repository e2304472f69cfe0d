"""Host-side logic that needs no GPU: builders, contact handling, layer merging,
omega_scale conventions and argument validation (reference negf.py:119-250,
models/*.py)."""

import numpy as np
import pytest

from negfpy_amd.core import negf as hn
from negfpy_amd.core.surface_gf import canonical_method, pack_lead
from negfpy_amd.core.types import Device1D, LeadBlocks
from negfpy_amd.models import (
    ChainParams,
    CubicLatticeParams,
    MultiDofCellParams,
    SquareLatticeParams,
    analytic_band_max,
    cubic_lattice_device,
    cubic_lattice_lead,
    device_mass_defect,
    device_perfect_chain,
    lead_blocks,
    multidof_device,
    multidof_lead_blocks,
    square_lattice_device,
    square_lattice_lead,
)
from oracle import negf_oracle as o


def test_chain_builders():
    p = ChainParams(mass=2.0, spring=3.0)
    lb = lead_blocks(p)
    assert lb.d00[0, 0] == 3.0 and lb.d01[0, 0] == -1.5 and lb.d10 is None
    d = device_perfect_chain(4, p)
    assert d.n_layers == 4 and len(d.coupling_blocks) == 3
    dd = device_mass_defect(4, p, 2, 6.0)
    assert dd.onsite_blocks[2][0, 0] == 1.0 and dd.coupling_blocks[1][0, 0] == -1.5
    assert analytic_band_max(p) == pytest.approx(2.0 * np.sqrt(1.5))
    with pytest.raises(ValueError):
        device_mass_defect(4, p, 4, 1.0)


def test_multidof_mass_weighting():
    rng = np.random.default_rng(0)
    m = np.array([1.0, 4.0])
    fc00 = rng.normal(size=(6, 6))
    fc01 = rng.normal(size=(6, 6))
    p = MultiDofCellParams(masses=m, fc00=fc00, fc01=fc01, onsite_pinning=0.5)
    lb = multidof_lead_blocks(p)
    s = 1.0 / np.sqrt(np.repeat(m, 3))
    assert np.allclose(lb.d00, np.diag(s) @ (fc00 + 0.5 * np.eye(6)) @ np.diag(s), atol=1e-15)
    assert np.allclose(lb.d01, np.diag(s) @ fc01 @ np.diag(s), atol=1e-15)
    lm = np.array([[1.0, 4.0], [2.0, 3.0], [1.0, 1.0]])
    dev = multidof_device(3, p, layer_masses=lm)
    s1 = 1.0 / np.sqrt(np.repeat(lm[1], 3))
    s2 = 1.0 / np.sqrt(np.repeat(lm[2], 3))
    assert np.allclose(dev.coupling_blocks[1], np.diag(s1) @ fc01 @ np.diag(s2), atol=1e-15)
    with pytest.raises(ValueError):
        multidof_device(3, p, layer_masses=np.ones((2, 2)))


def test_lattice_kspace_builders():
    sp = SquareLatticeParams(mass=2.0, spring_x=1.0, spring_y=0.5)
    d00, d01 = square_lattice_lead(sp).blocks((0.3,))
    assert d00[0, 0] == pytest.approx((2.0 + 1.0 - np.cos(0.3)) / 2.0)
    assert d01[0, 0] == -0.5
    dev = square_lattice_device(5, sp, defect_index=1, defect_mass=4.0).device((0.3,))
    assert dev.onsite_blocks[1][0, 0] == pytest.approx((2.0 + 1.0 - np.cos(0.3)) / 4.0)
    with pytest.raises(ValueError):
        square_lattice_lead(sp).blocks((0.1, 0.2))
    cp = CubicLatticeParams(mass=1.0, spring_x=1.0, spring_y=0.4, spring_z=0.3, onsite_pinning=0.1)
    d00, _ = cubic_lattice_lead(cp).blocks((0.2, 0.5))
    assert d00[0, 0] == pytest.approx(2.0 + 0.8 + 0.6 + 0.1 - 0.8 * np.cos(0.2) - 0.6 * np.cos(0.5))
    assert cubic_lattice_device(3, cp).device(None).n_layers == 3
    with pytest.raises(IndexError):  # the reference indexes kpar[0], kpar[1] of ()
        cubic_lattice_lead(cp).blocks(())


def test_method_aliases():
    assert canonical_method(" SR ") == "sancho_rubio"
    assert canonical_method("gep") == "generalized_eigen"
    assert canonical_method("ge_svd") == "generalized_eigen_svd"
    assert canonical_method("old_svd") == "legacy_eigen_svd"
    with pytest.raises(ValueError, match="Unknown surface GF method"):
        canonical_method("qz")


def test_contact_index_normalisation():
    assert hn._normalize_contact_indices(slice(0, 3), dim=10, default=np.arange(2)).tolist() == [0, 1, 2]
    for bad in ([], [10], [1, 1], [-1]):
        with pytest.raises(ValueError):
            hn._normalize_contact_indices(bad, dim=10, default=np.arange(2))


@pytest.mark.parametrize("idx_l,m", [([0, 1, 2], 2), ([0, 2, 4], 3), ([1], 1)])
def test_layer_merging_preserves_g1n(idx_l, m):
    """Contacts deeper than one layer merge m layers per block (zero-padded tail DOFs
    are decoupled); the contact block of G is unchanged."""
    rng = np.random.default_rng(1)
    n, nl = 2, 7
    ons = [(lambda q: q + q.T)(rng.normal(size=(n, n))) for _ in range(nl)]
    cps = [rng.normal(size=(n, n)) for _ in range(nl - 1)]
    dim = n * nl
    idx_l = np.array(idx_l)
    idx_r = np.array([dim - 2, dim - 1])
    ons_m, cps_m, pos_l, pos_r = hn._layout(ons, cps, idx_l, idx_r, max_block=48)
    assert ons_m.shape[1] == m * n and ons_m.shape[0] == -(-nl // m)
    z = (0.7 + 1e-3j) ** 2
    sl = rng.normal(size=(idx_l.size,) * 2) * 1j
    sr = rng.normal(size=(2, 2)) * 1j
    g_ref = np.linalg.inv(_full(z, ons, cps, idx_l, idx_r, sl, sr))
    big = ons_m.shape[1]
    sl_b = np.zeros((big, big), complex)
    sr_b = np.zeros((big, big), complex)
    sl_b[np.ix_(pos_l, pos_l)] = sl
    sr_b[np.ix_(pos_r, pos_r)] = sr
    g1n = o.rgf_g1n(z, list(ons_m), list(cps_m), sl_b, sr_b)
    assert np.allclose(g1n[np.ix_(pos_l, pos_r)], g_ref[np.ix_(idx_l, idx_r)], atol=1e-12)


def _full(z, ons, cps, idx_l, idx_r, sl, sr):
    from scipy.sparse import identity

    a = (z * identity(len(ons) * ons[0].shape[0]) - o.assemble_device(list(ons), list(cps))).toarray()
    a[np.ix_(idx_l, idx_l)] -= sl
    a[np.ix_(idx_r, idx_r)] -= sr
    return a


def test_resolve_k_conventions():
    d00 = np.array([[2.0, 0.1], [0.1, 2.0]])
    d01 = np.array([[-1.0, 0.2], [0.0, -1.0]])
    lead = LeadBlocks(d00=d00, d01=d01)
    dev = Device1D([d00] * 3, [d01] * 2)
    kp = hn.resolve_k(dev, lead, lead, None, eta=1e-8, eta_device=None, device_to_lead_left=None,
                      device_to_lead_right=None, contact_left_indices=None,
                      contact_right_indices=None, omega_scale=2.0, max_block=48)
    # blocks / s^2, default couplings / s^4 (negf.py:177-204)
    assert np.allclose(kp.onsite[0], d00 / 4.0)
    assert np.allclose(kp.v_l, d01.conj().T / 16.0)
    assert np.allclose(kp.v_r, d01 / 16.0)
    assert np.allclose(kp.lead_l, pack_lead(d00 / 4, d01 / 4))
    assert kp.pos_l.tolist() == [0, 1] and kp.pos_r.tolist() == [0, 1]
    with pytest.raises(ValueError, match="eta_device"):
        hn.resolve_k(dev, lead, lead, None, eta=1e-8, eta_device=-1.0, device_to_lead_left=None,
                     device_to_lead_right=None, contact_left_indices=None,
                     contact_right_indices=None, omega_scale=None, max_block=48)
    with pytest.raises(ValueError, match="omega_scale"):
        hn.resolve_k(dev, lead, lead, None, eta=1e-8, eta_device=None, device_to_lead_left=None,
                     device_to_lead_right=None, contact_left_indices=None,
                     contact_right_indices=None, omega_scale=0.0, max_block=48)
    with pytest.raises(ValueError, match="Left contact index count"):
        hn.resolve_k(dev, lead, lead, None, eta=1e-8, eta_device=None, device_to_lead_left=None,
                     device_to_lead_right=None, contact_left_indices=[0],
                     contact_right_indices=None, omega_scale=None, max_block=48)


def test_block_class_table_on_the_host():
    """The class override table is host state of the library (no GPU call)."""
    from negfpy_amd import _lib

    lib = _lib.load_symbols_only()
    fp64, mixed = 0, 1
    assert lib.negf_block_class(fp64, 20) == 24 and lib.negf_block_class(mixed, 50) == 64
    assert lib.negf_set_block_class(fp64, 20, 32) == 0
    assert lib.negf_block_class(fp64, 20) == 32
    assert lib.negf_block_class(mixed, 20) == 24          # per precision
    assert lib.negf_set_block_class(fp64, 20, 16) != 0    # class smaller than the block
    assert lib.negf_set_block_class(fp64, 20, 64) != 0    # 64 is a mixed-precision class only
    assert lib.negf_set_block_class(fp64, 20, 40) != 0    # not a class
    assert lib.negf_set_block_class(fp64, 97, 0) != 0     # block too large for FP64
    assert lib.negf_block_class(fp64, 60) == 96 and lib.negf_block_class(fp64, 96) == 96
    assert lib.negf_block_class(fp64, 97) == 0
    assert lib.negf_set_block_class(mixed, 20, 96) != 0   # 96 is an FP64 class only
    assert lib.negf_block_class(fp64, 20) == 32           # failed calls change nothing
    assert lib.negf_set_block_class(fp64, 20, 0) == 0
    assert lib.negf_block_class(fp64, 20) == 24


def test_transmission_spectra_rejects_malformed_problems():
    """Each problem is (omegas, device, lead_left, lead_right[, kpoints]); the check
    runs before anything touches the GPU."""
    import pytest

    from negfpy_amd.core import transmission_spectra

    with pytest.raises(ValueError, match="each problem is"):
        transmission_spectra([([1.0], None, None)])
    with pytest.raises(ValueError, match="at least one k-point"):
        transmission_spectra([([1.0], None, None, None, [])])
