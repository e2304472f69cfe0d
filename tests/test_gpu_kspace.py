"""GPU k-space block builder (k_kspace_blocks) and the batched k-averaged solves it feeds.

The builder's blocks must equal the reference builder's bit for bit (the vectors
the reference itself wrote), and transmission_kavg / transmission_spectrum on
force-constant models must give the reference's k-averaged T whether the blocks
are built on the GPU in one launch or resolved per k on the host.
"""

import numpy as np
import pytest

from negfpy_amd.core import transmission_kavg, transmission_spectrum
from negfpy_amd.core.kspace import kspace_blocks
from negfpy_amd.core.negf import kspace_sets
from negfpy_amd.core.types import DeviceKSpace, LeadKSpace
from negfpy_amd.models import MaterialKspaceParams, material_kspace_device, material_kspace_lead
from oracle import negf_oracle as o
from parity_util import T_TOL, oracle_spread
from test_reference_vectors import MK_K, REF, material_params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tag", ["h", "x"])
def test_gpu_builder_equals_reference_blocks(gpu, tag):
    mp = material_params(MaterialKspaceParams, explicit_10=(tag == "x"))
    kb = kspace_blocks(mp, MK_K, lead=True, n_layers=4, layer_masses=REF["mk_layer_masses"])
    np.testing.assert_array_equal(kb.lead.cpu().numpy(), REF[f"mk_{tag}_lead"])
    np.testing.assert_array_equal(kb.onsite.cpu().numpy(), REF[f"mk_{tag}_onsite"])
    np.testing.assert_array_equal(kb.coupling.cpu().numpy(), REF[f"mk_{tag}_coupling"])


def test_gpu_builder_many_k_matches_host_builder(gpu):
    """512 k points on a (ky, kz) mesh, uniform layer masses, one layer and many layers."""
    mp = material_params(MaterialKspaceParams)
    ky, kz = np.meshgrid(np.linspace(-np.pi, np.pi, 32), np.linspace(-np.pi, np.pi, 16))
    kpts = list(zip(ky.ravel().tolist(), kz.ravel().tolist()))
    for nl in (1, 7):
        kb = kspace_blocks(mp, kpts, lead=True, n_layers=nl)
        ml, md = material_kspace_lead(mp), material_kspace_device(nl, mp)
        lead = kb.lead.cpu().numpy()
        ons = kb.onsite.cpu().numpy()
        cpl = kb.coupling.cpu().numpy()
        assert cpl.shape == (len(kpts), nl - 1, 6, 6)
        for i in (0, 1, 100, 257, 511):
            np.testing.assert_array_equal(lead[i], np.stack(ml.blocks_builder(kpts[i])))
            d = md.device_builder(kpts[i])
            np.testing.assert_array_equal(ons[i], np.stack(d.onsite_blocks))
            if nl > 1:
                np.testing.assert_array_equal(cpl[i], np.stack(d.coupling_blocks))


def _untagged(lead, dev):
    """The same providers without the builder tag: forces per-k host resolution."""
    return (LeadKSpace(blocks_builder=lambda k: lead.blocks_builder(k)),
            DeviceKSpace(device_builder=lambda k: dev.device_builder(k)))


def test_kavg_on_gpu_built_blocks_matches_reference(gpu):
    mp = material_params(MaterialKspaceParams)
    ml = material_kspace_lead(mp)
    md = material_kspace_device(4, mp, layer_masses=REF["mk_layer_masses"])
    assert kspace_sets(md, ml, ml, MK_K) is not None
    ul, ud = _untagged(ml, md)
    assert kspace_sets(ud, ul, ul, MK_K) is None
    ons, cps, lead = REF["mk_h_onsite"], REF["mk_h_coupling"], REF["mk_h_lead"]
    for w, t_ref in zip(REF["mk_omegas"], REF["mk_t_kavg"]):
        t_gpu = transmission_kavg(w, md, ml, ml, kpoints=MK_K, eta=1e-6)
        t_host = transmission_kavg(w, ud, ul, ul, kpoints=MK_K, eta=1e-6)
        sp = max(oracle_spread(w, list(ons[i]), list(cps[i]), tuple(lead[i]), tuple(lead[i]),
                               0.0, eta=1e-6) for i in range(len(MK_K)))
        bound = max(T_TOL * max(abs(t_ref), 1e-3), 50.0 * sp)
        assert abs(t_gpu - t_ref) <= bound, (w, t_gpu, t_ref)
        assert abs(t_gpu - t_host) <= 1e-13 * max(abs(t_ref), 1e-3)


def test_spectrum_on_gpu_built_blocks(gpu):
    mp = material_params(MaterialKspaceParams)
    ml = material_kspace_lead(mp)
    md = material_kspace_device(4, mp, layer_masses=REF["mk_layer_masses"])
    ul, ud = _untagged(ml, md)
    w = np.linspace(0.05, 0.6, 9)
    t_gpu = transmission_spectrum(w, md, ml, ml, kpoints=MK_K, eta=1e-6)
    t_host = transmission_spectrum(w, ud, ul, ul, kpoints=MK_K, eta=1e-6)
    np.testing.assert_allclose(t_gpu, t_host, rtol=1e-12, atol=1e-15)
    ons, cps, lead = REF["mk_h_onsite"], REF["mk_h_coupling"], REF["mk_h_lead"]
    for a in (0, 4, 8):
        t_or = np.mean([o.transmission_sparse(w[a], list(ons[i]), list(cps[i]), tuple(lead[i]),
                                              tuple(lead[i]), eta=1e-6)
                        for i in range(len(MK_K))])
        assert abs(t_gpu[a] - t_or) <= 1e-9 * max(abs(t_or), 1e-3)


def test_builder_argument_errors(gpu):
    mp = material_params(MaterialKspaceParams)
    with pytest.raises(ValueError, match="at most two transverse"):
        kspace_blocks(mp, [(0.1, 0.2, 0.3)])
    with pytest.raises(ValueError, match="layer_masses shape"):
        kspace_blocks(mp, [(0.1,)], n_layers=3, layer_masses=np.ones((2, 2)))
    with pytest.raises(ValueError, match="nothing to build"):
        kspace_blocks(mp, [(0.1,)], lead=False)
