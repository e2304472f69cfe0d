"""Force-constant models for bulk transport (reference ``src/negfpy/modeling``)."""

from .builders import (
    build_fc_terms,
    build_material_kspace_params,
    enforce_translational_asr_on_self_term,
    full_realspace_grid,
    principal_layer_size,
)
from .conditioning import (
    apply_mass_mode,
    apply_transverse_stability_controls,
    drop_nyquist_transverse_terms,
    enforce_transverse_pm_symmetry,
    reorient_for_transport,
    transverse_cutoff,
)
from .contacts import build_interface_contacts
from .schema import BuildConfig, IFCData, IFCTerm
from .units import qe_omega_to_cm1, qe_omega_to_rad_s, qe_omega_to_thz

__all__ = [
    "BuildConfig", "IFCData", "IFCTerm", "apply_mass_mode", "apply_transverse_stability_controls",
    "build_fc_terms", "build_interface_contacts", "build_material_kspace_params",
    "drop_nyquist_transverse_terms", "enforce_translational_asr_on_self_term",
    "enforce_transverse_pm_symmetry", "full_realspace_grid", "principal_layer_size",
    "qe_omega_to_cm1", "qe_omega_to_rad_s", "qe_omega_to_thz", "reorient_for_transport",
    "transverse_cutoff",
]
