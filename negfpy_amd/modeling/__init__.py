"""Force-constant models for bulk transport (reference ``src/negfpy/modeling``)."""

from .atomic_masses import ATOMIC_MASS_AMU, get_atomic_mass_amu
from .builders import (
    build_fc_terms,
    build_material_kspace_params,
    build_transport_components,
    enforce_translational_asr_on_self_term,
    full_realspace_grid,
    infer_principal_layer_size,
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
from .contacts import build_interface_contacts, interface_contacts
from .schema import (
    BuildConfig,
    IFCData,
    IFCTerm,
    InterfaceData,
    validate_ifc_data,
    validate_material_kspace_params,
    validate_transport_connectivity,
)
from .units import qe_ev_to_omega, qe_omega_to_cm1, qe_omega_to_ev, qe_omega_to_rad_s, qe_omega_to_thz

__all__ = [
    "ATOMIC_MASS_AMU", "BuildConfig", "IFCData", "IFCTerm", "InterfaceData", "apply_mass_mode",
    "apply_transverse_stability_controls", "build_fc_terms", "build_interface_contacts",
    "build_material_kspace_params", "build_transport_components", "drop_nyquist_transverse_terms",
    "enforce_translational_asr_on_self_term", "enforce_transverse_pm_symmetry",
    "full_realspace_grid", "get_atomic_mass_amu", "infer_principal_layer_size",
    "interface_contacts", "principal_layer_size", "qe_ev_to_omega", "qe_omega_to_cm1",
    "qe_omega_to_ev", "qe_omega_to_rad_s", "qe_omega_to_thz", "reorient_for_transport",
    "transverse_cutoff", "validate_ifc_data", "validate_material_kspace_params",
    "validate_transport_connectivity",
]
