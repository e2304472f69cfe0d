"""GPU phonon-NEGF transport core with the API of the reference ``negfpy.core``
(``src/negfpy/core/__init__.py:1-78``, transport part)."""

from .adjoint import (
    thermal_conductance_torch,
    transmission_sets_torch,
    transmission_sum_torch,
    transmission_torch,
)
from .fcs import (
    HeatCurrentCumulants,
    HeatCurrentUncertainty,
    heat_current_cumulants_from_k_moments,
    heat_current_cumulants_from_spectrum,
    heat_current_cumulants_scan,
    heat_current_uncertainty,
)
from .dos import (
    lead_surface_dos,
    lead_surface_dos_kavg,
    lead_surface_dos_kavg_adaptive,
    lead_surface_dos_spectrum,
    leads_surface_dos_spectrum,
)
from .negf import (
    device_green_function,
    set_eta_ladders_in_kernel,
    transmission,
    transmission_kavg,
    transmission_kavg_adaptive,
    transmission_kavg_adaptive_global_eta,
    transmission_kavg_adaptive_spectrum,
    transmission_k_moments_spectrum,
)
from .surface_gf import (
    get_default_precision,
    set_default_precision,
    surface_gf,
    surface_gf_generalized_eigen,
    surface_gf_generalized_eigen_svd,
    surface_gf_legacy_eigen_svd,
    surface_gf_sancho_rubio,
)
from .thermal import qe_omega_to_rad_s, thermal_conductance_from_spectrum
from .thermal_current import (
    heat_current_1d,
    heat_current_density_3d,
    heat_current_from_spectrum,
    heat_current_per_length_2d,
    transmission_spectra,
    transmission_spectrum,
    transverse_area_from_vectors,
    transverse_length_from_vector,
)
from .types import Device1D, DeviceKSpace, LeadBlocks, LeadKSpace

__all__ = [
    "HeatCurrentCumulants",
    "HeatCurrentUncertainty",
    "heat_current_cumulants_from_k_moments",
    "heat_current_cumulants_from_spectrum",
    "heat_current_cumulants_scan",
    "heat_current_uncertainty",
    "transmission_k_moments_spectrum",
    "transmission_torch",
    "transmission_sets_torch",
    "transmission_sum_torch",
    "thermal_conductance_torch",
    "Device1D",
    "DeviceKSpace",
    "LeadBlocks",
    "LeadKSpace",
    "lead_surface_dos",
    "lead_surface_dos_kavg",
    "lead_surface_dos_kavg_adaptive",
    "lead_surface_dos_spectrum",
    "leads_surface_dos_spectrum",
    "device_green_function",
    "surface_gf",
    "surface_gf_generalized_eigen",
    "surface_gf_generalized_eigen_svd",
    "surface_gf_legacy_eigen_svd",
    "surface_gf_sancho_rubio",
    "transmission",
    "transmission_kavg",
    "transmission_kavg_adaptive",
    "transmission_kavg_adaptive_global_eta",
    "transmission_kavg_adaptive_spectrum",
    "transmission_spectra",
    "transmission_spectrum",
    "heat_current_from_spectrum",
    "heat_current_1d",
    "heat_current_per_length_2d",
    "heat_current_density_3d",
    "transverse_length_from_vector",
    "transverse_area_from_vectors",
    "qe_omega_to_rad_s",
    "thermal_conductance_from_spectrum",
    "set_default_precision",
    "set_eta_ladders_in_kernel",
    "get_default_precision",
]
