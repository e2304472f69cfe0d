"""Study drivers on top of the batched transport path (reference ``src/negfpy/workflows``)."""

from .ifc_transmission import (
    build_kpoints,
    build_omega_grid,
    build_transport_model,
    prepare_kmesh,
    run_transmission_config,
    save_spectrum_tsv,
    spectrum_metadata,
    effective_area_m2,
    estimate_omega_max,
    eta_values_for_omega,
    transmission_spectrum_from_config,
)
from .run_bulk import (
    load_transmission_spectrum,
    resolve_transmission_source,
    run_ifc_bulk,
    save_thermal_conductance_tsv,
    thermal_conductance_run,
)

__all__ = ["build_kpoints", "build_omega_grid", "build_transport_model", "prepare_kmesh",
           "run_transmission_config", "save_spectrum_tsv", "spectrum_metadata", "effective_area_m2",
           "estimate_omega_max", "eta_values_for_omega", "transmission_spectrum_from_config",
           "load_transmission_spectrum", "resolve_transmission_source", "run_ifc_bulk",
           "save_thermal_conductance_tsv", "thermal_conductance_run"]
