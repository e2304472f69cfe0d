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

__all__ = ["build_kpoints", "build_omega_grid", "build_transport_model", "prepare_kmesh",
           "run_transmission_config", "save_spectrum_tsv", "spectrum_metadata", "effective_area_m2",
           "estimate_omega_max", "eta_values_for_omega", "transmission_spectrum_from_config"]
