"""negfpy_amd: phonon-NEGF transport (surface Green's functions, self-energies,
recursive Green's function device solve, transmission, surface DOS, Landauer
heat current and thermal conductance) batched over (omega, k_par) points on
AMD Instinct MI355X (gfx950) through hand-written HIP kernels.

``negfpy_amd.core`` and ``negfpy_amd.models`` keep the API of the reference
package's ``negfpy.core`` / ``negfpy.models`` transport path.
"""

__version__ = "0.1.0"
