"""MI355X (gfx950) implementation of the ProteinSim2015 hot path: cell-list
neighbour search, the IForce pair forces (LJ/Yukawa, AO depletion,
Calibration soft shell) and the van Gunsteren-Berendsen integrator step,
as HIP kernels behind a C ABI (include/psim_gpu.h).

The native library must be built (``python -m proteinsim2015_amd.build``);
importing the engine loads it and fails loudly when it is missing.
"""
from ._lib import (ERR_BOUNDS, ERR_INIT, ERR_NAN, ERR_OVERLAP, FORCE_AO, FORCE_CAL,  # noqa: F401
                   FORCE_LJ, PsimError, load)
from .engine import (Adjoint, AOPotential, BrownianIntegrator, Calibration, Config, Ensemble,  # noqa: F401
                     ForceSet, GpuSystem, LennardJones, PLUGINS, Replica, SystemGeometry,
                     rsa_positions)

__all__ = [
    "Adjoint", "AOPotential", "BrownianIntegrator", "Calibration", "Config", "Ensemble", "ForceSet",
    "GpuSystem", "Replica",
    "LennardJones", "PLUGINS", "SystemGeometry", "rsa_positions", "PsimError", "load",
]
