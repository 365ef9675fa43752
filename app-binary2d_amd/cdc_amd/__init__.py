"""MI355X-native per-block Runge-Kutta update of the circumbinary disk code.

The product is the C-ABI library libcdc_hip.so (include/cdc_solver.h); this
package is its Python binding and the reference-shaped configuration types.
"""
from . import abi
from .config import Configuration, Euler, Isothermal, Mesh, Physics
from .solver import (CdcError, Ensemble, HydroError, Snapshot, Solver, orbital_state_from_time, resample_state,
                     seeded_cavity)
from . import workloads

__all__ = ["abi", "Configuration", "Euler", "Isothermal", "Mesh", "Physics", "CdcError", "Ensemble", "HydroError", "Snapshot",
           "Solver", "orbital_state_from_time", "resample_state", "seeded_cavity", "workloads"]
