"""MI355X-native ACO solver for the capacitated vehicle routing problem.

The hot path (candidate lists, tour construction, quantized pheromone
update, route costs) runs in hand-written gfx950 HIP kernels behind the C ABI
of include/aco_hip.h; this package is the Python view of that ABI.
"""
from .solver import AcoError, HipSolver, Instance, Params, solve, split_routes  # noqa: F401
from . import _native  # noqa: F401

__all__ = ["AcoError", "HipSolver", "Instance", "Params", "solve", "split_routes"]
