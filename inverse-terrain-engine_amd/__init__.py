"""MI355X-native batched f32 shallow-water evaluator for the inverse terrain engine.

The compute path is the HIP library built from ``csrc/`` (C ABI: include/itr_hip.h);
this package is its Python host mirror: terrain preparation (``terrain``), solver
parameters and synthetic workloads (``scenario``), and the batch backend (``gpu``).
"""
from .scenario import SolverParams, SplitMix64, mm_h  # noqa: F401
from .terrain import (FineTerrain, GeoTransform, Primitive, PreparedTerrain, TerrainBase,  # noqa: F401
                      TerrainEdit, materialize)


def __getattr__(name):
    # The GPU classes load the HIP library on use, not on import.
    if name in ("GpuSolver", "GpuRun", "BatchSession", "HipBatchBackend", "Declined", "DeviceError", "BatchRun"):
        from . import gpu
        return getattr(gpu, name)
    raise AttributeError(name)
