"""MI355X-native 2D SPH step (gfx950 HIP kernels behind a C ABI).

Import with importlib, the directory name carries a hyphen:
    sph = importlib.import_module("sph-simulations_amd")
"""
from .capi import (BC_PERIODIC, BC_WALLS, COMPUTE_LISTS, COMPUTE_SORTED, LIB_PATH,
                   ORDER_MORTON, ORDER_ROWMAJOR, PHYS_REFERENCE, PHYS_WCSPH, STAGES,
                   STEP_FUSED_BLOCK, STEP_PARTICLE, STEP_SPLIT_BLOCK, RESORT_FULL,
                   RESORT_INCREMENTAL, Batch, BatchTuning, Context, GridInfo, Params,
                   SphError, lib)
from . import diagnostics, scenes
from .capi import snapshot_info

__all__ = ["diagnostics", "snapshot_info", "Batch", "BatchTuning", "Context", "Params", "GridInfo", "SphError", "lib", "scenes", "LIB_PATH",
           "ORDER_MORTON", "ORDER_ROWMAJOR", "PHYS_REFERENCE", "PHYS_WCSPH", "BC_WALLS",
           "BC_PERIODIC", "COMPUTE_SORTED", "COMPUTE_LISTS", "STAGES", "STEP_SPLIT_BLOCK",
           "STEP_FUSED_BLOCK", "STEP_PARTICLE", "RESORT_FULL", "RESORT_INCREMENTAL"]
