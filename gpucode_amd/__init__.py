"""gpucode_amd -- MI355X-native DEM time-step engine (hand-written gfx950 HIP).

The compute path is libgpucode_amd.so (csrc/): HIP kernels behind the C ABI
declared in include/dem_capi.h and include/dem_batch.h.  This package is the Python host side:
ctypes bindings (_capi), the Engine wrapper (engine), seeded workload
generators (workloads) and parameter calibration (calibrate).
"""
from ._capi import (DemError, DemInput, DemSetup, DemStats, PREC_FP64, PREC_MIXED,  # noqa: F401
                    declared_symbols, lib)
from .engine import Batch, Engine, autotune, derive, load_para, particle_feed, read_deck, tune_instances, write_tec  # noqa: F401

__all__ = ["Engine", "Batch", "autotune", "DemInput", "DemSetup", "DemStats", "DemError", "PREC_FP64", "PREC_MIXED",
           "derive", "read_deck", "particle_feed", "write_tec", "load_para", "lib",
           "declared_symbols", "tune_instances"]
