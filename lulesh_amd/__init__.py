"""MI355X (gfx950) implementation of the LULESH 2.x Lagrange leapfrog step.

The device path is hand-written HIP behind a C ABI (include/lulesh_gpu.h);
the host keeps the reference's Domain class and per-cycle call structure
(include/lulesh_host.h).  This package only loads those libraries.
"""
from .native import (FIELDS, FP64, FP64_FMA, MIXED, DeviceContext, Ensemble, HostDomain,  # noqa: F401
                     SedovDevice, TimeState, check, default_constants, ensemble_draw, gpu_lib,
                     host_lib, region_runs)

__all__ = ["FIELDS", "FP64", "FP64_FMA", "MIXED", "DeviceContext", "Ensemble", "HostDomain",
           "SedovDevice", "TimeState", "check", "default_constants", "ensemble_draw", "gpu_lib",
           "host_lib", "region_runs"]
