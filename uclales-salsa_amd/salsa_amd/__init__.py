"""MI355X implementation of UCLALES-SALSA's coagulation + condensation step.

salsa_amd.capi     -- ctypes binding of the C ABI (include/salsa_amd.h)
salsa_amd.workload -- seeded synthetic LES fields
salsa_amd.opcount  -- algorithmic work per cell-step (flops, HBM bytes)
"""
from . import capi, workload  # noqa: F401

__all__ = ["capi", "workload"]
