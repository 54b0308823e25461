"""MI355X-native element-by-element linear-elasticity solve path.

Drop-in for the solver path of moarseniy/CudaFEMproject (femfunc: static
Jacobi-PCG, GN22 explicit/implicit dynamics) built on hand-written HIP
kernels for gfx950 behind the C-ABI of include/femgpu.h.
"""
from ._lib import (BATCH_LOCKSTEP, BATCH_PERSISTENT, CG_SINGLE_REDUCTION, CG_STANDARD, FP32, FP64, MIXED, IMPLICIT_PHYSICAL, IMPLICIT_REFERENCE, LOADS_PHYSICAL,
                   LOADS_REFERENCE, MASS_PHYSICAL, MASS_REFERENCE, TOL_REF_GAMMA, TOL_REL_RES,
                   RECV_ACCELERATION, RECV_DISPLACEMENT, RECV_VELOCITY, WAVELET_BERLAGE,
                   WAVELET_NONE, WAVELET_RICKER, Context, FemGpuError,
                   load_library, plan_info)
from .problem import Problem

__all__ = ["Context", "FemGpuError", "Problem", "load_library", "plan_info", "FP32", "FP64", "MIXED",
           "TOL_REF_GAMMA", "TOL_REL_RES", "LOADS_REFERENCE", "LOADS_PHYSICAL",
           "MASS_REFERENCE", "MASS_PHYSICAL", "WAVELET_NONE", "WAVELET_RICKER",
           "WAVELET_BERLAGE", "IMPLICIT_REFERENCE", "IMPLICIT_PHYSICAL", "CG_STANDARD",
           "CG_SINGLE_REDUCTION", "BATCH_LOCKSTEP", "BATCH_PERSISTENT", "RECV_DISPLACEMENT", "RECV_VELOCITY", "RECV_ACCELERATION"]
