"""disco_amd -- MI355X (gfx950) backend for DISCO's displacement hot path.

The per-atom sum of the isotropic elastic displacement of dislocation loops
over polygon segments (closed-form line integral) or loopons (Kroupa sum) and
periodic images, computed by hand-written HIP kernels in ``libdisco_hip.so``
behind a C ABI (``include/disco_hip.h``).  ``backend`` mirrors the reference's
``Code/backend.hpp`` interface; ``geometry`` produces the synthetic workloads.
"""
from .backend import (BackendTiming, Context, DiscoError, GpuOptions, LoopData, Plan,
                      PREC_FP32C, PREC_FP64, gpu_available, gpu_info, gpu_line_fp64,
                      kroupa_gpu, line_sum, loopon_sum, print_gpu_info)

__all__ = ["BackendTiming", "Context", "DiscoError", "GpuOptions", "LoopData", "Plan",
           "PREC_FP32C", "PREC_FP64", "gpu_available", "gpu_info", "gpu_line_fp64",
           "kroupa_gpu", "line_sum", "loopon_sum", "print_gpu_info"]
