"""MI355X (gfx950) implementation of the euler2d hot path.

The compute lives in lib/libeuler2d_hip.so (C-ABI: include/euler2d_hip.h);
this package is its Python binding and mirrors the reference's
HydroParameters / HydroRun interface.
"""
from .hydro import (  # noqa: F401
    FN_CMPFLX, FN_CONV_GT, FN_DIV, FN_DIV_NEAR, FN_DIV_ROOT, FN_EOS, FN_HLLC, FN_PRIM, FN_RIEMANN, FN_SLOPE, FN_SQRT, FN_TRACE, IMPL_FLUX_STORE, IMPL_FUSED,
    IMPL_MATERIALISED, LIB_PATH, PREC_FP32, PREC_FP32_NATIVE, PREC_FP64, PREC_MIXED, PROBLEM_BUMPS, PROBLEM_DIAG,
    PRECISIONS, PROBLEM_IMPLODE, PROBLEMS, RIEMANN_APPROX, RIEMANN_HLLC, E2DError, HydroParams, HydroRun,
    LibraryNotBuilt, device_available, exported_symbols, load_library, pointwise, read_namelist, run_steps_noalloc,
)
