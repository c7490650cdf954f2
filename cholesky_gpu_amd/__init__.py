"""cholesky_gpu_amd -- mixed-precision Cholesky-preconditioned CG on MI355X.

The compute path is libmpchol.so (hand-written gfx950 HIP kernels behind the C
ABI in include/mpchol.h); this package is its host-side binding.  Build the
library with `python -m cholesky_gpu_amd.build`.
"""
from . import _lib
from ._lib import (PREC_BF16, PREC_E4M3, PREC_E5M2, PREC_FP16, PREC_FP32, PREC_MXE4M3,
                   PREC_MXE5M2, PREC_NAMES, PREC_SWITCH, DIST_ARITHMETIC, DIST_CLUSTERED,
                   DIST_GEOMETRIC, DIST_LOGSPACED, DIST_RANDOM_LOG_UNIFORM, MpcError)

__all__ = ["_lib", "PREC_FP32", "PREC_BF16", "PREC_FP16", "PREC_E5M2", "PREC_E4M3",
           "PREC_MXE4M3", "PREC_MXE5M2", "PREC_SWITCH", "PREC_NAMES", "DIST_LOGSPACED",
           "DIST_RANDOM_LOG_UNIFORM", "DIST_CLUSTERED", "DIST_ARITHMETIC", "DIST_GEOMETRIC",
           "MpcError", "library_path"]


def library_path():
    return _lib.LIB_PATH
