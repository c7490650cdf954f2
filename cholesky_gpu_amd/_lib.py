"""ctypes binding of libmpchol.so (include/mpchol.h).

The library is the product: every compute call below goes to hand-written
gfx950 kernels through the C ABI.  There is no CPU fallback: if the shared
library is missing, importing the compute entry points raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libmpchol.so")

# ---- constants (include/mpchol.h)
ERR_ARG = -1
ERR_HIP = -100
ERR_NOT_CONVERGED = -101
ERR_NAN = -102
ERR_NOMEM = -103
ERR_UNSUPPORTED = -104
ERR_NO_DEVICE = -105

PREC_FP32, PREC_BF16, PREC_FP16, PREC_E5M2, PREC_E4M3, PREC_MXE4M3, PREC_MXE5M2 = range(7)
PREC_SWITCH = -1
RULE_MINPIVOT, RULE_FROBENIUS, RULE_SCHUR = 0, 1, 2
PREC_NAMES = {0: "fp32", 1: "bf16", 2: "fp16", 3: "e5m2", 4: "e4m3", 5: "mxe4m3", 6: "mxe5m2"}

NOALLOC_MAX_TILES = 1024        # MPC_NOALLOC_MAX_TILES of include/mpchol.h
BLOCK_PCG_MAX = 16              # MPC_BLOCK_PCG_MAX
DIST_RANDOM_LOG_UNIFORM, DIST_CLUSTERED, DIST_ARITHMETIC, DIST_GEOMETRIC, DIST_LOGSPACED = range(5)
MAT_HOUSEHOLDER_SPD = 0


class MatgenOpts(C.Structure):
    _fields_ = [("kind", C.c_int), ("dist", C.c_int), ("cond", C.c_double),
                ("lam_max", C.c_double), ("n_reflectors", C.c_int), ("seed", C.c_uint64),
                ("n_cluster", C.c_int64), ("cluster_lo", C.c_double), ("cluster_hi", C.c_double)]


class FactorOpts(C.Structure):
    _fields_ = [("nb", C.c_int), ("update_prec", C.c_int), ("ladder", C.c_int * 4),
                ("n_ladder", C.c_int), ("eps_prime", C.c_double), ("fallback", C.c_int),
                ("profile", C.c_int), ("agg_panels", C.c_int), ("rule", C.c_int)]


class FactorInfo(C.Structure):
    _fields_ = [("info", C.c_int), ("retries", C.c_int), ("eps_prime_used", C.c_double),
                ("min_pivot", C.c_double), ("flops_total", C.c_double),
                ("syrk_flops", C.c_double * 8), ("tiles", C.c_int64 * 8),
                ("ms_total", C.c_double), ("ms_potrf", C.c_double), ("ms_trsm", C.c_double),
                ("ms_quant", C.c_double), ("ms_syrk", C.c_double)]


class PcgOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxit", C.c_int), ("stop_rule", C.c_int),
                ("x0_mode", C.c_int), ("graph", C.c_int), ("check_every", C.c_int)]


class PcgInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("converged", C.c_int), ("relres", C.c_double),
                ("kappa_est", C.c_double), ("ms_total", C.c_double)]


class GpOpts(C.Structure):
    _fields_ = [("n_probes", C.c_int), ("probe_kind", C.c_int), ("probe_seed", C.c_uint64),
                ("factor", FactorOpts), ("pcg", PcgOpts)]


class GpInfo(C.Structure):
    _fields_ = [("loss", C.c_double), ("datafit", C.c_double), ("logdet", C.c_double),
                ("factor", FactorInfo), ("pcg", PcgInfo)]


class Settings(C.Structure):
    _fields_ = [("n", C.c_int64), ("condition_number", C.c_double), ("dist_type", C.c_int),
                ("block_size", C.c_int), ("eps_prime", C.c_double), ("floor", C.c_double),
                ("diag_pert", C.c_int), ("left", C.c_int)]


P = C.c_void_p
I64 = C.c_int64
D = C.c_double
I = C.c_int

SIGNATURES = {
    "mpc_handle_create": (I, [C.POINTER(P), I]),
    "mpc_handle_destroy": (I, [P]),
    "mpc_set_stream": (I, [P, P]),
    "mpc_get_stream": (P, [P]),
    "mpc_synchronize": (I, [P]),
    "mpc_last_error_string": (C.c_char_p, []),
    "mpc_version_string": (C.c_char_p, []),
    "mpc_matgen_opts_default": (None, [C.POINTER(MatgenOpts)]),
    "mpc_matgen": (I, [P, P, I64, I64, C.POINTER(MatgenOpts), P]),
    "mpc_rbf_points": (I, [P, P, I64, I, I64, I, D, C.c_uint64]),
    "mpc_rbf_assemble": (I, [P, P, I64, I, I64, C.POINTER(D), I, D, D, P, I64]),
    "mpc_median_pairwise_distance": (I, [P, P, I64, I, I64, I64, C.c_uint64, C.POINTER(D)]),
    "mpc_factor_opts_default": (None, [C.POINTER(FactorOpts)]),
    "mpc_factor": (I, [P, P, I64, I64, C.POINTER(FactorOpts), P, P, C.POINTER(FactorInfo)]),
    "mpc_pcg_opts_default": (None, [C.POINTER(PcgOpts)]),
    "mpc_pcg": (I, [P, P, I64, P, I64, I64, I, I, P, I64, P, I64, C.POINTER(PcgOpts),
                    C.POINTER(PcgInfo)]),
    "mpc_refine": (I, [P, P, I64, P, I64, I64, I, I, P, I64, P, I64, C.POINTER(PcgOpts),
                       C.POINTER(PcgInfo)]),
    "mpc_block_pcg": (I, [P, P, I64, P, I64, I64, I, I, P, I64, P, I64, C.POINTER(PcgOpts),
                          C.POINTER(PcgInfo)]),
    "mpc_precond_cg": (I, [P, P, I64, I64, I, P, I64, P, I64, C.POINTER(FactorOpts),
                           C.POINTER(PcgOpts), C.POINTER(FactorInfo), C.POINTER(PcgInfo)]),
    "mpc_pcg64": (I, [P, P, I64, P, I64, I64, I, I, P, I64, P, I64, C.POINTER(PcgOpts),
                      C.POINTER(PcgInfo)]),
    "mpc_rbf_assemble64": (I, [P, P, I64, I, I64, C.POINTER(D), I, D, D, P, I64, P, I64]),
    "mpc_probes": (I, [P, P, I64, I64, I, C.c_uint64, I]),
    "mpc_gp_opts_default": (None, [C.POINTER(GpOpts)]),
    "mpc_gp_loss_grad": (I, [P, P, I64, I, I64, C.POINTER(D), I, D, D, P, C.POINTER(GpOpts), P,
                             C.POINTER(D), C.POINTER(D), C.POINTER(D), C.POINTER(GpInfo)]),
    "mpc_quantize_panel": (I, [P, P, I64, I64, I64, I, P, P, I64, P]),
    "mpc_quantize_panel_noalloc": (I, [P, P, I64, I64, I64, I, P, P, I64, P]),
    "mpc_syrk": (I, [P, P, I64, I64, I64, P, I64, P, I, P]),
    "mpc_syrk_masked": (I, [P, P, I64, I64, I64, P, I64, P, I, P, C.c_uint, C.c_uint]),
    "mpc_apply_precond": (I, [P, P, I64, I64, I, P, I64, P, I64]),
    "mpc_potrf_block": (I, [P, P, I, I64, C.POINTER(I)]),
    "mpc_convert_f8": (I, [P, P, P, I64, I]),
    "mpc_pcg_last_coeffs": (I, [P, C.POINTER(D), C.POINTER(D), I, C.POINTER(I)]),
    "mpc_settings_default": (None, [C.POINTER(Settings)]),
    "mpc_settings_parse": (I, [C.c_char_p, C.c_size_t, C.POINTER(Settings)]),
    "mpc_settings_read": (I, [C.c_char_p, C.POINTER(Settings)]),
    "mpc_settings_apply": (I, [C.POINTER(Settings), C.POINTER(MatgenOpts), C.POINTER(FactorOpts)]),
    "mpc_idx_header": (I, [C.c_char_p, C.POINTER(I), C.POINTER(I), C.POINTER(I64), C.POINTER(I64)]),
    "mpc_idx_read_u8": (I, [C.c_char_p, I64, P, I64, C.POINTER(I64)]),
    "mpc_idx_points": (I, [P, C.c_char_p, I64, P, I64, C.POINTER(I64), C.POINTER(I)]),
}

_lib = None


def load():
    """Load libmpchol.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m cholesky_gpu_amd.build` "
                "(there is no CPU fallback)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


class MpcError(RuntimeError):
    def __init__(self, code, where):
        msg = load().mpc_last_error_string().decode(errors="replace")
        super().__init__(f"{where} failed with code {code}: {msg}")
        self.code = code


def check(code, where, allow=()):
    if code != 0 and code not in allow:
        raise MpcError(code, where)
    return code
