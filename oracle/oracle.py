"""ctypes view of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline.
The product (cholesky_gpu_amd) never imports this module.

The functions restate the reference's algorithm; each C function names the
reference file:line it follows (see oracle.c's header).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MPCHOL_ORACLE_LIB selects another build of oracle.c (the sanitizer build
# of `make -C oracle sanitize`)
_LIB = os.environ.get("MPCHOL_ORACLE_LIB") or os.path.join(_HERE, "_build", "liboracle.so")
_REF = os.path.join(_HERE, "_ref", "libref_micro_chol.so")

FP32, BF16, FP16, E5M2, E4M3, MXE4M3, MXE5M2 = range(7)
SWITCH = -1

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _declare(_lib)
    return _lib


def ref_lib():
    """oracle/_ref build of the reference's micro_cholesky, or None if absent."""
    if not os.path.exists(_REF):
        return None
    r = C.CDLL(_REF)
    r.ref_micro_cholesky.restype = C.c_int
    r.ref_micro_cholesky.argtypes = [C.c_void_p, C.c_int, C.c_int]
    return r


P = C.c_void_p
I64 = C.c_int64
U64 = C.c_uint64
D = C.c_double


class FactorOpts(C.Structure):
    _fields_ = [("nb", C.c_int), ("update_prec", C.c_int), ("ladder", C.c_int * 4),
                ("n_ladder", C.c_int), ("eps_prime", C.c_double), ("tile_rows", C.c_int),
                ("rule", C.c_int)]


class FactorInfo(C.Structure):
    _fields_ = [("info", C.c_int), ("min_pivot", C.c_double), ("flops", C.c_double * 8)]


class PcgOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxit", C.c_int), ("stop_rule", C.c_int),
                ("x0_mode", C.c_int)]


def _declare(L):
    sig = {
        "orc_f32_to_e4m3": (C.c_uint8, [C.c_float]),
        "orc_f32_to_e5m2": (C.c_uint8, [C.c_float]),
        "orc_e4m3_to_f32": (C.c_float, [C.c_uint8]),
        "orc_e5m2_to_f32": (C.c_float, [C.c_uint8]),
        "orc_f32_to_bf16": (C.c_uint16, [C.c_float]),
        "orc_f32_to_f16": (C.c_uint16, [C.c_float]),
        "orc_bf16_to_f32": (C.c_float, [C.c_uint16]),
        "orc_f16_to_f32": (C.c_float, [C.c_uint16]),
        "orc_f32_to_e4m3_n": (None, [P, P, I64]),
        "orc_f32_to_e5m2_n": (None, [P, P, I64]),
        "orc_sweep_e4m3": (I64, [U64, U64, P, D, C.POINTER(I64)]),
        "orc_sweep_e5m2": (I64, [U64, U64, P, D, C.POINTER(I64)]),
        "orc_closed_form_e4m3": (C.c_uint8, [C.c_float]),
        "orc_closed_form_e5m2": (C.c_uint8, [C.c_float]),
        "orc_selfcheck_f8": (I64, [C.c_int, U64, U64]),
        "orc_syrk_lower": (None, [P, I64, I64, P, I64, I64]),
        "orc_elem_bytes": (C.c_int, [C.c_int]),
        "orc_unit_roundoff": (D, [C.c_int]),
        "orc_scale_exp": (C.c_int, [C.c_float, C.c_int]),
        "orc_quantize_panel": (C.c_int, [P, I64, I64, I64, C.c_int, P, P, I64, P, P, I64]),
        "orc_philox4x32": (None, [P, P, P]),
        "orc_uniform": (D, [U64, U64, U64]),
        "orc_probes": (None, [I64, C.c_int, U64, P, I64]),
        "orc_normal": (D, [U64, U64, U64]),
        "orc_spectrum": (None, [C.c_int, I64, D, D, I64, D, D, U64, P]),
        "orc_hh_setup": (C.c_int, [I64, C.c_int, U64, P, P, P, P]),
        "orc_hh_entry": (D, [I64, I64, P, P, P, P]),
        "orc_hh_matrix": (None, [I64, P, P, P, P, P, I64]),
        "orc_rbf_matrix": (None, [I64, C.c_int, P, I64, P, C.c_int, D, D, P, I64]),
        "orc_points_gaussian": (None, [I64, C.c_int, U64, P, I64]),
        "orc_points_sparse_image": (None, [I64, C.c_int, D, U64, P, I64]),
        "orc_potrf_f64": (C.c_int, [P, I64, I64]),
        "orc_trsv_lower_f64": (None, [P, I64, I64, P, C.c_int]),
        "orc_symv_f32_f64": (None, [P, I64, I64, P, P]),
        "orc_potrf_f32_micro": (C.c_int, [P, C.c_int, C.c_int]),
        "orc_factor_mixed": (C.c_int, [P, I64, I64, C.POINTER(FactorOpts), P, P, P, C.c_int,
                                       C.POINTER(FactorInfo)]),
        "orc_rule_minpivot": (C.c_int, [P, C.c_int, D, D, D]),
        "orc_rule_minpivot_amax": (C.c_int, [P, C.c_int, D, D, D, D]),
        "orc_pcg": (C.c_int, [P, I64, P, I64, I64, P, P, C.POINTER(PcgOpts), P, P, P]),
        "orc_refine": (C.c_int, [P, I64, P, I64, I64, P, P, C.POINTER(PcgOpts), P, P]),
        "orc_block_pcg": (C.c_int, [P, I64, P, I64, I64, C.c_int, P, P, C.POINTER(PcgOpts), P]),
        "orc_lanczos_cond": (D, [P, P, C.c_int]),
        "orc_num_threads": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


# ---------------------------------------------------------------- formats
def e4m3(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty(x.shape, np.uint8)
    lib().orc_f32_to_e4m3_n(_p(x), _p(y), x.size)
    return y


def e5m2(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty(x.shape, np.uint8)
    lib().orc_f32_to_e5m2_n(_p(x), _p(y), x.size)
    return y


def sweep(fmt, lo, hi, expect, max_abs):
    """(in-domain mismatches, out-of-domain mismatches) of the conversion of the
    float bit patterns [lo, hi) against expect (uint8 array of length hi-lo)."""
    expect = np.ascontiguousarray(expect, np.uint8)
    other = C.c_int64(0)
    f = lib().orc_sweep_e4m3 if fmt == E4M3 else lib().orc_sweep_e5m2
    bad = f(lo, hi, _p(expect), max_abs, C.byref(other))
    return bad, other.value


def selfcheck_f8(fmt, lo=0, hi=1 << 32):
    return lib().orc_selfcheck_f8(1 if fmt == E4M3 else 0, lo, hi)


def syrk_lower(Cm, Q):
    """Cm: m x m float32 (Fortran order, lower updated in place); Q: m x k."""
    Q = np.asfortranarray(Q, np.float32)
    m, k = Q.shape
    lib().orc_syrk_lower(_p(Cm), m, m, _p(Q), m, k)
    return Cm


def decode_e4m3(v):
    L = lib()
    return np.array([L.orc_e4m3_to_f32(int(b)) for b in np.ravel(v)], np.float32).reshape(np.shape(v))


def decode_e5m2(v):
    L = lib()
    return np.array([L.orc_e5m2_to_f32(int(b)) for b in np.ravel(v)], np.float32).reshape(np.shape(v))


def scale_exp(amax, prec):
    return lib().orc_scale_exp(float(amax), prec)


def unit_roundoff(prec):
    return lib().orc_unit_roundoff(prec)


def quantize_panel(P, prec_of_tile, tile_rows):
    """P: m x k float32 (column-major semantics: P[r, c]).  Returns
    (payload[m, row_stride] uint8, scales[m, k/32] uint8, deq[m, k] float32)."""
    P = np.asfortranarray(P, dtype=np.float32)
    m, k = P.shape
    prec_of_tile = np.ascontiguousarray(prec_of_tile, dtype=np.uint8)
    stride = 4 * k
    pay = np.zeros((m, stride), np.uint8)
    sc = np.zeros((m, max(k // 32, 1)), np.uint8)
    deq = np.zeros((m, k), np.float32, order="F")
    rc = lib().orc_quantize_panel(_p(P), m, k, m, tile_rows, _p(prec_of_tile), _p(pay), stride,
                                  _p(sc), _p(deq), m)
    if rc != 0:
        raise FloatingPointError("NaN/Inf in panel")
    return pay, sc, deq


# ---------------------------------------------------------------- rng / generators
def philox(ctr, key):
    c = np.array(ctr, np.uint32)
    k = np.array(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().orc_philox4x32(_p(c), _p(k), _p(o))
    return o


def normal(seed, stream, idx):
    return lib().orc_normal(seed, stream, idx)


def uniform(seed, stream, idx):
    return lib().orc_uniform(seed, stream, idx)


def probes(n, nprobe, seed):
    """(nprobe, n) float64 Rademacher probes (orc_probes)."""
    Z = np.zeros((nprobe, n), np.float64)
    lib().orc_probes(n, nprobe, seed, _p(Z), n)
    return Z


def _gp_sqdiff(Xs, c):
    """(x_ic - x_jc)^2 of the scaled points for dimension c (n, n)."""
    x = Xs[:, c]
    return (x[:, None] - x[None, :]) ** 2


def gp_matrix_f64(X, ls, variance, noise):
    """A = variance * exp(-1/2 sum_c ((x_ic - x_jc)/ls_c)^2) + noise I in FP64
    (X: (n, d) float32), the noise-free kernel K0, and the scaled points."""
    X = np.asarray(X, np.float64)
    ls = np.atleast_1d(np.asarray(ls, np.float64))
    Xs = X / ls if ls.size > 1 else X / ls[0]
    S = np.zeros((X.shape[0], X.shape[0]))
    for c in range(X.shape[1]):
        S += _gp_sqdiff(Xs, c)
    K0 = variance * np.exp(-0.5 * S)
    return K0 + noise * np.eye(X.shape[0]), K0, Xs


def gp_loss_grad_f64(X, ls, variance, noise, b, Z=None, unit=False):
    """FP64 reference of the GP loss b^T A^-1 b / 2 + logdet(A) / 2 and its
    gradient.  With Z (nprobe, n) the logdet gradient is the Hutchinson estimate
    over those probes (weight 1/P, or 1 with unit=True); with Z None it is exact
    (tr(A^-1 dA)).  Returns loss, grad_b, grad_ls, grad_variance, grad_noise.
    dA_ij/dls_c = K_ij (x_ic - x_jc)^2 / ls_c^3, dA/dvariance = K / variance,
    dA/dnoise = I.  (No reference counterpart: the GP workload is defined by
    the differentiable-solve configuration only.)"""
    import scipy.linalg as sl
    ls = np.atleast_1d(np.asarray(ls, np.float64))
    A, K0, Xs = gp_matrix_f64(X, ls, variance, noise)
    cf = sl.cho_factor(A, lower=True)
    del A
    b = np.asarray(b, np.float64)
    alpha = sl.cho_solve(cf, b)
    logdet = 2.0 * np.log(np.diag(cf[0])).sum()
    loss = 0.5 * b @ alpha + 0.5 * logdet
    if Z is None:
        T = sl.cho_solve(cf, np.eye(K0.shape[0]))           # exact: tr(A^-1 dA)
    else:
        W = sl.cho_solve(cf, Z.T)                           # A^-1 z_p as columns
        T = W @ Z / (1.0 if unit else Z.shape[0])
    del cf
    M = 0.5 * T
    del T
    M -= 0.5 * np.outer(alpha, alpha)
    gnoise = np.trace(M)
    MK = M * K0                                             # weights of (dx/ls)^2 per dim
    del M, K0
    d = Xs.shape[1]
    if ls.size > 1:
        gls = np.array([(MK * _gp_sqdiff(Xs, c)).sum() / ls[c] for c in range(d)])
    else:
        gls = np.array([sum((MK * _gp_sqdiff(Xs, c)).sum() for c in range(d)) / ls[0]])
    return loss, alpha, gls, MK.sum() / variance, gnoise


def spectrum(dist, n, cond, lam_max=1.0, n_cluster=0, cluster_lo=0.0, cluster_hi=0.0, seed=0):
    lam = np.zeros(n, np.float64)
    lib().orc_spectrum(dist, n, cond, lam_max, n_cluster, cluster_lo, cluster_hi, seed, _p(lam))
    return lam


def hh_setup(n, nrefl, seed, lam):
    lam = np.ascontiguousarray(lam, np.float64)
    W = np.zeros((n, 4))
    Pm = np.zeros((n, 4))
    g = np.zeros((n, 4))
    rc = lib().orc_hh_setup(n, nrefl, seed, _p(lam), _p(W), _p(Pm), _p(g))
    assert rc == 0
    return W, Pm, g


def hh_matrix(n, lam, W, Pm, g):
    A = np.zeros((n, n), np.float32, order="F")
    lib().orc_hh_matrix(n, _p(np.ascontiguousarray(lam)), _p(W), _p(Pm), _p(g), _p(A), n)
    return A


def householder_spd(n, cond, lam_max, seed, nrefl=4, dist=4, n_cluster=0, cluster_lo=0.0,
                    cluster_hi=0.0):
    lam = spectrum(dist, n, cond, lam_max, n_cluster, cluster_lo, cluster_hi, seed)
    W, Pm, g = hh_setup(n, nrefl, seed, lam)
    return hh_matrix(n, lam, W, Pm, g), lam


def points_gaussian(n, d, seed):
    X = np.zeros((n, d), np.float32)
    lib().orc_points_gaussian(n, d, seed, _p(X), d)
    return X


def points_sparse_image(n, d, p_zero, seed):
    X = np.zeros((n, d), np.float32)
    lib().orc_points_sparse_image(n, d, p_zero, seed, _p(X), d)
    return X


def rbf_matrix(X, ls, variance=1.0, nugget=0.0):
    X = np.ascontiguousarray(X, np.float32)
    n, d = X.shape
    ls = np.atleast_1d(np.asarray(ls, np.float64))
    K = np.zeros((n, n), np.float32, order="F")
    lib().orc_rbf_matrix(n, d, _p(X), d, _p(ls), ls.size, variance, nugget, _p(K), n)
    return K


# ---------------------------------------------------------------- FP64 oracle
def potrf_f64(A):
    L = np.array(A, dtype=np.float64, order="F", copy=True)
    info = lib().orc_potrf_f64(_p(L), L.shape[0], L.shape[0])
    return np.tril(L), info


def potrf_f32_micro(A):
    L = np.array(A, dtype=np.float32, order="F", copy=True)
    info = lib().orc_potrf_f32_micro(_p(L), L.shape[0], L.shape[0])
    return L, info


def symv(A, x):
    A = np.asfortranarray(A, np.float32)
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros_like(x)
    lib().orc_symv_f32_f64(_p(A), A.shape[0], A.shape[0], _p(x), _p(y))
    return y


# ---------------------------------------------------------------- reference path
RULE_MINPIVOT, RULE_FROBENIUS, RULE_SCHUR = 0, 1, 2


def factor_mixed(A, nb=256, update_prec=FP32, ladder=(E4M3, FP32), eps_prime=1.0,
                 map_in=None, max_steps=-1, tile_rows=None, rule=RULE_MINPIVOT):
    """Returns (L [full array, lower valid], info, map, updated_diag, FactorInfo)."""
    L = np.array(A, dtype=np.float32, order="F", copy=True)
    n = L.shape[0]
    o = FactorOpts()
    o.nb = nb
    o.update_prec = update_prec
    for i, p in enumerate(ladder):
        o.ladder[i] = p
    o.n_ladder = len(ladder)
    o.eps_prime = eps_prime
    o.tile_rows = tile_rows or nb
    o.rule = rule
    nt = (n + nb - 1) // nb
    mp = np.full((nt, nt), 255, np.uint8, order="F")
    ud = np.zeros(n, np.float32)
    fi = FactorInfo()
    mi = None
    if map_in is not None:
        mi = np.asfortranarray(map_in, np.uint8)
    rc = lib().orc_factor_mixed(_p(L), n, n, C.byref(o), _p(mi), _p(mp), _p(ud), max_steps,
                                C.byref(fi))
    return L, rc, mp, ud, fi


def pcg(A, L, b, tol=1e-10, maxit=500, stop_rule=0, x0_mode=0):
    A = np.asfortranarray(A, np.float32)
    L = np.asfortranarray(L, np.float32)
    b = np.ascontiguousarray(b, np.float64)
    n = b.size
    x = np.zeros(n)
    o = PcgOpts(tol, maxit, stop_rule, x0_mode)
    rr = C.c_double(0)
    al = np.zeros(maxit + 1)
    be = np.zeros(maxit + 1)
    it = lib().orc_pcg(_p(A), n, _p(L), n, n, _p(b), _p(x), C.byref(o), C.byref(rr), _p(al), _p(be))
    k = it if it > 0 else (maxit if it < 0 else 0)
    kappa = lib().orc_lanczos_cond(_p(al), _p(be), k) if k > 0 else 1.0
    return x, it, rr.value, kappa


def refine(A, L, b, tol=1e-10, maxit=50, x0_mode=0):
    """orc_refine: x += M^-1 (b - A x) until the FP64 true residual meets
    tol.  Returns (x, corrections or -1, final relres, residual history)."""
    A = np.asfortranarray(A, np.float32)
    L = np.asfortranarray(L, np.float32)
    b = np.ascontiguousarray(b, np.float64)
    n = b.size
    x = np.zeros(n)
    o = PcgOpts(tol, maxit, 0, x0_mode)
    rr = C.c_double(0)
    hist = np.full(maxit + 1, np.nan)
    it = lib().orc_refine(_p(A), n, _p(L), n, n, _p(b), _p(x), C.byref(o), C.byref(rr), _p(hist))
    k = it if it >= 0 else maxit
    return x, it, rr.value, hist[:k + 1].copy()


def block_pcg(A, L, B, tol=1e-10, maxit=500, x0_mode=0):
    """orc_block_pcg over the columns of B (n x s) together.  Returns
    (X, iterations or -1 / -2, per-column final true relres)."""
    A = np.asfortranarray(A, np.float32)
    L = np.asfortranarray(L, np.float32)
    B = np.asfortranarray(B, np.float64)
    n, s = B.shape
    X = np.zeros((n, s), order="F")
    o = PcgOpts(tol, maxit, 0, x0_mode)
    rel = np.zeros(s)
    it = lib().orc_block_pcg(_p(A), n, _p(L), n, n, s, _p(B), _p(X), C.byref(o), _p(rel))
    return X, it, rel
