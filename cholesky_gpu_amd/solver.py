"""Host-side API over the C ABI, with torch tensors as device-memory plumbing.

Layout convention: a matrix argument is a CUDA (ROCm) torch tensor whose
memory holds the column-major matrix, i.e. for an n x n matrix M the tensor
T has shape (n, n) and T[j, i] == M[i, j] (T is M's transpose in torch's
row-major view).  Symmetric matrices are unaffected.  A block of nrhs
right-hand sides (column-major n x nrhs) is a contiguous tensor of shape
(nrhs, n).  Every call synchronises torch's device before handing pointers to
the library and the library's stream before returning, so results are ready
for torch ops.
"""
import ctypes as C

import torch

from . import _lib as L
from ._lib import check


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _dev(t):
    if not (t.is_cuda and t.is_contiguous()):
        raise ValueError("expected a contiguous device tensor")
    return t


class Handle:
    """Owns an mpc_handle (its stream and workspaces) on one device."""

    def __init__(self, device=0):
        self.lib = L.load()
        self.device = device
        h = C.c_void_p()
        torch.cuda.synchronize(device)
        check(self.lib.mpc_handle_create(C.byref(h), device), "mpc_handle_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.mpc_handle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(self.lib.mpc_synchronize(self.h), "mpc_synchronize")

    def stream_ptr(self):
        return self.lib.mpc_get_stream(self.h)


def _pre(h):
    torch.cuda.synchronize(h.device)


# ---------------------------------------------------------------- options
def factor_opts(nb=256, update_prec=L.PREC_SWITCH, ladder=(L.PREC_E4M3, L.PREC_FP32),
                eps_prime=0.25, fallback=1, profile=0, agg_panels=None, rule=L.RULE_MINPIVOT):
    o = L.FactorOpts()
    L.load().mpc_factor_opts_default(C.byref(o))
    o.rule = rule
    o.nb = nb
    o.update_prec = update_prec
    for i in range(4):
        o.ladder[i] = ladder[i] if i < len(ladder) else 0
    o.n_ladder = len(ladder)
    o.eps_prime = eps_prime
    o.fallback = fallback
    o.profile = profile
    if agg_panels is not None:
        o.agg_panels = agg_panels
    return o


def pcg_opts(tol=1e-10, maxit=500, stop_rule=0, x0_mode=0, graph=1, check_every=0):
    o = L.PcgOpts()
    L.load().mpc_pcg_opts_default(C.byref(o))
    o.tol, o.maxit, o.stop_rule, o.x0_mode = tol, maxit, stop_rule, x0_mode
    o.graph, o.check_every = graph, check_every
    return o


def info_dict(fi):
    if isinstance(fi, L.FactorInfo):
        return {"info": fi.info, "retries": fi.retries, "eps_prime_used": fi.eps_prime_used,
                "min_pivot": fi.min_pivot, "flops_total": fi.flops_total,
                "syrk_flops": {L.PREC_NAMES[k]: fi.syrk_flops[k] for k in range(7) if fi.syrk_flops[k]},
                "tiles": {L.PREC_NAMES[k]: fi.tiles[k] for k in range(7) if fi.tiles[k]},
                "ms_total": fi.ms_total, "ms_potrf": fi.ms_potrf, "ms_trsm": fi.ms_trsm,
                "ms_quant": fi.ms_quant, "ms_syrk": fi.ms_syrk}
    return {"iters": fi.iters, "converged": fi.converged, "relres": fi.relres,
            "kappa_est": fi.kappa_est, "ms_total": fi.ms_total}


# ---------------------------------------------------------------- generators
def matgen_householder(h, n, cond=1e6, lam_max=1e6, seed=20240531, dist=L.DIST_LOGSPACED,
                       n_reflectors=4, n_cluster=0, cluster_lo=0.0, cluster_hi=0.0, out=None):
    """A = H Lambda H^T (FP32, n x n) and its exact spectrum (FP64)."""
    o = L.MatgenOpts()
    h.lib.mpc_matgen_opts_default(C.byref(o))
    o.dist, o.cond, o.lam_max, o.n_reflectors, o.seed = dist, cond, lam_max, n_reflectors, seed
    o.n_cluster, o.cluster_lo, o.cluster_hi = n_cluster, cluster_lo, cluster_hi
    A = out if out is not None else torch.empty(n, n, dtype=torch.float32, device=h.device)
    lam = torch.empty(n, dtype=torch.float64, device=h.device)
    _pre(h)
    check(h.lib.mpc_matgen(h.h, _ptr(_dev(A)), n, n, C.byref(o), _ptr(lam)), "mpc_matgen")
    return A, lam


def rbf_points(h, n, d, kind=0, p_zero=0.8, seed=1):
    X = torch.empty(n, d, dtype=torch.float32, device=h.device)
    _pre(h)
    check(h.lib.mpc_rbf_points(h.h, _ptr(X), n, d, d, kind, p_zero, seed), "mpc_rbf_points")
    h.sync()
    return X


def rbf_assemble(h, X, ls, variance=1.0, nugget=0.0, out=None):
    n, d = X.shape
    ls = [float(v) for v in (ls if hasattr(ls, "__len__") else [ls])]
    arr = (C.c_double * len(ls))(*ls)
    K = out if out is not None else torch.empty(n, n, dtype=torch.float32, device=h.device)
    _pre(h)
    check(h.lib.mpc_rbf_assemble(h.h, _ptr(_dev(X)), n, d, d, arr, len(ls), variance, nugget,
                                 _ptr(_dev(K)), n), "mpc_rbf_assemble")
    return K


def median_pairwise_distance(h, X, nsub=1024, seed=3):
    n, d = X.shape
    out = C.c_double(0)
    _pre(h)
    check(h.lib.mpc_median_pairwise_distance(h.h, _ptr(_dev(X)), n, d, d, min(nsub, n), seed,
                                             C.byref(out)), "mpc_median_pairwise_distance")
    return out.value


# ---------------------------------------------------------------- factor / solve
def factor(h, A, opts=None, want_map=True, want_diag=False):
    """In-place mixed-precision Cholesky of A (lower).  Returns
    (info, FactorInfo, prec_map or None, updated_diag or None).  prec_map is
    an (nt, nt) uint8 tensor with prec_map[j, i] = precision of tile L_ij."""
    n = A.shape[0]
    opts = opts or factor_opts()
    nt = (n + opts.nb - 1) // opts.nb
    mp = torch.empty(nt, nt, dtype=torch.uint8, device=h.device) if want_map else None
    dg = torch.empty(n, dtype=torch.float32, device=h.device) if want_diag else None
    fi = L.FactorInfo()
    _pre(h)
    rc = h.lib.mpc_factor(h.h, _ptr(_dev(A)), n, n, C.byref(opts), _ptr(mp), _ptr(dg), C.byref(fi))
    if rc < 0:
        check(rc, "mpc_factor")
    return rc, fi, mp, dg


def pcg(h, A, Lf, B, nb=256, opts=None, X=None):
    """Solve A X = B with the factor Lf as preconditioner.  B: (nrhs, n) float64."""
    n = A.shape[0]
    B = B.reshape(-1, n).contiguous()
    nrhs = B.shape[0]
    X = X if X is not None else torch.zeros_like(B)
    pi = L.PcgInfo()
    opts = opts or pcg_opts()
    _pre(h)
    rc = h.lib.mpc_pcg(h.h, _ptr(_dev(A)), n, _ptr(_dev(Lf)), n, n, nb, nrhs, _ptr(_dev(B)), n,
                       _ptr(_dev(X)), n, C.byref(opts), C.byref(pi))
    check(rc, "mpc_pcg", allow=(L.ERR_NOT_CONVERGED,))
    return X, pi, rc


def refine(h, A, Lf, B, nb=256, opts=None, X=None):
    """Iterative refinement X += M^-1 (B - A X) with the factor Lf (mpc_refine).
    B: (nrhs, n) float64.  Returns (X, info, rc); info.iters = corrections."""
    n = A.shape[0]
    B = B.reshape(-1, n).contiguous()
    nrhs = B.shape[0]
    X = X if X is not None else torch.zeros_like(B)
    pi = L.PcgInfo()
    opts = opts or pcg_opts(maxit=50)
    _pre(h)
    rc = h.lib.mpc_refine(h.h, _ptr(_dev(A)), n, _ptr(_dev(Lf)), n, n, nb, nrhs, _ptr(_dev(B)), n,
                          _ptr(_dev(X)), n, C.byref(opts), C.byref(pi))
    check(rc, "mpc_refine", allow=(L.ERR_NOT_CONVERGED,))
    return X, pi, rc


def block_pcg(h, A, Lf, B, nb=256, opts=None, X=None):
    """Block PCG over the rows of B (nrhs <= 16, (nrhs, n) float64) together
    (mpc_block_pcg).  Returns (X, info, rc)."""
    n = A.shape[0]
    B = B.reshape(-1, n).contiguous()
    nrhs = B.shape[0]
    X = X if X is not None else torch.zeros_like(B)
    pi = L.PcgInfo()
    opts = opts or pcg_opts()
    _pre(h)
    rc = h.lib.mpc_block_pcg(h.h, _ptr(_dev(A)), n, _ptr(_dev(Lf)), n, n, nb, nrhs,
                             _ptr(_dev(B)), n, _ptr(_dev(X)), n, C.byref(opts), C.byref(pi))
    check(rc, "mpc_block_pcg", allow=(L.ERR_NOT_CONVERGED,))
    return X, pi, rc


def pcg_last_coeffs(h, maxn=4096):
    """(alphas, betas) of the first right-hand side of the handle's last PCG
    solve, as numpy float64 arrays (the CG coefficients behind kappa_est)."""
    import numpy as np
    a = np.zeros(maxn)
    b = np.zeros(maxn)
    cnt = C.c_int(0)
    check(h.lib.mpc_pcg_last_coeffs(h.h, a.ctypes.data_as(C.POINTER(C.c_double)),
                                    b.ctypes.data_as(C.POINTER(C.c_double)), maxn, C.byref(cnt)),
          "mpc_pcg_last_coeffs")
    return a[:cnt.value].copy(), b[:cnt.value].copy()


def precond_cg(h, A, B, fopts=None, popts=None, X=None):
    """Factor a workspace copy of A, then PCG.  Returns (X, FactorInfo, PcgInfo, rc)."""
    n = A.shape[0]
    B = B.reshape(-1, n).contiguous()
    nrhs = B.shape[0]
    X = X if X is not None else torch.zeros_like(B)
    fi, pi = L.FactorInfo(), L.PcgInfo()
    fopts = fopts or factor_opts()
    popts = popts or pcg_opts()
    _pre(h)
    rc = h.lib.mpc_precond_cg(h.h, _ptr(_dev(A)), n, n, nrhs, _ptr(_dev(B)), n, _ptr(_dev(X)), n,
                              C.byref(fopts), C.byref(popts), C.byref(fi), C.byref(pi))
    if rc < 0:
        check(rc, "mpc_precond_cg", allow=(L.ERR_NOT_CONVERGED,))
    return X, fi, pi, rc


# ---------------------------------------------------------------- building blocks
def quantize_panel(h, P, prec_of_tile, tile_rows):
    """P: (k, m) tensor holding the column-major m x k panel.  Returns
    (payload (m + 128, 4k) uint8, scales (m, k/32) uint8, tile formats).  An
    8-bit tile's payload is in the matrix-core operand order (whole 32-row
    blocks, hence the spare rows); other tiles are row-major."""
    k, m = P.shape
    pt = torch.as_tensor(list(prec_of_tile), dtype=torch.uint8).to(h.device)
    pay = torch.zeros(m + 128, 4 * k, dtype=torch.uint8, device=h.device)
    sc = torch.zeros(m + 128, k // 32, dtype=torch.uint8, device=h.device)
    _pre(h)
    check(h.lib.mpc_quantize_panel(h.h, _ptr(_dev(P)), m, k, m, tile_rows, _ptr(pt), _ptr(pay),
                                   4 * k, _ptr(sc)), "mpc_quantize_panel")
    return pay, sc[:m], pt


def syrk(h, Cm, payload, scales, k, tile_rows, prec_of_tile):
    """Cm (m, m) column-major, lower part updated in place: C -= Q Q^T."""
    m = Cm.shape[0]
    pt = prec_of_tile if torch.is_tensor(prec_of_tile) else torch.as_tensor(
        list(prec_of_tile), dtype=torch.uint8).to(h.device)
    _pre(h)
    check(h.lib.mpc_syrk(h.h, _ptr(_dev(Cm)), m, m, k, _ptr(payload), payload.stride(0),
                         _ptr(scales), tile_rows, _ptr(pt)), "mpc_syrk")
    h.sync()


def apply_precond(h, Lf, R):
    """Z = (L L^T)^{-1} R; R: (nrhs, n) float64, Lf the factor tensor."""
    n = Lf.shape[0]
    R = R.reshape(-1, n).contiguous()
    Z = torch.empty_like(R)
    _pre(h)
    check(h.lib.mpc_apply_precond(h.h, _ptr(_dev(Lf)), n, n, R.shape[0], _ptr(R), n, _ptr(Z), n),
          "mpc_apply_precond")
    return Z


def potrf_block(h, A):
    n = A.shape[0]
    info = C.c_int(0)
    _pre(h)
    rc = h.lib.mpc_potrf_block(h.h, _ptr(_dev(A)), n, n, C.byref(info))
    if rc < 0:
        check(rc, "mpc_potrf_block")
    return info.value


def convert_f8(h, x, prec):
    y = torch.empty(x.numel(), dtype=torch.uint8, device=h.device)
    _pre(h)
    check(h.lib.mpc_convert_f8(h.h, _ptr(_dev(x)), _ptr(y), x.numel(), prec), "mpc_convert_f8")
    h.sync()
    return y


# ---------------------------------------------------------------- GP loss / backward
def gp_opts(n_probes=32, probe_kind=0, probe_seed=20240601, factor=None, pcg=None):
    o = L.GpOpts()
    L.load().mpc_gp_opts_default(C.byref(o))
    o.n_probes, o.probe_kind, o.probe_seed = n_probes, probe_kind, probe_seed
    if factor is not None:
        o.factor = factor
    if pcg is not None:
        o.pcg = pcg
    return o


def gp_loss_grad(h, X, ls, variance, noise, b, opts=None):
    """loss = b^T A^-1 b / 2 + logdet(A) / 2 for the RBF kernel matrix A of the
    points X (n, d) float32, and its gradient.  Returns (loss, grad_b (device
    float64), grad_ls (numpy), grad_variance, grad_noise, GpInfo)."""
    import numpy as np
    n, d = X.shape
    ls = np.atleast_1d(np.asarray(ls, np.float64))
    arr = (C.c_double * len(ls))(*ls)
    gls = (C.c_double * len(ls))()
    gvar, gnoise = C.c_double(0), C.c_double(0)
    b = b.reshape(n).contiguous()
    gb = torch.empty_like(b)
    info = L.GpInfo()
    opts = opts or gp_opts()
    _pre(h)
    check(h.lib.mpc_gp_loss_grad(h.h, _ptr(_dev(X)), n, d, d, arr, len(ls), float(variance),
                                 float(noise), _ptr(_dev(b)), C.byref(opts), _ptr(gb), gls,
                                 C.byref(gvar), C.byref(gnoise), C.byref(info)),
          "mpc_gp_loss_grad")
    return info.loss, gb, np.array(gls[:], np.float64), gvar.value, gnoise.value, info


def probes(h, n, nprobe, seed, unit=False):
    Z = torch.empty(nprobe, n, dtype=torch.float64, device=h.device)
    _pre(h)
    check(h.lib.mpc_probes(h.h, _ptr(Z), n, n, nprobe, seed, 1 if unit else 0), "mpc_probes")
    h.sync()
    return Z


def rbf_assemble64(h, X, ls, variance=1.0, nugget=0.0, want32=False):
    n, d = X.shape
    ls = [float(v) for v in (ls if hasattr(ls, "__len__") else [ls])]
    arr = (C.c_double * len(ls))(*ls)
    K = torch.empty(n, n, dtype=torch.float64, device=h.device)
    K32 = torch.empty(n, n, dtype=torch.float32, device=h.device) if want32 else None
    _pre(h)
    check(h.lib.mpc_rbf_assemble64(h.h, _ptr(_dev(X)), n, d, d, arr, len(ls), variance, nugget,
                                   _ptr(K), n, _ptr(K32), n), "mpc_rbf_assemble64")
    return (K, K32) if want32 else K


def pcg64(h, A64, Lf, B, nb=256, opts=None, X=None):
    """PCG with the FP64 operator A64 and the FP32 factor Lf.  B: (nrhs, n) float64."""
    n = A64.shape[0]
    B = B.reshape(-1, n).contiguous()
    nrhs = B.shape[0]
    X = X if X is not None else torch.zeros_like(B)
    pi = L.PcgInfo()
    opts = opts or pcg_opts()
    _pre(h)
    rc = h.lib.mpc_pcg64(h.h, _ptr(_dev(A64)), n, _ptr(_dev(Lf)), n, n, nb, nrhs, _ptr(_dev(B)), n,
                         _ptr(_dev(X)), n, C.byref(opts), C.byref(pi))
    check(rc, "mpc_pcg64", allow=(L.ERR_NOT_CONVERGED,))
    return X, pi, rc


class GPLoss(torch.autograd.Function):
    """loss(ls, variance, noise, b) = b^T A^-1 b / 2 + logdet(A) / 2 for fixed
    points X, with the hand-written backward of libmpchol (adjoint PCG +
    fused dK/dtheta contraction, Hutchinson logdet gradient).  The forward
    computes the gradient too; backward scales it by the incoming gradient."""

    @staticmethod
    def forward(ctx, ls, variance, noise, b, X, handle, opts):
        loss, gb, gls, gvar, gnoise, info = gp_loss_grad(
            handle, X, ls.detach().cpu().numpy(), float(variance), float(noise), b.detach(), opts)
        ctx.save_for_backward(gb, torch.as_tensor(gls, dtype=ls.dtype, device=ls.device),
                              torch.tensor(gvar, dtype=variance.dtype, device=variance.device),
                              torch.tensor(gnoise, dtype=noise.dtype, device=noise.device))
        ctx.info = info
        return torch.tensor(loss, dtype=torch.float64, device=b.device)

    @staticmethod
    def backward(ctx, g):
        gb, gls, gvar, gnoise = ctx.saved_tensors
        gg = g.to(torch.float64)
        return (gg.to(gls.device) * gls, gg.to(gvar.device) * gvar, gg.to(gnoise.device) * gnoise,
                gg * gb, None, None, None)


def gp_loss(ls, variance, noise, b, X, handle, opts=None):
    """Differentiable GP loss (see GPLoss); ls is a 1-element or d-element tensor."""
    return GPLoss.apply(ls, variance, noise, b, X, handle, opts)


class SpdSolve(torch.autograd.Function):
    """X = A^-1 B for a fixed SPD A (FP32, device) and right-hand sides B
    ((nrhs, n) float64), differentiable with respect to B.  Forward: PCG
    preconditioned by the mixed-precision factor Lf of A.  Backward: A is
    symmetric, so dL/dB = A^-1 dL/dX, one more PCG with the same factor (the
    adjoint solve; no new factorisation)."""

    @staticmethod
    def forward(ctx, B, A, Lf, handle, nb, opts):
        X, pi, rc = pcg(handle, A, Lf, B, nb=nb, opts=opts)
        if rc != 0:
            raise RuntimeError(f"SpdSolve: PCG did not converge ({info_dict(pi)})")
        ctx.A, ctx.Lf, ctx.handle, ctx.nb, ctx.opts = A, Lf, handle, nb, opts
        ctx.info = pi
        return X.reshape(B.shape)

    @staticmethod
    def backward(ctx, G):
        Gb, pi, rc = pcg(ctx.handle, ctx.A, ctx.Lf, G.contiguous(), nb=ctx.nb, opts=ctx.opts)
        if rc != 0:
            raise RuntimeError(f"SpdSolve: adjoint PCG did not converge ({info_dict(pi)})")
        return Gb.reshape(G.shape), None, None, None, None, None


def spd_solve(B, A, Lf, handle, nb=256, opts=None):
    """Differentiable solve A X = B with the factor Lf (see SpdSolve)."""
    return SpdSolve.apply(B, A, Lf, handle, nb, opts or pcg_opts())


# ---------------------------------------------------------------- options file and idx data
def settings_parse(text):
    """Parse the JSON text of a settings file (the reference's settings.json
    keys) into an L.Settings structure."""
    lib = L.load()
    raw = text.encode() if isinstance(text, str) else bytes(text)
    s = L.Settings()
    check(lib.mpc_settings_parse(raw, len(raw), C.byref(s)), "mpc_settings_parse")
    return s


def settings_read(path):
    s = L.Settings()
    check(L.load().mpc_settings_read(str(path).encode(), C.byref(s)), "mpc_settings_read")
    return s


def settings_options(s, mo=None, fo=None):
    """(matgen_opts, factor_opts) configured from a settings structure:
    defaults where the file says nothing, the file's values applied on top."""
    lib = L.load()
    if mo is None:
        mo = L.MatgenOpts()
        lib.mpc_matgen_opts_default(C.byref(mo))
    if fo is None:
        fo = factor_opts()
    check(lib.mpc_settings_apply(C.byref(s), C.byref(mo), C.byref(fo)), "mpc_settings_apply")
    return mo, fo


def idx_header(path):
    """(type_code, dims tuple, data_offset) of a big-endian idx file."""
    t, nd = C.c_int(), C.c_int()
    dims = (C.c_int64 * 4)()
    off = C.c_int64()
    check(L.load().mpc_idx_header(str(path).encode(), C.byref(t), C.byref(nd), dims, C.byref(off)),
          "mpc_idx_header")
    return t.value, tuple(dims[i] for i in range(nd.value)), off.value


def idx_read_u8(path, count=-1):
    """The first count items (all if count < 0) of an unsigned-byte idx file
    as a numpy uint8 array of shape (items, *item_dims)."""
    import numpy as np
    _, dims, _ = idx_header(path)
    m = dims[0] if count < 0 else min(count, dims[0])
    per = int(np.prod(dims[1:])) if len(dims) > 1 else 1
    out = np.empty(m * per, np.uint8)
    items = C.c_int64()
    check(L.load().mpc_idx_read_u8(str(path).encode(), m, out.ctypes.data_as(C.c_void_p),
                                   out.size, C.byref(items)), "mpc_idx_read_u8")
    return out.reshape((items.value,) + tuple(dims[1:]))


def idx_points(h, path, n=-1):
    """Images of an idx3 file as an (n, rows*cols) FP32 device tensor of
    pixel / 255 (the real-data form of the image-shaped RBF points)."""
    _, dims, _ = idx_header(path)
    m = dims[0] if n < 0 else min(n, dims[0])
    d = dims[1] * dims[2]
    X = torch.empty(m, d, dtype=torch.float32, device=h.device)
    n_out, d_out = C.c_int64(), C.c_int()
    _pre(h)
    check(h.lib.mpc_idx_points(h.h, str(path).encode(), m, _ptr(X), d, C.byref(n_out),
                               C.byref(d_out)), "mpc_idx_points")
    return X
