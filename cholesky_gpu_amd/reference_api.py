"""The reference's driver functions, same names and argument meaning, on libmpchol.

    generatePSD(dA, n, condVal, dist)            matgen.cuh:188-194
    switching_chol(d_A, n, b, d_diagVals, updated_diag, eps_prime,
                   perturb_diag, microscal_size, scales)  switching_chol.cuh:19-22
    precond_CG(A, x, b, n, r, normA, eps_prime, flr, perturb_diag,
               microscal_size)                   mixed_cholesky.cu:76-78

Differences that follow from the drop-in boundary: the cuBLAS/cuSOLVER
handles and stream are replaced by an mpc Handle (created on demand and never
destroyed by these calls, unlike precond_CG at mixed_cholesky.cu:220-223);
matrices are FP32 device tensors (generatePSD's FP64 output is produced in
FP32 and widened, precond_CG's FP64 A is narrowed to FP32 exactly as
mixed_cholesky.cu:122 does); errors come back as return codes instead of
exit().  precond_CG returns 0 when converged and -1 when not, like the
reference; its rel_tol is the reference's hard-coded 0 unless tol is given.
"""
import enum

import torch

from . import _lib as L
from . import solver as S


class DistType(enum.IntEnum):
    """kernels_macros.cuh:24-29 (same ordinals)."""
    RandomLogUniform = 0
    Clustered = 1
    Arithmetic = 2
    Geometric = 3


_default = {}


def _handle(h):
    if h is not None:
        return h
    dev = torch.cuda.current_device()
    if dev not in _default:
        _default[dev] = S.Handle(dev)
    return _default[dev]


def generatePSD(dA, n, condVal, dist, h=None, seed=123444):
    """Fill dA (n x n, float32 or float64 device tensor) with an SPD matrix whose
    eigenvalues follow `dist` with condition number condVal (lambda_max = 1).
    Q is a product of four Householder reflectors instead of the QR of a
    random matrix; the spectrum semantics of construct_diag_* are kept."""
    h = _handle(h)
    A32 = dA if dA.dtype == torch.float32 else torch.empty(n, n, dtype=torch.float32,
                                                           device=dA.device)
    S.matgen_householder(h, n, cond=condVal, lam_max=1.0, seed=seed, dist=int(dist), out=A32)
    if A32 is not dA:
        dA.copy_(A32.to(dA.dtype))


def switching_chol(d_A, n, b, d_diagVals=None, updated_diag=None, eps_prime=0.25,
                   perturb_diag=False, microscal_size=32, scales=None, h=None,
                   ladder=(L.PREC_E4M3, L.PREC_FP32)):
    """In-place factor of d_A (lower).  updated_diag receives the Schur-diagonal
    tracker; d_diagVals (if given) receives diag(A) before the factor.
    microscal_size must be 32 (the OCP MX block).  Returns the LAPACK-style
    info (0 = success)."""
    if microscal_size != 32:
        raise ValueError("microscal_size must be 32 (OCP MX block size)")
    h = _handle(h)
    if d_diagVals is not None:
        d_diagVals.copy_(torch.diagonal(d_A))
    opts = S.factor_opts(nb=b, update_prec=L.PREC_SWITCH, ladder=ladder, eps_prime=eps_prime,
                         fallback=1 if perturb_diag else 0)
    rc, fi, _, dg = S.factor(h, d_A, opts, want_map=False, want_diag=updated_diag is not None)
    if updated_diag is not None:
        updated_diag.copy_(dg)
    return rc


def precond_CG(A, x, b, n, r, normA=0.0, eps_prime=0.25, flr=0.0, perturb_diag=True,
               microscal_size=32, h=None, tol=None, maxit=200):
    """Factor + PCG as in mixed_cholesky.cu:76-229.  A: device (n, n) float64 or
    float32; x, b: device float64 vectors.  x receives the solution.
    Returns 0 if converged, -1 if not, or the factor's info > 0."""
    h = _handle(h)
    A32 = A if A.dtype == torch.float32 else A.to(torch.float32)
    fo = S.factor_opts(nb=r, update_prec=L.PREC_SWITCH, eps_prime=eps_prime,
                       fallback=1 if perturb_diag else 0)
    po = S.pcg_opts(tol=tol if tol is not None else 1e-300, maxit=maxit, stop_rule=1, x0_mode=0)
    X, fi, pi, rc = S.precond_cg(h, A32, b.reshape(1, n).to(torch.float64), fo, po)
    x.copy_(X.reshape(x.shape))
    if rc > 0:
        return rc
    return 0 if rc == 0 else -1
