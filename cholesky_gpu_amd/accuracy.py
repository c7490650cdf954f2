"""FP64 accuracy reports for a solve: forward error, backward error of the
factor and the true residual.

These replace the reference's FP64 direct solve used for comparison
(solveWithCuSolver, src/mixed_cholesky.cu:297-349).  They are checks, not
part of the solve: the exact solution comes from an FP64 Cholesky of the
FP32-stored matrix promoted to FP64 (LAPACK on the host up to
HOST_MAX_N, the device's FP64 Cholesky beyond that, where a host
factorisation would take minutes).  Conventions follow solver.py: an n x n
device tensor holds a column-major matrix, a block of right-hand sides is
(nrhs, n).
"""
import numpy as np
import torch

HOST_MAX_N = 16384


def true_relres(A, X, B, rows=2048):
    """max_j ||b_j - A x_j|| / ||b_j|| with A (FP32, symmetric, device) taken
    in FP64 row block by row block."""
    n = A.shape[0]
    Xt = X.reshape(-1, n).to(torch.float64).T.contiguous()
    Bt = B.reshape(-1, n).to(torch.float64).T
    num = torch.zeros(Xt.shape[1], dtype=torch.float64, device=A.device)
    for r0 in range(0, n, rows):
        r1 = min(n, r0 + rows)
        R = Bt[r0:r1] - A[r0:r1].to(torch.float64) @ Xt
        num += (R * R).sum(0)
    return (num.sqrt() / Bt.norm(dim=0)).cpu().numpy()


BLK = 4096     # block edge of the FP64 checks (no tensor of 2^31 or more elements)


def _lblock(Lf, r0, r1, c0, c1):
    """L[r0:r1, c0:c1] in FP64 (zero above the diagonal) from the factor
    buffer Lf (tensor = L^T; its strict upper part holds A)."""
    B = Lf[c0:c1, r0:r1].T.to(torch.float64)
    if r0 < c1 and c0 < r1:          # a block on the diagonal
        B = torch.tril(B, diagonal=c0 - r0) if r0 == c0 else B * (
            torch.arange(r0, r1, device=Lf.device)[:, None] >=
            torch.arange(c0, c1, device=Lf.device)[None, :])
    return B


def backward_error(A, Lf):
    """||L L^T - A||_F / ||A||_F in FP64 for the lower triangle L of the
    factor buffer Lf (the strict upper part of Lf still holds A and is
    ignored), A the FP32 matrix that was factored.  Both are device tensors in
    the column-major convention (tensor = matrix^T).  Block pairs (I, J) of
    the lower block triangle: (L L^T)_IJ = sum_K L_IK L_JK^T over K <= J."""
    n = A.shape[0]
    nbk = (n + BLK - 1) // BLK
    num = den = 0.0
    for bi in range(nbk):
        r0, r1 = bi * BLK, min(n, (bi + 1) * BLK)
        for bj in range(bi + 1):
            c0, c1 = bj * BLK, min(n, (bj + 1) * BLK)
            P = torch.zeros(r1 - r0, c1 - c0, dtype=torch.float64, device=A.device)
            for bk in range(bj + 1):
                k0, k1 = bk * BLK, min(n, (bk + 1) * BLK)
                P += _lblock(Lf, r0, r1, k0, k1) @ _lblock(Lf, c0, c1, k0, k1).T
            Ab = A[c0:c1, r0:r1].T.to(torch.float64)      # A[I, J]
            if bi == bj:                                   # lower part of a diagonal block
                m = (torch.arange(r0, r1, device=A.device)[:, None] >=
                     torch.arange(c0, c1, device=A.device)[None, :])
                w = torch.where(m, 1.0, 0.0).to(torch.float64)
                w = 2.0 * w - torch.eye(r1 - r0, dtype=torch.float64, device=A.device)
            else:
                w = 2.0
            E = P - Ab
            num += float((w * E * E).sum())
            den += float((w * Ab * Ab).sum())
    return (num / den) ** 0.5


def _chol64_blocked(A):
    """Lower FP64 Cholesky of the FP32-stored A as column panels
    P_j = L[j0:, j0:j1] (left-looking, blocks of BLK), with the device's
    dense FP64 routines on blocks only."""
    n = A.shape[0]
    panels = []
    for j0 in range(0, n, BLK):
        j1 = min(n, j0 + BLK)
        P = A[j0:j1, j0:].T.to(torch.float64)            # A[j0:, j0:j1]
        for i, Pi in enumerate(panels):
            i0 = i * BLK
            P -= Pi[j0 - i0:] @ Pi[j0 - i0:j0 - i0 + (j1 - j0)].T
        Ljj = torch.linalg.cholesky(P[:j1 - j0])
        if j1 < n:
            P[j1 - j0:] = torch.linalg.solve_triangular(Ljj, P[j1 - j0:].T, upper=False).T
        P[:j1 - j0] = Ljj
        panels.append(P)
    return panels


def _chol64_solve(panels, n, Bm):
    """X = (L L^T)^-1 B for B (n x nrhs) with the panels of _chol64_blocked."""
    Y = Bm.clone()
    for j, P in enumerate(panels):                         # L y = b
        j0 = j * BLK
        j1 = min(n, j0 + BLK)
        Y[j0:j1] = torch.linalg.solve_triangular(P[:j1 - j0], Y[j0:j1], upper=False)
        if j1 < n:
            Y[j1:] -= P[j1 - j0:] @ Y[j0:j1]
    for j in range(len(panels) - 1, -1, -1):               # L^T x = y
        P = panels[j]
        j0 = j * BLK
        j1 = min(n, j0 + BLK)
        if j1 < n:
            Y[j0:j1] -= P[j1 - j0:].T @ Y[j1:]
        Y[j0:j1] = torch.linalg.solve_triangular(P[:j1 - j0].T, Y[j0:j1], upper=True)
    return Y


def fp64_solution(A, B):
    """X* = A^-1 B from an FP64 Cholesky of the FP32-stored A (host LAPACK for
    n <= HOST_MAX_N, else a blocked FP64 Cholesky on the device).  Returns a
    (nrhs, n) float64 device tensor and where it was computed."""
    n = A.shape[0]
    Bm = B.reshape(-1, n)
    if n <= HOST_MAX_N:
        import scipy.linalg as sl
        A64 = A.cpu().numpy().astype(np.float64)        # symmetric: layout immaterial
        c = sl.cho_factor(A64, lower=True, overwrite_a=True, check_finite=False)
        Xs = sl.cho_solve(c, Bm.cpu().numpy().T, check_finite=False).T
        return torch.from_numpy(np.ascontiguousarray(Xs)).to(A.device), "host"
    panels = _chol64_blocked(A)
    Xs = _chol64_solve(panels, n, Bm.T.contiguous().to(torch.float64)).T.contiguous()
    del panels
    return Xs, "device"


def forward_error(X, Xs):
    """max_j ||x_j - x*_j|| / ||x*_j||."""
    n = Xs.shape[-1]
    D = X.reshape(-1, n).to(torch.float64) - Xs
    return float((D.norm(dim=1) / Xs.norm(dim=1)).max())


def report(A, Lf, X, B, pcg_info=None):
    """Forward error, backward error, true residual and (when the PCG info
    is given) the Lanczos estimate of cond(M^-1 A), as a dict."""
    Xs, where = fp64_solution(A, B)
    out = {"fwd_err": forward_error(X, Xs), "fwd_ref": f"fp64 cholesky ({where})",
           "bwd_err": backward_error(A, Lf), "relres": float(true_relres(A, X, B).max())}
    if pcg_info is not None:
        out["kappa_est"] = float(pcg_info.kappa_est)
        out["iters"] = int(pcg_info.iters)
    return out
