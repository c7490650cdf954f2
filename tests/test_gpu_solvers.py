"""Iterative refinement and block PCG (mpc_refine, mpc_block_pcg) against the
oracle's restatements on the same FP32 factor, and the SYMV / sweep kernels
at right-hand-side counts that take several passes."""
import numpy as np
import pytest
import torch

from cholesky_gpu_amd import _lib as L  # noqa: E402
from cholesky_gpu_amd import solver as S  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def to_dev_cm(A):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(A, np.float32).T)).cuda()


def from_dev_cm(T):
    return T.cpu().numpy().T


def gpu_factor(handle, A, nb, ladder, eps):
    """The GPU's factor of A (lower, FP32) on the device and on the host."""
    T = to_dev_cm(A)
    fo = S.factor_opts(nb=nb, update_prec=L.PREC_SWITCH if len(ladder) > 1 else ladder[0],
                       ladder=ladder, eps_prime=eps)
    rc, _, _, _ = S.factor(handle, T, fo)
    assert rc == 0
    return T, np.tril(from_dev_cm(T))


def relres(A64, X, B):
    return max(np.linalg.norm(B[j] - A64 @ X[j]) / np.linalg.norm(B[j]) for j in range(len(B)))


@pytest.mark.parametrize("n,cond,nrhs", [(768, 1e3, 1), (1536, 1e4, 3)])
def test_refine_matches_oracle(handle, oracle, n, cond, nrhs):
    """Same number of corrections as the oracle on the GPU's factor (+-1),
    every column's FP64 true residual <= tol, solution equal to the oracle's
    to 1e-9.  (eps' = 0.25 keeps the spectral radius of I - M^-1 A near
    0.05, where refinement converges in about ten steps.)"""
    say("refine", n, cond, nrhs)
    A, _ = oracle.householder_spd(n, cond, cond, n + 1)
    Ld, Lh = gpu_factor(handle, A, 128, (L.PREC_E4M3, L.PREC_FP32), 0.25)
    B = np.random.default_rng(n).standard_normal((nrhs, n))
    X, pi, rc = S.refine(handle, to_dev_cm(A), Ld, torch.from_numpy(B).cuda(), nb=128,
                         opts=S.pcg_opts(tol=1e-12, maxit=40))
    assert rc == 0, S.info_dict(pi)
    Xh = X.cpu().numpy()
    assert relres(A.astype(np.float64), Xh, B) <= 1e-12 and pi.relres <= 1e-12
    its = []
    for j in range(nrhs):
        xo, it_o, rr_o, hist = oracle.refine(A, Lh, B[j], tol=1e-12, maxit=40)
        assert it_o >= 1
        its.append(it_o)
        assert np.abs(Xh[j] - xo).max() <= 1e-9 * np.abs(xo).max()
    # the step count is the slowest column's
    assert abs(pi.iters - max(its)) <= 1, (pi.iters, its)


def test_refine_not_converged_and_x0_modes(handle, oracle):
    n = 512
    A, _ = oracle.householder_spd(n, 1e3, 1e3, 9)
    Ld, Lh = gpu_factor(handle, A, 64, (L.PREC_E4M3, L.PREC_FP32), 0.25)
    B = torch.from_numpy(np.random.default_rng(3).standard_normal((1, n))).cuda()
    Ad = to_dev_cm(A)
    # maxit 1 cannot reach 1e-14: MPC_ERR_NOT_CONVERGED with the last residual
    X1, p1, rc1 = S.refine(handle, Ad, Ld, B, nb=64, opts=S.pcg_opts(tol=1e-15, maxit=1))
    assert rc1 == L.ERR_NOT_CONVERGED and p1.iters == 1 and p1.converged == 0
    # x0 = 0 takes at least as many steps as x0 = M^-1 b; x0 from X reuses the iterate
    _, p0, rc0 = S.refine(handle, Ad, Ld, B, nb=64, opts=S.pcg_opts(tol=1e-12, maxit=40))
    _, pz, rcz = S.refine(handle, Ad, Ld, B, nb=64, opts=S.pcg_opts(tol=1e-12, maxit=40, x0_mode=1))
    assert rc0 == 0 and rcz == 0 and pz.iters >= p0.iters
    X2 = X1.clone()
    _, p2, rc2 = S.refine(handle, Ad, Ld, B, nb=64, X=X2,
                          opts=S.pcg_opts(tol=1e-12, maxit=40, x0_mode=2))
    assert rc2 == 0 and p2.iters <= p0.iters


@pytest.mark.parametrize("n,nrhs", [(1024, 1), (1024, 4), (2304, 8), (3000, 16)])
def test_block_pcg_matches_oracle(handle, oracle, n, nrhs):
    """Block PCG on the GPU's factor: iterations within 1 of the oracle's
    block PCG, true residual <= tol, X equal to the oracle's to 1e-7."""
    say("block_pcg", n, nrhs)
    A, _ = oracle.householder_spd(n, 1e6, 1e6, n + nrhs)
    Ld, Lh = gpu_factor(handle, A, 128, (L.PREC_E4M3, L.PREC_FP32), 0.25)
    B = np.random.default_rng(nrhs).standard_normal((nrhs, n))
    X, pi, rc = S.block_pcg(handle, to_dev_cm(A), Ld, torch.from_numpy(B).cuda(), nb=128,
                            opts=S.pcg_opts(tol=1e-10, maxit=200))
    assert rc == 0, S.info_dict(pi)
    Xh = X.cpu().numpy()
    assert relres(A.astype(np.float64), Xh, B) <= 1e-10
    Xo, it_o, rel_o = oracle.block_pcg(A, Lh, B.T, tol=1e-10, maxit=200)
    assert it_o > 0 and abs(pi.iters - it_o) <= 1, (pi.iters, it_o)
    assert np.abs(Xh - Xo.T).max() <= 1e-7 * np.abs(Xo).max()
    if nrhs > 1:
        # no more steps than the slowest independent PCG
        its = [oracle.pcg(A, Lh, B[j], tol=1e-10, maxit=200)[1] for j in range(nrhs)]
        assert pi.iters <= max(its) + 1


@pytest.mark.parametrize("n,s", [(64, 8), (64, 16), (128, 16)])
def test_block_pcg_finite_termination(handle, oracle, n, s):
    """n distinct eigenvalues, identity preconditioner: at most ceil(n/s)
    block steps (single-vector CG needs about 20 on these matrices)."""
    say("block_pcg termination", n, s)
    lam = np.geomspace(1.0, 4.0, n).astype(np.float32)
    np.random.default_rng(n + s).shuffle(lam)
    A = np.diag(lam).astype(np.float32)
    Ld = to_dev_cm(np.eye(n, dtype=np.float32))
    B = np.random.default_rng(s).standard_normal((s, n))
    X, pi, rc = S.block_pcg(handle, to_dev_cm(A), Ld, torch.from_numpy(B).cuda(), nb=64,
                            opts=S.pcg_opts(tol=1e-10, maxit=100, x0_mode=1))
    assert rc == 0 and 0 < pi.iters <= -(-n // s), S.info_dict(pi)
    np.testing.assert_allclose(X.cpu().numpy(), B / lam.astype(np.float64)[None, :], rtol=1e-10)


def test_block_pcg_breakdown_falls_back(handle, oracle):
    """Dependent right-hand sides: the block recurrence breaks down (the
    oracle says so too) and the columns finish as independent PCGs."""
    n = 800
    A, _ = oracle.householder_spd(n, 1e4, 1e4, 21)
    Ld, Lh = gpu_factor(handle, A, 64, (L.PREC_FP32,), 0.25)
    B = np.random.default_rng(6).standard_normal((3, n))
    B[2] = B[0]
    assert oracle.block_pcg(A, Lh, B.T, tol=1e-10)[1] == -2
    X, pi, rc = S.block_pcg(handle, to_dev_cm(A), Ld, torch.from_numpy(B).cuda(), nb=64,
                            opts=S.pcg_opts(tol=1e-10, maxit=100))
    assert rc == 0 and pi.converged == 3, S.info_dict(pi)
    Xh = X.cpu().numpy()
    assert relres(A.astype(np.float64), Xh, B) <= 1e-10
    assert np.array_equal(Xh[0], Xh[2])
    with pytest.raises(RuntimeError):
        S.block_pcg(handle, to_dev_cm(A), Ld,
                    torch.zeros(L.BLOCK_PCG_MAX + 1, n, dtype=torch.float64).cuda(), nb=64)


@pytest.mark.parametrize("nrhs", [5, 9, 17])
def test_pcg_multi_pass_symv(handle, oracle, nrhs):
    """More right-hand sides than one SYMV pass holds (8): PCG on the FP32
    matrix equals an FP64 solve per column, each column as its own solve."""
    say("pcg multi-pass", nrhs)
    n = 1100
    A, _ = oracle.householder_spd(n, 1e4, 1e4, 31)
    Ld, Lh = gpu_factor(handle, A, 128, (L.PREC_E4M3, L.PREC_FP32), 0.25)
    B = np.random.default_rng(nrhs).standard_normal((nrhs, n))
    Ad = to_dev_cm(A)
    X, pi, rc = S.pcg(handle, Ad, Ld, torch.from_numpy(B).cuda(), nb=128,
                      opts=S.pcg_opts(tol=1e-10, maxit=200))
    assert rc == 0
    Xh = X.cpu().numpy()
    assert relres(A.astype(np.float64), Xh, B) <= 1e-10
    X1, _, rc1 = S.pcg(handle, Ad, Ld, torch.from_numpy(B[-1:].copy()).cuda(), nb=128,
                       opts=S.pcg_opts(tol=1e-10, maxit=200))
    assert rc1 == 0
    np.testing.assert_allclose(Xh[-1], X1.cpu().numpy()[0], rtol=0, atol=1e-8 * np.abs(Xh[-1]).max())
