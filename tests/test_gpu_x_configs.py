"""The five workloads at their full sizes on the GPU, through the C ABI.

C1 (n = 2048 RBF, FP16 update, tol 1e-8) and C4 (n = 8192 GP) are small
enough for the oracle to run beside them.  At C2 (n = 16384), C3 (60000
image-shaped points, MXFP8) and C5 (n = 65536, cond 1e8, four-format
switching) the checks are size-independent properties: the factor
succeeds, every right-hand side reaches the FP64 true residual (computed here
in FP64 from the FP32 matrix, in row blocks), the iteration count stays
bounded, and the generated matrices have their defining features (unit
diagonal plus nugget, symmetry, the prescribed spectrum's trace)."""
import numpy as np
import pytest
import torch

from cholesky_gpu_amd import _lib as L
from cholesky_gpu_amd import solver as S

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def relres_fp64(A, X, B, rows=2048):
    """||b_j - A x_j|| / ||b_j|| for every column, A (n x n, FP32, symmetric,
    on the GPU) taken row block by row block in FP64."""
    n = A.shape[0]
    Xt = X.reshape(-1, n).to(torch.float64).T.contiguous()         # n x nrhs
    Bt = B.reshape(-1, n).to(torch.float64).T
    num = torch.zeros(Xt.shape[1], dtype=torch.float64, device=A.device)
    for r0 in range(0, n, rows):
        r1 = min(n, r0 + rows)
        R = Bt[r0:r1] - A[r0:r1].to(torch.float64) @ Xt
        num += (R * R).sum(0)
    return (num.sqrt() / Bt.norm(dim=0)).cpu().numpy()


def test_c1_rbf_fp16_update(handle, oracle):
    """C1: 2048 points ~ N(0, I_16), lengthscale 4, nugget 1e-2, block 128,
    FP16 trailing update, PCG to 1e-8; the oracle factors and solves the same
    matrix."""
    n, d = 2048, 16
    say("C1", n, d)
    X = S.rbf_points(handle, n, d, kind=0, seed=11)
    K = S.rbf_assemble(handle, X, [4.0], 1.0, 1e-2)
    b = torch.from_numpy(np.random.default_rng(1).standard_normal((1, n))).cuda()
    fo = S.factor_opts(nb=128, update_prec=L.PREC_FP16, ladder=(L.PREC_FP16,), fallback=0)
    Xs, fi, pi, rc = S.precond_cg(handle, K, b, fo, S.pcg_opts(tol=1e-8, maxit=200))
    assert rc == 0, (rc, S.info_dict(fi), S.info_dict(pi))
    assert relres_fp64(K, Xs, b)[0] <= 1e-8
    Kh = K.cpu().numpy()
    assert np.allclose(np.diag(Kh), 1.01, rtol=1e-6)
    # the oracle's factor of the same matrix with the same (fixed) update format
    Lg = K.clone()
    rc2, _, _, _ = S.factor(handle, Lg, fo)
    assert rc2 == 0
    Lo, rco, _, _, _ = oracle.factor_mixed(Kh, nb=128, update_prec=L.PREC_FP16,
                                           ladder=(L.PREC_FP16,), tile_rows=128)
    assert rco == 0
    Lgh = np.tril(Lg.cpu().numpy().T)
    # an FP16 rounding decision can flip between the two update orders (the
    # GPU sums a group's products before subtracting), moving an element by
    # up to one FP16 ulp (2^-11): a bar of about 4 u_fp16
    rel = np.linalg.norm(Lgh - np.tril(Lo)) / np.linalg.norm(np.tril(Lo))
    assert rel < 2e-3, rel
    K64 = Kh.astype(np.float64)
    bg = np.linalg.norm(Lgh.astype(np.float64) @ Lgh.T.astype(np.float64) - K64) / np.linalg.norm(K64)
    Lo64 = np.tril(Lo).astype(np.float64)
    bo = np.linalg.norm(Lo64 @ Lo64.T - K64) / np.linalg.norm(K64)
    assert bg < 2 * bo + 1e-6, (bg, bo)
    _, it_o, _, _ = oracle.pcg(Kh, Lgh, b.cpu().numpy()[0], tol=1e-8, maxit=200)
    assert abs(pi.iters - it_o) <= 2, (pi.iters, it_o)


def test_c2_conditioned_e4m3(handle):
    """C2: n = 16384, A = H Lambda H^T, Lambda logspace(0, 6), block 256, E4M3
    with FP32 switching, PCG to a 1e-10 FP64 true residual."""
    n = 16384
    say("C2", n)
    A, lam = S.matgen_householder(handle, n, cond=1e6, lam_max=1e6, seed=20240531)
    assert abs(float(torch.diagonal(A).double().sum()) / float(lam.sum()) - 1.0) < 1e-5
    B = torch.from_numpy(np.random.default_rng(7).standard_normal((1, n))).cuda()
    fo = S.factor_opts(nb=256, update_prec=L.PREC_SWITCH, ladder=(L.PREC_E4M3, L.PREC_FP32),
                       eps_prime=0.125)
    Xs, fi, pi, rc = S.precond_cg(handle, A, B, fo, S.pcg_opts(tol=1e-10, maxit=500))
    assert rc == 0, (rc, S.info_dict(pi))
    assert fi.syrk_flops[L.PREC_E4M3] > 0.5 * sum(fi.syrk_flops)
    assert pi.iters <= 30
    assert relres_fp64(A, Xs, B)[0] <= 1e-10


def test_c3_image_rbf_mxfp8(handle):
    """C3: 60000 points in 784 dims (each coordinate 0 with probability 0.8,
    else uniform in (0, 1]), lengthscale = median pairwise distance of a
    1024-point subsample, nugget 1e-3 (14.4 GB in FP32), block 256, MXFP8
    (E4M3 + E8M0 per 32) with FP32 switching, 4 right-hand sides to 1e-10."""
    n, d = 60000, 784
    say("C3", n, d)
    X = S.rbf_points(handle, n, d, kind=1, p_zero=0.8, seed=60000)
    ls = S.median_pairwise_distance(handle, X, nsub=1024, seed=3)
    assert ls > 0
    K = S.rbf_assemble(handle, X, [ls], 1.0, 1e-3)
    del X
    dg = torch.diagonal(K)
    assert torch.allclose(dg, torch.full_like(dg, 1.001), rtol=1e-6)
    idx = torch.randint(0, n, (2, 4096), generator=torch.Generator().manual_seed(0)).cuda()
    assert torch.equal(K[idx[0], idx[1]], K[idx[1], idx[0]])
    B = torch.from_numpy(np.random.default_rng(3).standard_normal((4, n))).cuda()
    fo = S.factor_opts(nb=256, update_prec=L.PREC_SWITCH, ladder=(L.PREC_MXE4M3, L.PREC_FP32),
                       eps_prime=0.125)
    Xs, fi, pi, rc = S.precond_cg(handle, K, B, fo, S.pcg_opts(tol=1e-10, maxit=500))
    say("C3 factor ms", fi.ms_total, "pcg ms", pi.ms_total, "iters", pi.iters)
    assert rc == 0, (rc, S.info_dict(fi), S.info_dict(pi))
    rr = relres_fp64(K, Xs, B)
    assert np.all(rr <= 1e-10), rr


def test_c4_gp_full_size(handle, oracle):
    """C4: 8192 points ~ N(0, I_8), ARD lengthscales in [0.5, 2], noise 1e-2,
    32 Rademacher probes: every gradient within 1e-6 of the FP64 oracle with
    the same probes."""
    n, d, P = 8192, 8, 32
    say("C4", n, d, P)
    rng = np.random.default_rng(8192)
    Xh = rng.standard_normal((n, d)).astype(np.float32)
    ls = rng.uniform(0.5, 2.0, d)
    b = rng.standard_normal(n)
    o = S.gp_opts(n_probes=P, probe_seed=17, pcg=S.pcg_opts(tol=1e-11, maxit=500))
    loss, gb, gls, gvar, gno, info = S.gp_loss_grad(handle, torch.from_numpy(Xh).cuda(), ls, 1.0,
                                                    1e-2, torch.from_numpy(b).cuda(), o)
    Z = oracle.probes(n, P, 17)
    rl, rb, rls, rvar, rno = oracle.gp_loss_grad_f64(Xh, ls, 1.0, 1e-2, b, Z=Z)
    rel = lambda a, r: np.linalg.norm(np.asarray(a) - r) / np.linalg.norm(r)
    assert rel(gb.cpu().numpy(), rb) < 1e-6
    assert rel(gls, rls) < 1e-6, (gls, rls)
    assert abs(gvar - rvar) <= 1e-6 * abs(rvar)
    assert abs(gno - rno) <= 1e-6 * abs(rno)


def test_c5_switching_at_scale(handle):
    """C5: n = 65536 (17 GB FP32), H Lambda H^T with 64 eigenvalues clustered
    near 1e-8 and the rest logspace(-4, 0) (cond 1e8), block 256, per-tile
    switching among E5M2 / MXFP8 / BF16 / FP32, 16 right-hand sides to 1e-10."""
    n = 65536
    say("C5", n)
    A, lam = S.matgen_householder(handle, n, cond=1e4, lam_max=1.0, seed=65536, n_cluster=64,
                                  cluster_lo=1e-8, cluster_hi=1.1e-8)
    lamh = lam.cpu().numpy()
    assert np.isclose(lamh.max() / lamh.min(), 1e8, rtol=0.11)
    assert abs(float(torch.diagonal(A).double().sum()) / float(lamh.sum()) - 1.0) < 1e-5
    B = torch.from_numpy(np.random.default_rng(5).standard_normal((16, n))).cuda()
    fo = S.factor_opts(nb=256, update_prec=L.PREC_SWITCH,
                       ladder=(L.PREC_E5M2, L.PREC_MXE4M3, L.PREC_BF16, L.PREC_FP32),
                       eps_prime=0.125)
    Xs, fi, pi, rc = S.precond_cg(handle, A, B, fo, S.pcg_opts(tol=1e-10, maxit=1000))
    say("C5 factor ms", fi.ms_total, "pcg ms", pi.ms_total, "iters", pi.iters,
        "flops by format", list(fi.syrk_flops)[:7])
    assert rc == 0, (rc, S.info_dict(fi), S.info_dict(pi))
    rr = relres_fp64(A, Xs, B)
    assert np.all(rr <= 1e-10), rr
