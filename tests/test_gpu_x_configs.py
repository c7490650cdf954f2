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
from cholesky_gpu_amd import accuracy as ACC
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
    rep = ACC.report(K, Lg, Xs, b, pi)
    say("C1 accuracy", rep)
    assert rep["bwd_err"] < 2 * bo + 1e-6
    # ||x - x*|| / ||x*|| <= cond(K) * relres; cond(K) from the FP64 spectrum
    ev = np.linalg.eigvalsh(K64)
    assert rep["fwd_err"] <= ev.max() / ev.min() * rep["relres"] * 1.01, rep


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
    # accuracy report: the same factor in a buffer of its own (the solve's
    # lives in the library's workspace), forward error against the device's
    # FP64 Cholesky solve (a host one at n = 60000 would take minutes)
    Lf = K.clone()
    rc2, fi2, _, _ = S.factor(handle, Lf, fo)
    assert rc2 == 0 and list(fi2.syrk_flops) == list(fi.syrk_flops)
    rep = ACC.report(K, Lf, Xs, B, pi)
    say("C3 accuracy", rep)
    assert rep["bwd_err"] < 1e-2, rep
    assert rep["fwd_err"] < 1e-3, rep
    assert 1.0 <= rep["kappa_est"] < 1e4


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


def test_c5_gradient_wrt_b(handle):
    """C5's gradients with respect to b: loss = 1/2 sum_j b_j^T A^-1 b_j over
    the 16 right-hand sides at n = 65536, differentiated through the
    hand-written backward of solver.spd_solve (an adjoint PCG with the same
    mixed-precision factor).  The exact gradient is A^-1 B = X: the autograd
    gradient equals it to the solve's accuracy, and a central difference of
    the loss along a random direction D matches <D, grad>."""
    n, nrhs = 65536, 16
    say("C5 grad b", n, nrhs)
    A, lam = S.matgen_householder(handle, n, cond=1e4, lam_max=1.0, seed=65536, n_cluster=64,
                                  cluster_lo=1e-8, cluster_hi=1.1e-8)
    fo = S.factor_opts(nb=256, update_prec=L.PREC_SWITCH,
                       ladder=(L.PREC_E5M2, L.PREC_MXE4M3, L.PREC_BF16, L.PREC_FP32),
                       eps_prime=0.125)
    Lf = A.clone()
    rc, fi, _, _ = S.factor(handle, Lf, fo)
    assert rc == 0
    po = S.pcg_opts(tol=1e-10, maxit=1000)
    B = torch.from_numpy(np.random.default_rng(5).standard_normal((nrhs, n))).cuda()
    Bv = B.clone().requires_grad_(True)
    X = S.spd_solve(Bv, A, Lf, handle, opts=po)
    loss = 0.5 * (Bv * X).sum()
    loss.backward()
    G = Bv.grad.detach()
    Xd = X.detach()
    rel = float((G - Xd).norm() / Xd.norm())
    say("C5 grad b: ||grad - X|| / ||X||", rel)
    # both are A^-1 B to the PCG tolerance: residuals of 1e-10 each, which
    # the conditioning (1e8) can turn into up to 1e-2 in the solution
    assert rel < 2e-2, rel
    assert np.all(relres_fp64(A, G, B) <= 1e-10)
    # the loss is exactly quadratic in B, so a central difference has no
    # truncation error at any step; a unit step keeps the solves' own
    # errors (|x^T r| <= |x| |b| tol per column) far below <D, grad>
    D = torch.from_numpy(np.random.default_rng(6).standard_normal((nrhs, n))).cuda()
    eps = 1.0

    def f(Bp):
        with torch.no_grad():
            return float(0.5 * (Bp * S.spd_solve(Bp, A, Lf, handle, opts=po)).sum())
    fd = (f(B + eps * D) - f(B - eps * D)) / (2 * eps)
    dd = float((D * G).sum())
    say("C5 grad b: central difference", fd, "<D, grad>", dd)
    assert abs(fd - dd) <= 1e-5 * abs(dd), (fd, dd)
    del Bv, X, G
    rep = ACC.report(A, Lf, Xd, B)
    say("C5 accuracy", rep)
    lamh = lam.cpu().numpy()
    # ||x - x*|| / ||x*|| <= cond(A) * relres (cond from the generated spectrum)
    assert rep["fwd_err"] <= lamh.max() / lamh.min() * rep["relres"] * 1.01, rep
    assert rep["bwd_err"] < 1e-2, rep
