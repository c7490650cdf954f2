"""Accuracy reports and error paths on the GPU, through the C ABI.

- Forward error against an FP64 Cholesky solve of the same FP32 matrix,
  backward error of the factor, and the Lanczos estimate of cond(M^-1 A)
  (the reference compares with an FP64 direct solve, mixed_cholesky.cu:297-349,
  and estimates the condition number with an eigensolve, matgen.cuh:9-95).
- At n = 16384 (the E4M3/FP32 switching factor and the FP32-update factor)
  the PCG iteration counts against the oracle's PCG on the GPU's own factor.
- The failure paths of the factor and the solve: NaN in a panel, PCG
  non-convergence, an FP8 breakdown recovered by the fallback, and FP16
  tiles whose values exceed the FP16 range.
"""
import numpy as np
import pytest
import torch

from cholesky_gpu_amd import _lib as L
from cholesky_gpu_amd import accuracy as ACC
from cholesky_gpu_amd import solver as S

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def to_dev_cm(A):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(A, np.float32).T)).cuda()


def hh(oracle, n, seed=5, cond=1e6, dist=L.DIST_LOGSPACED):
    A, _ = oracle.householder_spd(n, cond, cond, seed, dist=dist)
    return A


# ---------------------------------------------------------------- condition estimate
def test_kappa_est_matches_lanczos_and_spectrum(handle, oracle):
    """kappa_est = theta_max / theta_min of the Lanczos tridiagonal T_k built
    from the CG coefficients (matgen.cuh:9-95 computes cond with an
    eigensolve instead): equal to the oracle's Sturm bisection on the same
    coefficients and to a dense eigensolve of T_k; the Ritz values interlace
    the spectrum of L^-1 A L^-T, so the estimate never exceeds the exact
    cond(L^-1 A L^-T) and, after CG has converged, is close to it."""
    n = 768
    say("kappa_est", n)
    A = hh(oracle, n, seed=31, cond=1e2)
    Ad = to_dev_cm(A)
    Lf = Ad.clone()
    # an all-E4M3 factor: a visibly imperfect preconditioner
    fo = S.factor_opts(nb=256, update_prec=L.PREC_E4M3, ladder=(L.PREC_E4M3,), fallback=0)
    rc, fi, _, _ = S.factor(handle, Lf, fo)
    assert rc == 0
    B = torch.from_numpy(np.random.default_rng(2).standard_normal((1, n))).cuda()
    X, pi, rc = S.pcg(handle, Ad, Lf, B, nb=256, opts=S.pcg_opts(tol=1e-13, maxit=400))
    assert rc == 0, S.info_dict(pi)
    al, be = S.pcg_last_coeffs(handle)
    k = len(al)
    assert k >= 3 and k >= pi.iters - 1, (k, pi.iters)
    k_or = oracle.lib().orc_lanczos_cond(al.ctypes.data, be.ctypes.data, k)
    assert abs(pi.kappa_est - k_or) <= 1e-9 * k_or, (pi.kappa_est, k_or)
    # T_k: diagonal 1/a_i + b_{i-1}/a_{i-1}, off-diagonal sqrt(b_i)/a_i
    d = 1.0 / al + np.concatenate([[0.0], be[:-1] / al[:-1]])
    e = np.sqrt(be[:-1]) / al[:-1]
    th = np.linalg.eigvalsh(np.diag(d) + np.diag(e, 1) + np.diag(e, -1))
    assert abs(pi.kappa_est - th.max() / th.min()) <= 1e-8 * pi.kappa_est
    Lh = np.tril(Lf.cpu().numpy().T).astype(np.float64)
    import scipy.linalg as sl
    Y = sl.solve_triangular(Lh, A.astype(np.float64), lower=True)
    Mt = sl.solve_triangular(Lh, Y.T, lower=True)             # L^-1 A L^-T
    ev = np.linalg.eigvalsh((Mt + Mt.T) / 2)
    k_true = ev.max() / ev.min()
    say("kappa_est", pi.kappa_est, "exact", k_true, "iters", pi.iters)
    assert k_true > 1.01                     # the preconditioner is not exact
    assert 0.5 * k_true <= pi.kappa_est <= (1 + 1e-6) * k_true, (pi.kappa_est, k_true)


# ---------------------------------------------------------------- full-size C2 accuracy
@pytest.mark.parametrize("mode", ["e4m3_fp32", "fp32"])
def test_c2_accuracy_and_iterations(handle, oracle, mode):
    """n = 16384, cond 1e6: forward error <= cond(A) * tol (the perturbation
    bound for a residual of tol), backward error of the factor within its
    format's bar, and the GPU's PCG iteration count within 2 of the oracle's
    PCG run on the GPU's own factor."""
    n = 16384
    say("C2 accuracy", mode)
    A, lam = S.matgen_householder(handle, n, cond=1e6, lam_max=1e6, seed=20240531)
    B = torch.from_numpy(np.random.default_rng(7).standard_normal((1, n))).cuda()
    if mode == "fp32":
        fo = S.factor_opts(nb=256, update_prec=L.PREC_FP32, ladder=(L.PREC_FP32,), fallback=0)
    else:
        fo = S.factor_opts(nb=256, update_prec=L.PREC_SWITCH, ladder=(L.PREC_E4M3, L.PREC_FP32),
                           eps_prime=0.125)
    Lf = A.clone()
    rc, fi, _, _ = S.factor(handle, Lf, fo)
    assert rc == 0
    tol = 1e-10
    X, pi, rc = S.pcg(handle, A, Lf, B, nb=256, opts=S.pcg_opts(tol=tol, maxit=500))
    assert rc == 0, S.info_dict(pi)
    rep = ACC.report(A, Lf, X, B, pi)
    say("C2", mode, rep)
    kappa_a = float(lam.max() / lam.min())
    assert rep["relres"] <= tol
    assert rep["fwd_err"] <= kappa_a * tol, rep
    # the FP32-update factor: backward error at FP32 rounding level; the
    # 8-bit one: the switching rule bounds each tile's error by eps' times
    # the smallest pivot, far below the 8-bit unit roundoff times |A|
    assert rep["bwd_err"] < (1e-5 if mode == "fp32" else 1e-2), rep
    assert 1.0 <= rep["kappa_est"] < 1e3
    # the oracle's PCG (FP64 recurrence, same stopping rule) on the same L
    Ah = A.cpu().numpy()
    Lh = np.tril(Lf.cpu().numpy().T)
    _, it_o, rr_o, kap_o = oracle.pcg(Ah, Lh, B.cpu().numpy()[0], tol=tol, maxit=500)
    say("C2", mode, "gpu iters", pi.iters, "oracle iters", it_o, "oracle kappa", kap_o)
    assert it_o > 0
    assert abs(pi.iters - it_o) <= 2, (pi.iters, it_o)


# ---------------------------------------------------------------- error paths
def test_nan_in_panel_reports_nan(handle, oracle):
    """A NaN below the diagonal block (reaching the panel TRSM) or on the
    diagonal (a NaN pivot): the factor stops with MPC_ERR_NAN, with and
    without the fallback (a retry cannot remove a NaN from A)."""
    n = 768
    say("nan panel")
    A = hh(oracle, n, seed=3, cond=1e3)
    A[600, 10] = np.nan
    A[10, 600] = np.nan
    D = hh(oracle, n, seed=3, cond=1e3)
    D[300, 300] = np.nan                    # a NaN pivot
    for fb, A in ((0, A), (1, A), (0, D), (1, D)):
        T = to_dev_cm(A)
        fi = L.FactorInfo()
        S._pre(handle)
        rc = handle.lib.mpc_factor(handle.h, S._ptr(T), n, n,
                                   S.C.byref(S.factor_opts(nb=256, fallback=fb)), None, None,
                                   S.C.byref(fi))
        assert rc == L.ERR_NAN, (fb, rc)
        assert "NaN" in L.load().mpc_last_error_string().decode()


def test_pcg_not_converged(handle, oracle):
    """maxit = 1 at tol 1e-14: MPC_ERR_NOT_CONVERGED from mpc_precond_cg, and
    the reference's return convention (-1) from reference_api.precond_CG."""
    from cholesky_gpu_amd import reference_api as R
    n = 1024
    say("pcg not converged")
    A = hh(oracle, n, seed=4, cond=1e6)
    Ad = to_dev_cm(A)
    B = torch.from_numpy(np.random.default_rng(4).standard_normal((1, n))).cuda()
    fo = S.factor_opts(nb=256, update_prec=L.PREC_SWITCH, ladder=(L.PREC_E4M3, L.PREC_FP32),
                       eps_prime=0.25, fallback=0)
    X, fi, pi, rc = S.precond_cg(handle, Ad, B, fo, S.pcg_opts(tol=1e-14, maxit=1))
    assert rc == L.ERR_NOT_CONVERGED, rc
    assert pi.converged == 0 and pi.iters == 1
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    rc2 = R.precond_CG(Ad.T.contiguous().to(torch.float64), x, B[0], n, 256, h=handle, tol=1e-14,
                       maxit=1)
    assert rc2 == -1


def test_fp8_breakdown_recovered_by_fallback(handle, oracle):
    """cond 1e6 with every trailing update in E4M3 (eps' = 1e12 admits E4M3
    everywhere, also after the two eps'/16 retries) loses positive
    definiteness; the fallback restores A, retries with eps'/16 twice, then
    with FP32 updates, and the solve converges.
    Through mpc_precond_cg with agg_panels = 2 and n > 2 nb, whose first
    attempt reads A through the deferred copy: the answer is bitwise that of
    factoring a full copy with the same options and then solving."""
    n = 1280
    say("fp8 breakdown + fallback")
    A, _ = oracle.householder_spd(n, 1e6, 1e6, 12)
    Ad = to_dev_cm(A)
    B = torch.from_numpy(np.random.default_rng(12).standard_normal((2, n))).cuda()
    ladder = (L.PREC_E4M3, L.PREC_FP32)
    # without the fallback the all-E4M3 schedule breaks down
    T = to_dev_cm(A)
    rc0, fi0, _, _ = S.factor(handle, T, S.factor_opts(nb=256, update_prec=L.PREC_SWITCH,
                                                        ladder=ladder, eps_prime=1e12, fallback=0))
    assert rc0 > 0, rc0
    fo = S.factor_opts(nb=256, update_prec=L.PREC_SWITCH, ladder=ladder, eps_prime=1e12,
                       fallback=1, agg_panels=2)
    po = S.pcg_opts(tol=1e-10, maxit=500)
    X1, fi1, pi1, rc1 = S.precond_cg(handle, Ad, B, fo, po)
    say("fallback retries", fi1.retries, "eps used", fi1.eps_prime_used, "iters", pi1.iters)
    assert rc1 == 0, (rc1, S.info_dict(fi1), S.info_dict(pi1))
    assert fi1.retries == 3 and fi1.eps_prime_used == 1e12 / 256
    assert np.all(ACC.true_relres(Ad, X1, B) <= 1e-10)
    Lf = Ad.clone()
    rc2, fi2, _, _ = S.factor(handle, Lf, fo)
    assert rc2 == 0 and fi2.retries == fi1.retries
    X2, pi2, rc3 = S.pcg(handle, Ad, Lf, B, nb=256, opts=po)
    assert rc3 == 0 and pi2.iters == pi1.iters
    assert torch.equal(X1, X2)
    assert torch.equal(Ad, to_dev_cm(A))


def test_fp16_range_excluded_by_rule(handle, oracle):
    """A matrix scaled by 1e10: some panel tiles hold entries beyond FP16's
    largest finite value (65504), so the switching rule gives those tiles the
    next format of the ladder instead of overflowing, and the factor succeeds
    without a retry; the oracle's rule makes the same choices."""
    n = 768
    say("fp16 range")
    A = hh(oracle, n, seed=21, cond=1e3) * np.float32(1e10)
    T = to_dev_cm(A)
    fo = S.factor_opts(nb=256, update_prec=L.PREC_SWITCH, ladder=(L.PREC_FP16, L.PREC_FP32),
                       eps_prime=1e6, fallback=0)
    rc, fi, mp, _ = S.factor(handle, T, fo, want_map=True)
    assert rc == 0, rc
    gmap = mp.cpu().numpy().T
    _, rco, omap, _, _ = oracle.factor_mixed(A, nb=256, update_prec=L.PREC_SWITCH,
                                             ladder=(L.PREC_FP16, L.PREC_FP32), eps_prime=1e6,
                                             tile_rows=256)
    assert rco == 0
    low = np.tril(np.ones_like(gmap, bool), -1)
    assert np.array_equal(gmap[low], omap[low])
    assert np.any(gmap[low] == L.PREC_FP32)     # eps' = 1e6 alone would admit FP16 everywhere
    # the unscaled matrix keeps FP16 for every tile
    T2 = to_dev_cm(A / np.float32(1e10))
    rc2, fi2, mp2, _ = S.factor(handle, T2, fo, want_map=True)
    assert rc2 == 0 and np.all(mp2.cpu().numpy().T[low] == L.PREC_FP16)


# ---------------------------------------------------------------- matgen spectra
@pytest.mark.parametrize("dist", [L.DIST_RANDOM_LOG_UNIFORM, L.DIST_CLUSTERED, L.DIST_ARITHMETIC,
                                  L.DIST_GEOMETRIC, L.DIST_LOGSPACED])
def test_matgen_every_dist(handle, oracle, dist):
    """Every DistType ordinal (kernels_macros.cuh:24-29) on the GPU: the same
    spectrum as the oracle (bitwise: both take the seeded Philox streams) and
    the same matrix to FP32 rounding; cond(A) as requested."""
    n, cond = 640, 1e4
    say("matgen dist", dist)
    A, lam = S.matgen_householder(handle, n, cond=cond, lam_max=1.0, seed=44, dist=dist)
    Ao, lam_o = oracle.householder_spd(n, cond, 1.0, 44, dist=dist)
    assert np.allclose(lam.cpu().numpy(), lam_o, rtol=1e-14, atol=0)
    Ag = A.cpu().numpy().T
    assert np.array_equal(Ag, Ag.T)
    assert np.abs(Ag - Ao).max() <= 1e-6 * np.abs(Ao).max()
    ev = np.linalg.eigvalsh(Ag.astype(np.float64))
    assert np.allclose(np.sort(ev)[-5:], np.sort(lam_o)[-5:], rtol=1e-5)
    if dist != L.DIST_RANDOM_LOG_UNIFORM:
        assert np.isclose(lam_o.max() / lam_o.min(), cond, rtol=1e-9)
