"""CPU tests that pin the oracle (oracle/oracle.c) before it is trusted.

Each check is chosen so that a plausible mistake in the restatement (wrong
tie rule, wrong subnormal threshold, transposed operand, dropped term) fails.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


# ---------------------------------------------------------------- 8-bit formats
# lo_float.h:789-847 (e4m3fn) and :962-1025 (e5m2) special values and limits
E4M3_TABLE = [
    (0.0, 0x00), (-0.0, 0x80), (448.0, 0x7E), (-448.0, 0xFE), (464.0, 0x7E),
    (464.0001, 0x7F), (1e9, 0x7F), (-1e9, 0xFF), (math.inf, 0x7F), (-math.inf, 0xFF),
    (math.nan, 0x80), (2.0 ** -9, 0x01), (2.0 ** -10, 0x00), (3 * 2.0 ** -11, 0x01),
    (3 * 2.0 ** -10, 0x02), (2.0 ** -6, 0x08), (1.0, 0x38), (0.125, 0x20), (1.0625, 0x38), (1.1875, 0x3A),
    (-3.25, 0xC5), (2.0 ** -10 * 1.001, 0x01), (-(2.0 ** -12), 0x80),
]
E5M2_TABLE = [
    (0.0, 0x00), (-0.0, 0x80), (57344.0, 0x7B), (61440.0, 0x7C), (61439.0, 0x7B),
    (math.inf, 0x7C), (-math.inf, 0xFC), (math.nan, 0x7E), (2.0 ** -16, 0x01),
    (2.0 ** -17, 0x00), (2.0 ** -14, 0x04), (1.0, 0x3C), (1.125, 0x3C), (1.375, 0x3E),
]


@pytest.mark.parametrize("x,code", E4M3_TABLE)
def test_e4m3_table(x, code):
    assert O.lib().orc_f32_to_e4m3(x) == code


@pytest.mark.parametrize("x,code", E5M2_TABLE)
def test_e5m2_table(x, code):
    assert O.lib().orc_f32_to_e5m2(x) == code


@pytest.mark.parametrize("fmt", [O.E4M3, O.E5M2])
def test_decode_roundtrip(fmt):
    """Every finite code decodes and re-encodes to itself (lo_float's -0 = 0x80
    is its NaN code, so 0x80 is excluded for E4M3)."""
    for v in range(256):
        x = O.lib().orc_e4m3_to_f32(v) if fmt == O.E4M3 else O.lib().orc_e5m2_to_f32(v)
        if not math.isfinite(x):
            continue
        enc = O.lib().orc_f32_to_e4m3(x) if fmt == O.E4M3 else O.lib().orc_f32_to_e5m2(x)
        if x == 0.0:
            assert enc in (0x00, 0x80)
        else:
            assert enc == v, (v, x, enc)


@pytest.mark.parametrize("fmt", [O.E4M3, O.E5M2])
def test_rne_against_closed_form_sample(fmt):
    """Independent frexp/ldexp rounding on 2^22 stratified inputs (the full 2^32
    sweep is test_rne_against_closed_form_exhaustive)."""
    rng = np.random.default_rng(fmt)
    bits = rng.integers(0, 2 ** 32, size=1 << 22, dtype=np.uint64)
    lo = 0
    # the C self-check over a contiguous window plus random windows
    assert O.selfcheck_f8(fmt, 0x3F000000, 0x3F000000 + (1 << 22)) == 0
    assert O.selfcheck_f8(fmt, 0x00000000, 1 << 22) == 0     # zero and FP32 subnormals
    for b in bits[:16]:
        start = int(b) & ~0xFFFF
        assert O.selfcheck_f8(fmt, start, start + 0x10000) == 0
    del lo


@pytest.mark.slow
@pytest.mark.parametrize("fmt", [O.E4M3, O.E5M2])
def test_rne_against_closed_form_exhaustive(fmt):
    """All 2^32 float bit patterns (about 15 s per format on 8 threads)."""
    assert O.selfcheck_f8(fmt) == 0


def test_bf16_f16_against_numpy():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(20000) * np.exp(rng.uniform(-30, 30, 20000)),
                        [0.0, -0.0, 65504.0, 65520.0, 65519.99, 6e-8, 3e-8, 2.0 ** -24,
                         2.0 ** -25, 2.0 ** -25 * 1.0001]]).astype(np.float32)
    f16 = np.array([O.lib().orc_f32_to_f16(float(v)) for v in x], np.uint16)
    with np.errstate(over="ignore"):
        expect = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(f16, expect)
    # bfloat16 RNE closed form: round(x / ulp) * ulp on the 8-bit mantissa grid
    for v in x[:2000]:
        b = O.lib().orc_f32_to_bf16(float(v))
        back = O.lib().orc_bf16_to_f32(b)
        if v == 0:
            assert back == 0
            continue
        e = math.floor(math.log2(abs(float(v))))
        ulp = 2.0 ** (e - 7)
        q = abs(float(v)) / ulp
        r = math.floor(q)
        if q - r > 0.5 or (q - r == 0.5 and r % 2 == 1):
            r += 1
        assert abs(back) == r * ulp, (v, back)


def test_scale_exp_closed_form():
    for fmt, mx in [(O.E4M3, 448.0), (O.E5M2, 57344.0), (O.MXE4M3, 448.0)]:
        for am in [1.0, 448.0, 449.0, 0.001, 3e-39, 1e30, 57344.0, 1.75, 1.7500001]:
            e = O.scale_exp(am, fmt)
            assert am / 2.0 ** e <= mx
            if e > -127:
                assert am / 2.0 ** (e - 1) > mx      # smallest such power of two
    assert O.scale_exp(0.0, O.E4M3) == 0
    assert O.scale_exp(5.0, O.FP32) == 0


def test_quantize_panel_values():
    rng = np.random.default_rng(1)
    P = (rng.standard_normal((300, 64)) * 10).astype(np.float32)
    for prec in [O.E4M3, O.E5M2, O.MXE4M3, O.MXE5M2, O.BF16, O.FP16, O.FP32]:
        pay, sc, deq = O.quantize_panel(P, [prec] * 3, 128)
        rel = np.abs(deq - P) / np.maximum(np.abs(P), 1e-30)
        u = O.unit_roundoff(prec)
        big = np.abs(P) > np.abs(P).max() * 2.0 ** -5    # away from the subnormal range
        assert rel[big].max() <= u * (1 + 1e-6), O.unit_roundoff(prec)
        if prec in (O.MXE4M3, O.MXE5M2):
            # every 32-block of a row has its own power-of-two scale
            e = sc.astype(int) - 127
            blk = np.abs(P).reshape(300, 2, 32).max(axis=2)
            mx = 448.0 if prec == O.MXE4M3 else 57344.0
            assert np.all(blk / 2.0 ** e <= mx)


# ---------------------------------------------------------------- RNG and generators
def test_philox_known_answers():
    """Random123 kat_vectors for philox4x32-10."""
    assert list(O.philox([0, 0, 0, 0], [0, 0])) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert list(O.philox([0xffffffff] * 4, [0xffffffff] * 2)) == [0x408f276d, 0x41c83b0e,
                                                                  0xa20bc7c6, 0x6d5451fd]
    assert list(O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                         [0xa4093822, 0x299f31d0])) == [0xd16cfe09, 0x94fdcceb, 0x5001e420,
                                                        0x24126ea1]


def test_normal_moments():
    z = np.array([O.normal(5, 0, i) for i in range(40000)])
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02
    u = np.array([O.uniform(5, 1, i) for i in range(40000)])
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.01


@pytest.mark.parametrize("dist", [0, 1, 2, 3, 4])
def test_householder_spectrum(dist):
    """H Lambda H^T has exactly the spectrum Lambda (H orthogonal)."""
    n = 300
    cond = 1e4
    A, lam = O.householder_spd(n, cond, 1.0, 17, dist=dist)
    assert np.array_equal(A, A.T)
    Hl = np.zeros((n, n))
    W, Pm, g = O.hh_setup(n, 4, 17, lam)
    for i in range(n):
        for j in range(n):
            Hl[i, j] = O.lib().orc_hh_entry(i, j, lam.ctypes.data, W.ctypes.data, Pm.ctypes.data,
                                            g.ctypes.data)
    ev = np.linalg.eigvalsh(Hl)
    assert np.allclose(np.sort(ev), np.sort(lam), rtol=1e-9, atol=1e-12 * lam.max())
    if dist == 4:
        assert np.isclose(lam.max() / lam.min(), cond)


def test_spectrum_cluster():
    lam = O.spectrum(4, 1000, 1e4, 1.0, 64, 1e-8, 1.1e-8, 0)
    assert np.all(lam[:64] >= 1e-8 * (1 - 1e-12)) and np.all(lam[:64] <= 1.1e-8 * (1 + 1e-12))
    assert np.isclose(lam[64:].min(), 1e-4) and np.isclose(lam.max(), 1.0)


def test_rbf_closed_forms():
    from scipy.spatial.distance import cdist
    X = O.points_gaussian(200, 16, 3)
    K = O.rbf_matrix(X, 4.0, 1.0, 1e-2)
    assert np.allclose(np.diag(K), 1.0 + 1e-2)
    assert np.array_equal(K, K.T)
    D2 = cdist(X.astype(np.float64), X.astype(np.float64), "sqeuclidean")
    Kr = np.exp(-D2 / (2 * 16.0)) + 1e-2 * np.eye(200)
    assert np.abs(K - Kr).max() < 1e-6
    ls = np.linspace(0.5, 2.0, 16)
    Ka = O.rbf_matrix(X, ls, 1.0, 0.0)
    D2a = cdist(X / ls, X / ls, "sqeuclidean")
    assert np.abs(Ka - np.exp(-0.5 * D2a)).max() < 1e-6
    Xs = O.points_sparse_image(1000, 50, 0.8, 9)
    frac0 = np.mean(Xs == 0)
    assert abs(frac0 - 0.8) < 0.02 and Xs.max() <= 1.0 and Xs.min() >= 0.0


# ---------------------------------------------------------------- factor + PCG
def test_micro_cholesky_matches_reference_build():
    """orc_potrf_f32_micro vs the reference's own micro_cholesky
    (micro_chol.hpp:14-44 compiled from the reference tree into oracle/_ref)."""
    r = O.ref_lib()
    if r is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(0)
    for n in [1, 2, 17, 64, 128]:
        M = rng.standard_normal((n, n))
        A = (M @ M.T + n * np.eye(n)).astype(np.float32)
        L1, info = O.potrf_f32_micro(A)
        L2 = np.array(A, order="F", copy=True)
        r.ref_micro_cholesky(L2.ctypes.data, n, n)
        assert info == 0
        assert np.array_equal(L1.view(np.uint32), L2.view(np.uint32))


def test_potrf_closed_forms():
    L, info = O.potrf_f64(np.array([[4.0, 2.0], [2.0, 3.0]]))
    assert info == 0 and np.allclose(L, [[2, 0], [1, math.sqrt(2)]])
    A = np.array([[25.0, 15, -5], [15, 18, 0], [-5, 0, 11]])
    L, info = O.potrf_f64(A)
    assert np.allclose(L, [[5, 0, 0], [3, 3, 0], [-1, 1, 3]])
    _, info = O.potrf_f64(np.array([[1.0, 2.0], [2.0, 1.0]]))
    assert info == 2
    Lm, info = O.potrf_f32_micro(np.array([[4.0, 2.0], [2.0, -3.0]], np.float32))
    assert info == 2


def test_potrf_f64_vs_numpy():
    rng = np.random.default_rng(2)
    M = rng.standard_normal((200, 200))
    A = M @ M.T + 200 * np.eye(200)
    L, info = O.potrf_f64(A)
    assert info == 0 and np.allclose(L, np.linalg.cholesky(A), rtol=1e-12, atol=1e-12)


def test_syrk_ones_kat():
    """syrk_gemm.cuh:252-279: ones(256 x 256) A A^T = 256 everywhere (exact)."""
    Cm = np.zeros((256, 256), np.float32, order="F")
    O.syrk_lower(Cm, np.ones((256, 256), np.float32))
    assert np.all(Cm[np.tril_indices(256)] == -256.0)
    assert np.all(np.triu(Cm, 1) == 0.0)


def test_factor_fp32_matches_numpy():
    A, _ = O.householder_spd(512, 1e4, 1.0, 3)
    L, rc, mp, ud, fi = O.factor_mixed(A, nb=128, update_prec=O.FP32)
    assert rc == 0
    Ln = np.linalg.cholesky(A.astype(np.float64))
    assert np.linalg.norm(np.tril(L) - Ln) / np.linalg.norm(Ln) < 1e-4
    # the Schur-diagonal tracker (update_diag) ends at a_ii minus the squares of
    # row i over the panels before i's own block: the diagonal handed to POTRF
    c0 = (np.arange(512) // 128) * 128
    expect = np.array([A[i, i] - np.sum(Ln[i, :c0[i]] ** 2) for i in range(512)])
    assert np.allclose(ud, expect, rtol=1e-3, atol=1e-5)
    assert np.isclose(fi.min_pivot, (np.diag(Ln) ** 2).min(), rtol=1e-4)
    # flop accounting: SYRK flops sum_j s(s+1)b over the steps (lower incl. diag)
    s = np.arange(512 - 128, 0, -128)
    assert np.isclose(sum(fi.flops), float(np.sum(s * (s + 1) * 128)), rtol=1e-12)


def test_factor_mixed_rule_and_parity():
    A, _ = O.householder_spd(1024, 1e6, 1e6, 11)
    L, rc, mp, ud, fi = O.factor_mixed(A, nb=128, update_prec=-1, ladder=(O.E4M3, O.FP32),
                                       eps_prime=0.25)
    assert rc == 0
    assert fi.flops[O.E4M3] > 0 and fi.flops[O.FP32] > 0
    # replaying the schedule reproduces the factor exactly
    L2, rc2, mp2, _, _ = O.factor_mixed(A, nb=128, update_prec=-1, ladder=(O.E4M3, O.FP32),
                                        eps_prime=0.25, map_in=mp)
    assert rc2 == 0 and np.array_equal(mp, mp2) and np.array_equal(L, L2)
    # the rule itself
    lad = np.array([O.E4M3, O.FP32], np.int32)
    assert O.lib().orc_rule_minpivot(lad.ctypes.data, 2, 1.0, 16.0, 1.0) == O.E4M3
    assert O.lib().orc_rule_minpivot(lad.ctypes.data, 2, 1.0, 16.1, 1.0) == O.FP32


def test_pcg_exact_and_iterations():
    n = 600
    A, _ = O.householder_spd(n, 1e6, 1e6, 5)
    b = np.random.default_rng(1).standard_normal(n)
    L, rc, _, _, _ = O.factor_mixed(A, nb=128, update_prec=O.FP32)
    x, it, rr, kap = O.pcg(A, np.tril(L), b, tol=1e-10)
    assert 0 < it <= 6 and rr <= 1e-10
    xe = np.linalg.solve(A.astype(np.float64), b)
    assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-4
    assert 1.0 <= kap < 1.5
    # identity preconditioner (L = I): plain CG does not reach 1e-6 in 100 iterations
    x2, it2, rr2, kap2 = O.pcg(A, np.eye(n, dtype=np.float32), b, tol=1e-6, maxit=100)
    assert it2 == -1 and rr2 > 1e-6
    # reference stopping rule sqrt(r.z)/||b||
    _, it3, _, _ = O.pcg(A, np.tril(L), b, tol=1e-10, stop_rule=1)
    assert it3 > 0


def test_lanczos_cond_matches_tridiagonal_eigenvalues():
    """The Lanczos tridiagonal of CG (diagonal 1/a_k + b_{k-1}/a_{k-1},
    off-diagonal sqrt(b_k)/a_k): Sturm bisection's extreme-eigenvalue ratio
    equals numpy's on the same matrix, and run to convergence on a small SPD
    matrix it gives cond(A) to 1e-3."""
    n = 24
    lam = np.geomspace(1.0, 50.0, n)
    A = np.diag(lam).astype(np.float32)
    b = np.random.default_rng(8).standard_normal(n)
    _, it, _, _ = O.pcg(A, np.eye(n, dtype=np.float32), b, tol=1e-14, maxit=n, x0_mode=1)
    al = np.zeros(n + 1)
    be = np.zeros(n + 1)
    o = O.PcgOpts(1e-14, n, 0, 1)
    x = np.zeros(n)
    rr = O.C.c_double(0)
    A32 = np.asfortranarray(A)
    I32 = np.asfortranarray(np.eye(n, dtype=np.float32))
    k = O.lib().orc_pcg(O._p(A32), n, O._p(I32), n, n, O._p(b), O._p(x), O.C.byref(o), O.C.byref(rr),
                        O._p(al), O._p(be))
    k = k if k > 0 else n
    for m in (4, 10, k):
        d = np.array([1 / al[i] + (be[i - 1] / al[i - 1] if i else 0.0) for i in range(m)])
        e = np.array([np.sqrt(be[i]) / al[i] for i in range(m - 1)])
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        ev = np.linalg.eigvalsh(T)
        assert abs(O.lib().orc_lanczos_cond(O._p(al), O._p(be), m) - ev[-1] / ev[0]) <= 1e-9 * ev[-1] / ev[0]
    assert abs(O.lib().orc_lanczos_cond(O._p(al), O._p(be), k) - 50.0) <= 1e-3 * 50.0


def _diag_spd(n, kappa, seed):
    """Diagonal SPD with n distinct eigenvalues (exact in FP32)."""
    lam = np.geomspace(1.0, kappa, n).astype(np.float32)
    np.random.default_rng(seed).shuffle(lam)
    return np.diag(lam).astype(np.float32)


def test_block_pcg_one_column_is_pcg():
    """s = 1: alpha = z.r / p.q and beta = z'.r' / z.r, the single-vector loop."""
    n = 400
    A, _ = O.householder_spd(n, 1e4, 1e4, 7)
    L, rc, _, _, _ = O.factor_mixed(A, nb=64, update_prec=O.FP32)
    b = np.random.default_rng(2).standard_normal(n)
    x, it, rr, _ = O.pcg(A, np.tril(L), b, tol=1e-10)
    X, itb, rel = O.block_pcg(A, np.tril(L), b[:, None], tol=1e-10)
    assert itb == it and rel[0] <= 1e-10
    assert np.abs(X[:, 0] - x).max() <= 1e-12 * np.abs(x).max()


@pytest.mark.parametrize("n,s", [(32, 4), (32, 8), (64, 8), (64, 16), (32, 32)])
def test_block_pcg_finite_termination(n, s):
    """Block CG with s columns spans s new directions per step: on n distinct
    eigenvalues it terminates within ceil(n/s) steps (one step for s = n),
    where single-vector CG on the same matrix needs about 20 or more.  A
    transposed gram or a wrong alpha/beta order loses this property."""
    A = _diag_spd(n, 4.0, n + s)
    B = np.random.default_rng(s).standard_normal((n, s))
    eye = np.eye(n, dtype=np.float32)
    X, it, rel = O.block_pcg(A, eye, B, tol=1e-10, maxit=200, x0_mode=1)
    assert 0 < it <= -(-n // s)
    assert rel.max() <= 1e-12
    np.testing.assert_allclose(X, B / np.diag(A).astype(np.float64)[:, None], rtol=1e-11, atol=0)


def test_block_pcg_matches_fp64_solve():
    n, s = 500, 6
    A, _ = O.householder_spd(n, 1e6, 1e6, 11)
    L, rc, _, _, _ = O.factor_mixed(A, nb=128, update_prec=-1, ladder=(O.E4M3, O.FP32),
                                    eps_prime=0.5)
    assert rc == 0
    B = np.random.default_rng(3).standard_normal((n, s))
    X, it, rel = O.block_pcg(A, np.tril(L), B, tol=1e-10)
    assert it > 0 and rel.max() <= 1e-10
    Xe = np.linalg.solve(A.astype(np.float64), B)
    assert np.abs(X - Xe).max() / np.abs(Xe).max() < 1e-3
    # no slower than the independent PCGs it replaces
    its = [O.pcg(A, np.tril(L), B[:, j], tol=1e-10)[1] for j in range(s)]
    assert it <= max(its)


def test_block_pcg_dependent_columns_break_down():
    """Two equal right-hand sides make P^T A P singular: the block
    recurrence reports breakdown (-2) instead of dividing by a zero pivot."""
    n = 200
    A, _ = O.householder_spd(n, 1e2, 1e2, 17)
    B = np.random.default_rng(5).standard_normal((n, 3))
    B[:, 2] = B[:, 0]
    X, it, rel = O.block_pcg(A, np.eye(n, dtype=np.float32), B, tol=1e-10, x0_mode=1)
    assert it == -2 and np.isfinite(X).all()


def test_refine_residual_polynomial_and_fixed_point():
    """x0 = M^-1 b and x += M^-1 r give r_k = (I - A M^-1)^(k+1) b in closed
    form (a sign error, a transposed solve or a skipped x0 breaks it); the
    rate is the spectral radius of I - M^-1 A; an exact factor needs one
    correction; the fixed point is the FP64 solution of the FP32 matrix."""
    n = 300
    A, _ = O.householder_spd(n, 1e3, 1e3, 13)
    A64 = A.astype(np.float64)
    b = np.random.default_rng(4).standard_normal(n)
    xe = np.linalg.solve(A64, b)
    # exact factor rounded to FP32: one correction
    Lx = np.linalg.cholesky(A64).astype(np.float32)
    x, it, rr, hist = O.refine(A, Lx, b, tol=1e-12, maxit=10)
    assert it == 1 and rr <= 1e-12
    # an E4M3/FP32 switching factor
    L, rc, _, _, _ = O.factor_mixed(A, nb=64, update_prec=-1, ladder=(O.E4M3, O.FP32),
                                    eps_prime=0.5)
    assert rc == 0
    L = np.tril(L).astype(np.float64)
    E = np.eye(n) - A64 @ np.linalg.inv(L @ L.T)          # residual propagator
    rho = np.abs(np.linalg.eigvals(E)).max()
    assert rho < 0.5
    x, it, rr, hist = O.refine(A, L.astype(np.float32), b, tol=1e-13, maxit=40)
    assert it > 1 and rr <= 1e-13
    r = E @ b
    for k in range(len(hist)):
        want = np.linalg.norm(r) / np.linalg.norm(b)
        if want > 1e-11:                                 # above the rounding floor
            assert abs(hist[k] - want) <= 1e-6 * want + 1e-13
        r = E @ r
    rate = (hist[it - 1] / hist[1]) ** (1.0 / (it - 2)) if it > 2 else hist[1] / hist[0]
    assert rate <= 1.05 * rho
    assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-11
    # maxit reached: -1 and the last residual reported
    _, it2, rr2, hist2 = O.refine(A, L.astype(np.float32), b, tol=1e-16, maxit=2)
    assert it2 == -1 and len(hist2) == 3 and rr2 == hist2[-1]
    # an all-E4M3 factor has rho(I - M^-1 A) > 1 here: refinement diverges
    # where PCG (whose rate depends on the spread of M^-1 A) still converges
    L8, rc, _, _, _ = O.factor_mixed(A, nb=64, update_prec=O.E4M3, eps_prime=1.0)
    L8 = np.tril(L8)
    E8 = np.eye(n) - A64 @ np.linalg.inv(L8.astype(np.float64) @ L8.T.astype(np.float64))
    assert np.abs(np.linalg.eigvals(E8)).max() > 1.0
    assert O.refine(A, L8, b, tol=1e-10, maxit=30)[1] == -1
    assert O.pcg(A, L8, b, tol=1e-10)[1] > 0


def _rules_case(c=8.0):
    """A = L L^T for L = [[2 I, 0], [B, c I]] (256 x 256 blocks, B = +-1/2):
    every step of the factor is exact in FP32.  Panel 0's tile (256 rows)
    has row sums of squares 64, Frobenius^2 256 * 64, L_11 = 2 I
    (Frobenius^2 1024, pivots^2 4), and the Schur tracker below it is c^2
    against the original diagonal 64 + c^2."""
    m = 256
    B = np.where(np.random.default_rng(12).random((m, m)) < 0.5, -0.5, 0.5)
    Lx = np.zeros((2 * m, 2 * m))
    Lx[:m, :m] = 2 * np.eye(m)
    Lx[m:, :m] = B
    Lx[m:, m:] = c * np.eye(m)
    return (Lx @ Lx.T).astype(np.float32), Lx


@pytest.mark.parametrize("rule,eps_hi,eps_lo", [
    (O.RULE_MINPIVOT, 1.05, 0.95),       # u * 64 <= eps * 4
    (O.RULE_FROBENIUS, 1.05, 0.95),      # u * 16384 <= eps * 1024
    (O.RULE_SCHUR, 0.04, 0.03),          # u * 64 / 128 <= eps
])
def test_switching_rules_closed_form(rule, eps_hi, eps_lo):
    """Each rule's threshold lands where its closed form says (u_E4M3 = 1/16):
    E4M3 just above it, FP32 just below; the factor is exact either way."""
    A, Lx = _rules_case()
    for eps, want in ((eps_hi, O.E4M3), (eps_lo, O.FP32)):
        L, rc, mp, ud, _ = O.factor_mixed(A, nb=256, update_prec=-1, ladder=(O.E4M3, O.FP32),
                                          eps_prime=eps, rule=rule)
        assert rc == 0 and mp[1, 0] == want, (rule, eps, mp[1, 0])
        if want == O.FP32:
            np.testing.assert_array_equal(np.tril(L), Lx.astype(np.float32))
        np.testing.assert_array_equal(ud[256:], np.float32(64.0))


def test_factor_breakdown_info():
    A = np.eye(256, dtype=np.float32)
    A[100, 100] = -1
    L, rc, _, _, _ = O.factor_mixed(A, nb=64, update_prec=O.FP32)
    assert rc == 101


# ---------------------------------------------------------------- GP loss and gradient
def _gp_case(n=96, d=3, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, d)).astype(np.float32)
    return X, np.array([0.7, 1.3, 1.9][:d]), 1.3, 0.05, rng.standard_normal(n)


@pytest.mark.parametrize("ard", [True, False])
def test_gp_gradient_matches_finite_differences(ard):
    """The exact FP64 gradient formula (dA/dls = K dx^2/ls^3, dA/dvar = K/var,
    dA/dnoise = I, dloss/db = A^-1 b) against central differences of the
    exact loss: a dropped term, a wrong power of ls or a missing 1/2 fails."""
    X, ls, var, noise, b = _gp_case()
    if not ard:
        ls = np.array([1.1])
    loss, gb, gls, gvar, gno = O.gp_loss_grad_f64(X, ls, var, noise, b)

    def f(ls_, var_, no_, b_):
        return O.gp_loss_grad_f64(X, ls_, var_, no_, b_)[0]

    h = 1e-5
    for c in range(ls.size):
        e = np.zeros(ls.size)
        e[c] = h
        fd = (f(ls + e, var, noise, b) - f(ls - e, var, noise, b)) / (2 * h)
        assert abs(fd - gls[c]) <= 1e-6 * max(1.0, abs(fd)), (c, fd, gls[c])
    fd = (f(ls, var + h, noise, b) - f(ls, var - h, noise, b)) / (2 * h)
    assert abs(fd - gvar) <= 1e-6 * max(1.0, abs(fd))
    hn = 1e-7
    fd = (f(ls, var, noise + hn, b) - f(ls, var, noise - hn, b)) / (2 * hn)
    assert abs(fd - gno) <= 1e-5 * max(1.0, abs(fd))
    for i in (0, 17, 95):
        e = np.zeros(b.size)
        e[i] = h
        fd = (f(ls, var, noise, b + e) - f(ls, var, noise, b - e)) / (2 * h)
        assert abs(fd - gb[i]) <= 1e-6 * max(1.0, abs(fd))


def test_gp_unit_probes_give_exact_trace():
    """Hutchinson with the unit vectors e_1..e_n (weight 1) is the exact trace."""
    X, ls, var, noise, b = _gp_case(n=64)
    ex = O.gp_loss_grad_f64(X, ls, var, noise, b)
    un = O.gp_loss_grad_f64(X, ls, var, noise, b, Z=np.eye(64), unit=True)
    for a, c in zip(ex[2:], un[2:]):
        assert np.allclose(a, c, rtol=1e-12, atol=1e-12)


def test_probes_are_rademacher_and_seeded():
    Z = O.probes(4096, 8, 99)
    assert set(np.unique(Z)) == {-1.0, 1.0}
    assert abs(Z.mean()) < 0.02                     # balanced
    assert np.array_equal(Z, O.probes(4096, 8, 99))
    assert not np.array_equal(Z, O.probes(4096, 8, 100))
    # definition: sign of the Philox uniform of stream 3 at index p*n + i
    for p, i in [(0, 0), (3, 17), (7, 4095)]:
        u = O.uniform(99, 3, p * 4096 + i)
        assert Z[p, i] == (-1.0 if u < 0.5 else 1.0)


def test_rule_skips_fp16_out_of_range():
    """FP16 (stored unscaled) is taken only for a tile whose largest entry is
    finite in FP16 (<= 65504); otherwise the next format of the ladder.  With
    eps' = 1e6 the pivot test admits FP16 for every tile: unscaled, every
    tile is FP16; scaled by 1e10, the tiles with entries beyond FP16's range
    take FP32 and the factor completes without an overflow (the quantiser
    would report a non-finite FP16 value)."""
    lad = np.array([O.FP16, O.FP32], np.int32)
    f = O.lib().orc_rule_minpivot_amax
    assert f(lad.ctypes.data, 2, 1e6, 1.0, 1.0, 65504.0) == O.FP16
    assert f(lad.ctypes.data, 2, 1e6, 1.0, 1.0, 65505.0) == O.FP32
    assert f(lad.ctypes.data, 2, 1e6, 1.0, 1.0, float("nan")) == O.FP32
    A, _ = O.householder_spd(512, 1e3, 1e3, 21)
    low = np.tril(np.ones((4, 4), bool), -1)
    for scale, all16 in ((1e10, False), (1.0, True)):
        L, rc, mp, _, fi = O.factor_mixed(A * np.float32(scale), nb=128, update_prec=-1,
                                          ladder=(O.FP16, O.FP32), eps_prime=1e6)
        assert rc == 0
        assert np.all(np.isfinite(np.tril(L)))
        assert np.all(mp[low] == O.FP16) == all16
        assert np.all((mp[low] == O.FP16) | (mp[low] == O.FP32))
