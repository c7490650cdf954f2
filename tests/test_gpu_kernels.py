"""GPU parity tests: libmpchol's kernels (through the C ABI) against the oracle.

Run on an MI355X: python -u -m pytest tests -m gpu -x -v --timeout 120 --timeout-method thread
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from cholesky_gpu_amd import _lib as L  # noqa: E402
from cholesky_gpu_amd import solver as S  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def to_dev_cm(A):
    """numpy matrix -> device tensor holding it column-major (tensor = A^T)."""
    return torch.from_numpy(np.ascontiguousarray(np.asarray(A, np.float32).T)).cuda()


def from_dev_cm(T):
    return T.cpu().numpy().T


def spd(n, seed, cond=1e3):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.logspace(0, np.log10(cond), n)
    A = (Q * lam) @ Q.T
    return ((A + A.T) / 2).astype(np.float32)


# ---------------------------------------------------------------- conversions
@pytest.mark.parametrize("fmt", [L.PREC_E4M3, L.PREC_E5M2])
def test_convert_f8_exhaustive(handle, oracle, fmt):
    """Hardware OCP conversion == lo_float restatement on every float whose
    magnitude is within the quantiser's domain (|x| <= max finite)."""
    maxabs = 448.0 if fmt == L.PREC_E4M3 else 57344.0
    chunk = 1 << 28
    bad = other = 0
    for c in range(16):
        lo = c * chunk
        say("convert", L.PREC_NAMES[fmt], "chunk", c)
        bits = torch.arange(lo, lo + chunk, dtype=torch.int64, device="cuda").to(torch.int32)
        y = S.convert_f8(handle, bits.view(torch.float32), fmt).cpu().numpy()
        del bits
        b, o = oracle.sweep(fmt, lo, lo + chunk, y, maxabs)
        bad += b
        other += o
    print("out-of-domain encodings that differ (NaN/overflow):", other)
    assert bad == 0


# ---------------------------------------------------------------- POTRF
@pytest.mark.parametrize("n", [256, 200, 64, 33])
def test_potrf_block(handle, oracle, n):
    say("potrf_block n =", n)
    A = spd(n, n)
    T = to_dev_cm(A)
    info = S.potrf_block(handle, T)
    assert info == 0
    Lg = np.tril(from_dev_cm(T))
    Lo, info_o = oracle.potrf_f32_micro(A)
    assert info_o == 0
    Lo = np.tril(Lo)
    rel = np.linalg.norm(Lg - Lo) / np.linalg.norm(Lo)
    assert rel < 2e-5, rel
    bwd = np.linalg.norm(Lg.astype(np.float64) @ Lg.T.astype(np.float64) - A) / np.linalg.norm(A)
    assert bwd < 4 * n * 6e-8, bwd
    # strict upper triangle untouched
    assert np.array_equal(np.triu(from_dev_cm(T), 1), np.triu(A, 1))


def test_potrf_block_indefinite(handle):
    say("potrf_block indefinite")
    A = np.eye(64, dtype=np.float32)
    A[40, 40] = -1.0
    info = S.potrf_block(handle, to_dev_cm(A))
    assert info == 41


# ---------------------------------------------------------------- quantiser
def f8_rowmajor(flat, r0, r1, k, row_stride):
    """Rows r0..r1-1 (k codes each) of an 8-bit tile stored in the
    matrix-core operand order: per 32-row block and 64-wide chunk c, lane
    l = n + 32 h holds bytes 0-15 = row n, K [64c + 16h, +16) and bytes 16-31 =
    row n, K [64c + 32 + 16h, +16), the block's chunks from its first row's
    offset."""
    rows = np.arange(r0, r1)[:, None]
    kk = np.arange(k)[None, :]
    c, kp = kk >> 6, kk & 63
    lane = (rows & 31) + 32 * ((kp >> 4) & 1)
    off = (rows & ~31) * row_stride + c * 2048 + lane * 32 + (kp >> 5) * 16 + (kp & 15)
    return flat[off]


ALL_FORMATS = [L.PREC_E4M3, L.PREC_E5M2, L.PREC_MXE4M3, L.PREC_MXE5M2, L.PREC_BF16,
               L.PREC_FP16, L.PREC_FP32]


@pytest.mark.parametrize("m,k,tile_rows", [(7 * 128 + 37, 256, 128), (512, 128, 256), (300, 64, 128)])
def test_quantize_panel_bitexact(handle, oracle, m, k, tile_rows):
    say("quantize", m, k, tile_rows)
    rng = np.random.default_rng(m + k)
    P = (rng.standard_normal((m, k)) * np.exp(rng.uniform(-8, 8, (m, 1)))).astype(np.float32)
    P[3, 5] = 0.0
    P[4, 7] = -0.0
    P[5, :] = 1e-30   # tiny row: subnormal outputs
    nt = (m + tile_rows - 1) // tile_rows
    precs = [ALL_FORMATS[t % len(ALL_FORMATS)] for t in range(nt)]
    pay_g, sc_g, _ = S.quantize_panel(handle, to_dev_cm(P), precs, tile_rows)
    pay_o, sc_o, _ = oracle.quantize_panel(P, precs, tile_rows)
    pay_g = pay_g.cpu().numpy()
    sc_g = sc_g.cpu().numpy()
    flat = pay_g.reshape(-1)
    for t in range(nt):
        r0, r1 = t * tile_rows, min(m, (t + 1) * tile_rows)
        nbytes = k * {L.PREC_FP32: 4, L.PREC_BF16: 2, L.PREC_FP16: 2}.get(precs[t], 1)
        got = pay_g[r0:r1, :nbytes]
        if nbytes == k:                # 8-bit tile: matrix-core operand order
            got = f8_rowmajor(flat, r0, r1, k, pay_g.shape[1])
            # rows of the last 32-row block beyond m hold zeros
            r32 = min(((r1 + 31) // 32) * 32, (t + 1) * tile_rows)
            if r1 == m and r32 > m:
                assert not f8_rowmajor(flat, m, r32, k, pay_g.shape[1]).any()
        assert np.array_equal(got, pay_o[r0:r1, :nbytes]), L.PREC_NAMES[precs[t]]
        assert np.array_equal(sc_g[r0:r1], sc_o[r0:r1]), L.PREC_NAMES[precs[t]]


# ---------------------------------------------------------------- SYRK
SYRK_CASES = [
    (L.PREC_E4M3, L.PREC_E4M3), (L.PREC_MXE4M3, L.PREC_MXE4M3), (L.PREC_E5M2, L.PREC_E5M2),
    (L.PREC_MXE5M2, L.PREC_E4M3), (L.PREC_BF16, L.PREC_BF16), (L.PREC_BF16, L.PREC_E4M3),
    (L.PREC_FP16, L.PREC_FP16), (L.PREC_FP32, L.PREC_FP32), (L.PREC_FP32, L.PREC_MXE4M3),
    (L.PREC_FP32, L.PREC_BF16),
]


def flush_fp8_subnormals(deq, pay, precs, tr):
    """deq with every FP8 element whose code is subnormal replaced by zero."""
    out = deq.copy()
    for t, p in enumerate(precs):
        r0, r1 = t * tr, min(deq.shape[0], (t + 1) * tr)
        if p in (L.PREC_E4M3, L.PREC_MXE4M3):
            codes = pay[r0:r1, :deq.shape[1]]
            sub = ((codes & 0x78) == 0) & ((codes & 0x07) != 0)
        elif p in (L.PREC_E5M2, L.PREC_MXE5M2):
            codes = pay[r0:r1, :deq.shape[1]]
            sub = ((codes & 0x7C) == 0) & ((codes & 0x03) != 0)
        else:
            continue
        out[r0:r1][sub] = 0.0
    return out


@pytest.mark.parametrize("pa,pb", SYRK_CASES)
def test_syrk_classes(handle, oracle, pa, pb):
    """C -= Q Q^T with the tiles of Q in formats (pa, pb, pa, ...): exact
    products of the dequantised operands, FP32 accumulation (error bound
    k * u32 relative to sum |C| + |Q||Q|^T)."""
    m, k, tr = 3 * 128 + 70, 256, 128
    say("syrk", L.PREC_NAMES[pa], L.PREC_NAMES[pb], m, k)
    rng = np.random.default_rng(pa * 10 + pb)
    P = (rng.standard_normal((m, k)) * np.exp(rng.uniform(-3, 3, (m, 1)))).astype(np.float32)
    Cm = rng.standard_normal((m, m)).astype(np.float32)
    nt = (m + tr - 1) // tr
    precs = [pa if t % 2 == 0 else pb for t in range(nt)]
    pay, sc, pt = S.quantize_panel(handle, to_dev_cm(P), precs, tr)
    pay_o, _, deq = oracle.quantize_panel(P, precs, tr)
    Cd = to_dev_cm(Cm)
    S.syrk(handle, Cd, pay, sc, k, tr, pt)
    got = from_dev_cm(Cd)
    low = np.tril(np.ones((m, m), bool))
    errs = {}
    for name, D in [("keep", deq), ("flush", flush_fp8_subnormals(deq, pay_o, precs, tr))]:
        D = D.astype(np.float64)
        ref = Cm.astype(np.float64) - D @ D.T
        scale = np.abs(Cm).astype(np.float64) + np.abs(D) @ np.abs(D).T
        errs[name] = (np.abs(got - ref)[low] / scale[low]).max()
    print("syrk max rel err", L.PREC_NAMES[pa], L.PREC_NAMES[pb], errs, flush=True)
    # the FP8 matrix-core path accumulates with slightly looser rounding than a
    # sequential FP32 sum (measured ~2.4e-5 of sum |terms| at k = 256)
    tol = 1e-4 if max(pa, pb) >= L.PREC_E5M2 else 2e-5
    assert errs["keep"] < tol, errs
    assert np.array_equal(np.triu(got, 1), np.triu(Cm, 1))   # upper untouched


def test_syrk_ones_kat(handle, oracle):
    """syrk_gemm.cuh:252-279 generalised: A = ones(m x k) -> A A^T = k exactly."""
    m, k = 256, 256
    say("syrk ones KAT")
    for prec in [L.PREC_E4M3, L.PREC_BF16, L.PREC_FP16, L.PREC_FP32, L.PREC_MXE5M2]:
        P = np.ones((m, k), np.float32)
        pay, sc, pt = S.quantize_panel(handle, to_dev_cm(P), [prec, prec], 128)
        Cd = to_dev_cm(np.zeros((m, m), np.float32))
        S.syrk(handle, Cd, pay, sc, k, 128, pt)
        got = from_dev_cm(Cd)
        assert np.all(got[np.tril_indices(m)] == -float(k)), L.PREC_NAMES[prec]


def test_syrk_masked_matches_syrk(handle):
    """mpc_syrk_masked with the masks of the formats present gives mpc_syrk's
    result bit for bit; a mask naming no compute class is refused."""
    import ctypes as C
    m, k, tr = 3 * 128 + 70, 256, 128
    rng = np.random.default_rng(5)
    P = rng.standard_normal((m, k)).astype(np.float32)
    Cm = rng.standard_normal((m, m)).astype(np.float32)
    cases = [([L.PREC_E4M3] * 4, 1 << L.PREC_E4M3, 1),
             ([L.PREC_FP32, L.PREC_E4M3, L.PREC_E5M2, L.PREC_FP32],
              (1 << L.PREC_FP32) | (1 << L.PREC_E4M3), 3)]
    for precs, cmask, fmask in cases:
        say("syrk_masked", [L.PREC_NAMES[p] for p in precs])
        pay, sc, pt = S.quantize_panel(handle, to_dev_cm(P), precs, tr)
        C1, C2 = to_dev_cm(Cm), to_dev_cm(Cm)
        S.syrk(handle, C1, pay, sc, k, tr, pt)
        rc = handle.lib.mpc_syrk_masked(handle.h, C.c_void_p(C2.data_ptr()), m, m, k,
                                        C.c_void_p(pay.data_ptr()), pay.stride(0),
                                        C.c_void_p(sc.data_ptr()), tr, C.c_void_p(pt.data_ptr()),
                                        cmask, fmask)
        assert rc == 0, rc
        torch.cuda.synchronize()
        assert torch.equal(C1, C2)
    rc = handle.lib.mpc_syrk_masked(handle.h, C.c_void_p(C2.data_ptr()), m, m, k,
                                    C.c_void_p(pay.data_ptr()), pay.stride(0),
                                    C.c_void_p(sc.data_ptr()), tr, C.c_void_p(pt.data_ptr()),
                                    1 << 7, 1)
    assert rc == L.ERR_ARG


# ---------------------------------------------------------------- generators
def test_matgen_householder(handle, oracle):
    n = 1000
    say("matgen householder", n)
    A, lam = S.matgen_householder(handle, n, cond=1e6, lam_max=1e6, seed=7)
    Ag = from_dev_cm(A)
    Ao, lam_o = oracle.householder_spd(n, 1e6, 1e6, 7)
    assert np.array_equal(Ag, Ag.T)
    assert np.allclose(lam.cpu().numpy(), lam_o, rtol=1e-15, atol=0)
    rel = np.abs(Ag - Ao).max() / np.abs(Ao).max()
    assert rel < 1e-6, rel
    ev = np.linalg.eigvalsh(Ag.astype(np.float64))
    assert np.allclose(np.sort(ev)[-10:], np.sort(lam_o)[-10:], rtol=1e-6)


def test_rbf_points_and_kernel(handle, oracle):
    say("rbf points + assembly")
    X1 = S.rbf_points(handle, 500, 20, kind=1, p_zero=0.8, seed=5).cpu().numpy()
    assert np.array_equal(X1, oracle.points_sparse_image(500, 20, 0.8, 5))
    X0 = S.rbf_points(handle, 300, 16, kind=0, seed=9).cpu().numpy()
    assert np.allclose(X0, oracle.points_gaussian(300, 16, 9), rtol=1e-6, atol=1e-6)
    for ls, nug, X in [(4.0, 1e-2, X0), ([0.5 + 0.1 * i for i in range(16)], 1e-3, X0),
                       (3.0, 0.0, X1)]:
        Xd = torch.from_numpy(np.ascontiguousarray(X)).cuda()
        K = from_dev_cm(S.rbf_assemble(handle, Xd, ls, 1.0, nug))
        Ko = np.asarray(oracle.rbf_matrix(X, ls, 1.0, nug))
        assert np.abs(K - Ko).max() < 2e-6
        assert np.array_equal(K, K.T)


# ---------------------------------------------------------------- factor + PCG
def hh(oracle, n, seed=11):
    A, lam = oracle.householder_spd(n, 1e6, 1e6, seed)
    return A


@pytest.mark.parametrize("n,nb,prec,ladder,agg,rule", [
    (1024, 256, L.PREC_FP32, (L.PREC_FP32,), 4, 0),
    (1024, 256, L.PREC_SWITCH, (L.PREC_E4M3, L.PREC_FP32), 4, 0),
    (1100, 128, L.PREC_SWITCH, (L.PREC_MXE4M3, L.PREC_BF16, L.PREC_FP32), 4, 0),
    (2048, 256, L.PREC_SWITCH, (L.PREC_E5M2, L.PREC_FP32), 4, 0),
    (700, 64, L.PREC_SWITCH, (L.PREC_E4M3, L.PREC_FP32), 4, 0),
    # updates per panel (no aggregation), and groups of 3 (a ragged last group)
    (2048, 256, L.PREC_SWITCH, (L.PREC_E4M3, L.PREC_FP32), 1, 0),
    (2560, 256, L.PREC_SWITCH, (L.PREC_E4M3, L.PREC_FP32), 3, 0),
    (1100, 128, L.PREC_SWITCH, (L.PREC_MXE4M3, L.PREC_BF16, L.PREC_FP32), 2, 0),
])
def test_factor_matches_oracle(handle, oracle, n, nb, prec, ladder, agg, rule):
    say("factor", n, nb, prec, ladder, agg, rule)
    factor_parity(handle, oracle, hh(oracle, n), nb, prec, ladder, agg, rule, 0.25)


def lowrank_spd(n, rank, seed):
    """I + V V^T with |v_i|^2 ~ 64: the Schur diagonal stays near a_ii until
    the first rank rows are eliminated and then drops to about 1/65 of it."""
    V = np.random.default_rng(seed).standard_normal((n, rank)) * (8 / np.sqrt(rank))
    return (np.eye(n) + V @ V.T).astype(np.float32)


@pytest.mark.parametrize("case,nb,ladder,rule,eps", [
    ("hh1536", 256, (L.PREC_E4M3, L.PREC_FP32), L.RULE_FROBENIUS, 0.01),
    ("hh1100", 128, (L.PREC_E5M2, L.PREC_FP32), L.RULE_FROBENIUS, 0.05),
    ("lowrank1536", 256, (L.PREC_E4M3, L.PREC_FP32), L.RULE_SCHUR, 0.003),
])
def test_factor_rules_match_oracle(handle, oracle, case, nb, ladder, rule, eps):
    """The tile-Frobenius and Schur-diagonal rules against the oracle, on
    matrices and thresholds where they give a mixed schedule and the factor
    succeeds (both rules ignore the pivots and break down on these matrices
    at eps' = 0.25, in the oracle as on the GPU)."""
    say("factor rule", case, rule, eps)
    A = lowrank_spd(1536, 384, 3) if case.startswith("lowrank") else hh(oracle, int(case[2:]))
    gmap = factor_parity(handle, oracle, A, nb, L.PREC_SWITCH, ladder, 2, rule, eps)
    below = np.tril(np.ones_like(gmap, bool), -1)
    assert len(np.unique(gmap[below])) == 2          # a mixed schedule


def factor_parity(handle, oracle, A, nb, prec, ladder, agg, rule, eps):
    n = A.shape[0]
    T = to_dev_cm(A)
    opts = S.factor_opts(nb=nb, update_prec=prec, ladder=ladder, eps_prime=eps, fallback=0,
                         agg_panels=agg, rule=rule)
    rc, fi, mp, dg = S.factor(handle, T, opts, want_map=True, want_diag=True)
    assert rc == 0, rc
    Lg = np.tril(from_dev_cm(T))
    gmap = mp.cpu().numpy().T          # gmap[i, j] = precision of tile L_ij
    # oracle on the GPU's schedule (threshold flips would otherwise differ)
    Lo, rco, omap, udo, ofi = oracle.factor_mixed(A, nb=nb, update_prec=prec, ladder=ladder,
                                                  eps_prime=eps, map_in=gmap,
                                                  tile_rows=max(nb, 128))
    assert rco == 0
    # the map holds one entry per nb block of L below the diagonal, also when
    # a row tile spans two blocks (nb = 64, 128-row tiles)
    below = np.tril(np.ones_like(gmap, bool), -1)
    assert (gmap[below] != 255).all() and np.array_equal(omap[below], gmap[below])
    Lo = np.tril(Lo)
    rel = np.linalg.norm(Lg - Lo) / np.linalg.norm(Lo)
    # FP32-only schedules agree to rounding; with 8-bit tiles an occasional RNE
    # decision flips between the two rounding orders and moves an element by
    # up to u_fp8, which the conditioning (1e6) carries through the factor
    has8 = any(p >= L.PREC_E5M2 for p in ladder) or prec >= L.PREC_E5M2
    assert rel < (5e-3 if has8 else 1e-4), rel
    # both factors are equally good approximations of A
    A64 = A.astype(np.float64)
    bg = np.linalg.norm(Lg.astype(np.float64) @ Lg.T.astype(np.float64) - A64) / np.linalg.norm(A64)
    bo = np.linalg.norm(Lo.astype(np.float64) @ Lo.T.astype(np.float64) - A64) / np.linalg.norm(A64)
    assert bg < 2 * bo + 1e-6, (bg, bo)
    # the oracle's own rule gives (nearly) the same schedule
    _, _, omap2, _, _ = oracle.factor_mixed(A, nb=nb, update_prec=prec, ladder=ladder,
                                            eps_prime=eps, tile_rows=max(nb, 128), rule=rule)
    low = np.tril(np.ones_like(gmap, bool), -1) & (omap2 != 255)
    f = max(nb, 128) // nb                    # map entries per row tile
    flips = np.sum(gmap[low] != omap2[low]) // f
    assert flips <= max(1, low.sum() // f // 50), (flips, low.sum() // f)
    # Schur-diagonal tracker
    assert np.allclose(dg.cpu().numpy(), udo, rtol=1e-3, atol=1e-3 * np.abs(np.diag(A)).max())
    # flop accounting, per compute class (same schedule: the oracle ran on the GPU's map)
    assert abs(sum(fi.syrk_flops) - sum(ofi.flops)) / sum(ofi.flops) < 1e-9
    for c in range(8):
        assert abs(fi.syrk_flops[c] - ofi.flops[c]) <= 1e-9 * sum(ofi.flops), (c, list(fi.syrk_flops),
                                                                              list(ofi.flops))
    return gmap


@pytest.mark.parametrize("rule,eps_hi,eps_lo", [
    (L.RULE_MINPIVOT, 1.05, 0.95), (L.RULE_FROBENIUS, 1.05, 0.95), (L.RULE_SCHUR, 0.04, 0.03),
])
def test_switching_rules_closed_form(handle, oracle, rule, eps_hi, eps_lo):
    """A = L L^T with L = [[2I, 0], [B, 8I]], B = +-1/2 (256-blocks): each rule's
    threshold falls where its closed form puts it (u_E4M3 = 1/16), on the GPU
    as in the oracle, and the factor is exact in both formats."""
    m = 256
    B = np.where(np.random.default_rng(12).random((m, m)) < 0.5, -0.5, 0.5)
    Lx = np.zeros((2 * m, 2 * m))
    Lx[:m, :m] = 2 * np.eye(m)
    Lx[m:, :m] = B
    Lx[m:, m:] = 8 * np.eye(m)
    A = (Lx @ Lx.T).astype(np.float32)
    for eps, want in ((eps_hi, L.PREC_E4M3), (eps_lo, L.PREC_FP32)):
        say("rule", rule, eps)
        T = to_dev_cm(A)
        opts = S.factor_opts(nb=256, update_prec=L.PREC_SWITCH, ladder=(L.PREC_E4M3, L.PREC_FP32),
                             eps_prime=eps, fallback=0, rule=rule)
        rc, fi, mp, dg = S.factor(handle, T, opts, want_map=True, want_diag=True)
        assert rc == 0
        gmap = mp.cpu().numpy().T
        assert gmap[1, 0] == want, (rule, eps, gmap)
        np.testing.assert_allclose(np.tril(from_dev_cm(T)), Lx, rtol=0, atol=4e-6)
        np.testing.assert_allclose(dg.cpu().numpy()[m:], 64.0, rtol=0, atol=1e-4)


@pytest.mark.parametrize("prec,agg", [
    (L.PREC_E4M3, 2), (L.PREC_MXE4M3, 2), (L.PREC_E5M2, 1), (L.PREC_MXE5M2, 3),
])
def test_factor_fixed_fp8_matches_oracle(handle, oracle, prec, agg):
    """Every trailing update in one 8-bit format (the all-FP8 factor of
    fp8_chol.cuh:125-152, here with per-tile or per-32 E8M0 scales): the GPU
    factor against the oracle's, tile formats all equal to the fixed one."""
    n = 1536
    say("factor fixed", L.PREC_NAMES[prec], agg)
    A, _ = oracle.householder_spd(n, 1e2, 1e2, 23)
    T = to_dev_cm(A)
    opts = S.factor_opts(nb=256, update_prec=prec, ladder=(prec,), fallback=0, agg_panels=agg)
    rc, fi, mp, dg = S.factor(handle, T, opts, want_map=True, want_diag=True)
    assert rc == 0, rc
    gmap = mp.cpu().numpy().T
    low = np.tril(np.ones_like(gmap, bool), -1)
    assert np.all(gmap[low] == prec)
    assert fi.syrk_flops[L.PREC_E4M3] == sum(fi.syrk_flops)      # FP8 class only
    Lg = np.tril(from_dev_cm(T))
    Lo, rco, omap, udo, ofi = oracle.factor_mixed(A, nb=256, update_prec=prec, ladder=(prec,),
                                                  tile_rows=256)
    assert rco == 0
    assert np.array_equal(omap[low], gmap[low])
    Lo = np.tril(Lo)
    rel = np.linalg.norm(Lg - Lo) / np.linalg.norm(Lo)
    assert rel < 5e-3, rel
    A64 = A.astype(np.float64)
    bg = np.linalg.norm(Lg.astype(np.float64) @ Lg.T.astype(np.float64) - A64) / np.linalg.norm(A64)
    bo = np.linalg.norm(Lo.astype(np.float64) @ Lo.T.astype(np.float64) - A64) / np.linalg.norm(A64)
    assert bg < 2 * bo + 1e-6, (bg, bo)
    assert abs(sum(fi.syrk_flops) - sum(ofi.flops)) <= 1e-9 * sum(ofi.flops)


@pytest.mark.parametrize("n,nb,ladder,nrhs", [
    (1024, 256, (L.PREC_FP32,), 1),
    (1536, 256, (L.PREC_E4M3, L.PREC_FP32), 1),
    (1024, 128, (L.PREC_MXE4M3, L.PREC_FP32), 3),
])
def test_pcg_matches_oracle(handle, oracle, n, nb, ladder, nrhs):
    say("pcg", n, nb, ladder, nrhs)
    A = hh(oracle, n, seed=n)
    rng = np.random.default_rng(n)
    B = rng.standard_normal((nrhs, n))
    Ad = to_dev_cm(A)
    Bd = torch.from_numpy(B).cuda()
    fo = S.factor_opts(nb=nb, update_prec=L.PREC_SWITCH if len(ladder) > 1 else ladder[0],
                       ladder=ladder, eps_prime=0.25)
    po = S.pcg_opts(tol=1e-10, maxit=200)
    X, fi, pi, rc = S.precond_cg(handle, Ad, Bd, fo, po)
    assert rc == 0, (rc, S.info_dict(pi))
    Xh = X.cpu().numpy()
    A64 = A.astype(np.float64)
    for j in range(nrhs):
        rr = np.linalg.norm(B[j] - A64 @ Xh[j]) / np.linalg.norm(B[j])
        assert rr <= 1e-10, rr
    # same factor on the CPU: run the oracle PCG on the GPU's L
    Lg = torch.empty_like(Ad)
    Lg.copy_(Ad)
    rc2, fi2, _, _ = S.factor(handle, Lg, fo)
    assert rc2 == 0
    Lnp = np.tril(from_dev_cm(Lg))
    _, it_o, rr_o, kap = oracle.pcg(A, Lnp, B[0], tol=1e-10, maxit=200)
    assert it_o > 0
    assert abs(pi.iters - it_o) <= 2, (pi.iters, it_o)


@pytest.mark.parametrize("n", [1000, 3000, 9000])
def test_apply_precond_matches_fp64(handle, oracle, n):
    """z = (L L^T)^{-1} r through the blocked big-block-inverse sweeps vs an FP64
    triangular solve with the same L (n = 3000 and 9000: ragged big blocks)."""
    import scipy.linalg as sl
    say("apply_precond", n)
    A = hh(oracle, n, seed=n)
    T = to_dev_cm(A)
    rc, fi, _, _ = S.factor(handle, T, S.factor_opts(nb=256, update_prec=L.PREC_FP32,
                                                       ladder=(L.PREC_FP32,)))
    assert rc == 0
    Lh = np.tril(from_dev_cm(T)).astype(np.float64)
    R = np.random.default_rng(n).standard_normal((3, n))
    Z = S.apply_precond(handle, T, torch.from_numpy(R).cuda()).cpu().numpy()
    for j in range(3):
        ze = sl.solve_triangular(Lh.T, sl.solve_triangular(Lh, R[j], lower=True), lower=False)
        err = np.linalg.norm(Z[j] - ze) / np.linalg.norm(ze)
        assert err < 1e-4, (j, err)


def test_factor_breakdown_reports_info(handle):
    """An indefinite matrix: info = failing column (no fallback), and the
    fallback path ends with the same info after trying FP32."""
    say("factor breakdown")
    n = 512
    A = np.eye(n, dtype=np.float32) * 4
    A[300, 300] = -2.0
    T = to_dev_cm(A)
    rc, fi, _, _ = S.factor(handle, T, S.factor_opts(nb=256, fallback=0))
    assert rc == 301
    T = to_dev_cm(A)
    rc, fi, _, _ = S.factor(handle, T, S.factor_opts(nb=256, fallback=1))
    assert rc == 301 and fi.retries >= 1


def test_reference_api(handle, oracle):
    from cholesky_gpu_amd import reference_api as R
    say("reference api")
    n = 512
    dA = torch.empty(n, n, dtype=torch.float64, device="cuda")
    R.generatePSD(dA, n, 100.0, R.DistType.Arithmetic, h=handle)
    lam = np.linalg.eigvalsh(dA.cpu().numpy())
    assert abs(lam.max() / lam.min() - 100.0) / 100.0 < 1e-3
    b = torch.from_numpy(np.random.default_rng(0).standard_normal(n)).cuda()
    x = torch.zeros_like(b)
    rc = R.precond_CG(dA, x, b, n, 128, h=handle, tol=1e-12)
    assert rc == 0
    r = b.cpu().numpy() - dA.cpu().numpy() @ x.cpu().numpy()
    assert np.linalg.norm(r) / np.linalg.norm(b.cpu().numpy()) < 1e-10


@pytest.mark.parametrize("check_every", [0, 1, 3])
def test_pcg_graph_replay_matches_eager(handle, oracle, check_every):
    """Each iteration replayed as a HIP graph (several replays between the
    host's eager convergence checks) gives bitwise the eager result."""
    say("pcg graph vs eager", check_every)
    n = 1536
    A = hh(oracle, n, seed=77)
    Ad = to_dev_cm(A)
    Lf = Ad.clone()
    rc, fi, _, _ = S.factor(handle, Lf, S.factor_opts(nb=256))
    assert rc == 0
    B = torch.from_numpy(np.random.default_rng(1).standard_normal((3, n))).cuda()
    res = []
    for graph in (0, 1):
        X, pi, rc = S.pcg(handle, Ad, Lf, B, nb=256,
                          opts=S.pcg_opts(tol=1e-10, maxit=200, graph=graph, check_every=check_every))
        assert rc == 0, S.info_dict(pi)
        res.append((X.cpu().numpy(), pi.iters))
    assert res[0][1] == res[1][1]
    assert np.array_equal(res[0][0], res[1][0])


def test_pcg_cached_graph_matches_eager(handle, oracle):
    """A second solve in the same buffers replays the iteration graph kept in
    the handle (no new capture); with new right-hand sides in those buffers
    it still gives bitwise the eager result, and a third solve with another
    tolerance (a different graph) too."""
    say("pcg cached graph")
    n = 1536
    A = hh(oracle, n, seed=78)
    Ad = to_dev_cm(A)
    Lf = Ad.clone()
    rc, fi, _, _ = S.factor(handle, Lf, S.factor_opts(nb=256))
    assert rc == 0
    B = torch.empty((2, n), dtype=torch.float64, device="cuda")
    X = torch.zeros_like(B)
    for seed, tol in ((2, 1e-10), (3, 1e-10), (4, 1e-9)):
        B.copy_(torch.from_numpy(np.random.default_rng(seed).standard_normal((2, n))))
        out = []
        for graph in (1, 0):
            X.zero_()
            _, pi, rc = S.pcg(handle, Ad, Lf, B, nb=256, X=X,
                              opts=S.pcg_opts(tol=tol, maxit=200, graph=graph, check_every=0))
            assert rc == 0, S.info_dict(pi)
            out.append((X.cpu().numpy().copy(), pi.iters))
        assert out[0][1] == out[1][1], (seed, out[0][1], out[1][1])
        assert np.array_equal(out[0][0], out[1][0]), seed


@pytest.mark.parametrize("agg,ladder", [
    (2, (L.PREC_E4M3, L.PREC_FP32)),
    (3, (L.PREC_E5M2, L.PREC_MXE4M3, L.PREC_BF16, L.PREC_FP32)),
])
def test_precond_cg_deferred_copy(handle, oracle, agg, ladder):
    """precond_cg copies only the first agg_panels block columns of A and lets
    the first aggregated update read the rest from A: the solution and the
    factor statistics are bitwise those of factoring a full copy and then
    running the PCG on it."""
    n = 2304
    say("precond_cg deferred copy", agg, [L.PREC_NAMES[p] for p in ladder])
    A = hh(oracle, n, seed=91)
    Ad = to_dev_cm(A)
    B = torch.from_numpy(np.random.default_rng(9).standard_normal((2, n))).cuda()
    fo = S.factor_opts(nb=256, update_prec=L.PREC_SWITCH, ladder=ladder, eps_prime=0.25,
                       fallback=0, agg_panels=agg)
    po = S.pcg_opts(tol=1e-10, maxit=300)
    X1, fi1, pi1, rc1 = S.precond_cg(handle, Ad, B, fo, po)
    assert rc1 == 0, S.info_dict(pi1)
    Lf = Ad.clone()
    rc, fi2, _, _ = S.factor(handle, Lf, fo)
    assert rc == 0
    X2, pi2, rc2 = S.pcg(handle, Ad, Lf, B, nb=256, opts=po)
    assert rc2 == 0
    assert list(fi1.syrk_flops) == list(fi2.syrk_flops)
    assert sum(fi1.syrk_flops[p] for p in range(8) if p != L.PREC_FP32) > 0   # mixed classes
    assert fi1.syrk_flops[L.PREC_FP32] > 0
    assert pi1.iters == pi2.iters
    assert torch.equal(X1, X2)
    assert torch.equal(Ad, to_dev_cm(A))            # A itself untouched


def test_factor_schedule_invariant(handle, oracle):
    """Per-panel updates (agg_panels = 1) give the same factor bit for bit
    under the look-ahead schedule (block column on its own stream in 64-row
    tiles, the rest per precision class) and the sequential schedule with
    per-phase timing: each element sees the same updates in the same order and
    each tile product the same K order.  The aggregated schedule sums a
    group's products before subtracting them, so it agrees to rounding, and
    is itself deterministic."""
    say("factor schedule invariance")
    A = hh(oracle, 2304)
    outs = []
    for agg, prof in [(1, 0), (1, 1), (4, 0), (4, 0)]:
        T = to_dev_cm(A)
        o = S.factor_opts(nb=256, update_prec=L.PREC_SWITCH, ladder=(L.PREC_E4M3, L.PREC_FP32),
                          eps_prime=0.25, fallback=0, agg_panels=agg, profile=prof)
        rc, fi, mp, _ = S.factor(handle, T, o, want_map=True)
        assert rc == 0
        outs.append((np.tril(T.cpu().numpy()), mp.cpu().numpy()))
    assert np.array_equal(outs[1][1], outs[0][1])
    assert np.array_equal(outs[1][0], outs[0][0])
    assert np.array_equal(outs[3][1], outs[2][1])
    assert np.array_equal(outs[3][0], outs[2][0])
    rel = np.linalg.norm(outs[2][0] - outs[0][0]) / np.linalg.norm(outs[0][0])
    assert rel < 5e-3, rel
