/*
 * oracle.h -- CPU restatement of the mixed-precision Cholesky-PCG path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in this directory is part of the
 * product: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it, and only as the checker (or the timed CPU baseline).  The HIP
 * library never links or calls it.
 *
 * Every function names the reference file:line it restates.  Where the
 * reference is defective or silent (see the deviations list in oracle.c) the
 * restatement follows the intended semantics and says so.
 *
 * Pinning status (what each function is checked against, tests/test_oracle_*.py):
 *   orc_f32_to_e4m3 / e5m2 ... lo_float special-value table, closed-form RNE
 *                              (frexp/ldexp), exhaustive decode round trip
 *   orc_f32_to_bf16 / f16 .... numpy float16 (RNE) and bfloat16 closed form
 *   orc_scale_exp ............ closed form (smallest power of two)
 *   orc_potrf_f32_micro ...... oracle/_ref build of micro_chol.hpp (bitwise)
 *   orc_potrf_f64 ............ 2x2/3x3 closed forms, numpy.linalg.cholesky
 *   orc_factor_mixed ......... FP64 oracle backward error, all-FP32 path vs
 *                              numpy, all-ones SYRK known answer
 *   orc_pcg .................. exact FP64 solve, iteration-count invariants
 *   orc_refine ............... exact factor converges in one correction; error
 *                              contraction bounded by ||I - M^-1 A|| (numpy);
 *                              fixed point equals the FP64 solve
 *   orc_block_pcg ............ s = 1 equals orc_pcg; k distinct eigenvalues
 *                              converge in ceil(k / s) block steps; FP64 solve
 *   orc_hh_* ................. spectrum of H*Lambda*H^T equals Lambda (numpy
 *                              eigvalsh), symmetry
 *   orc_philox4x32 ........... Random123 published known-answer vectors
 *   orc_rbf_* ................ closed form K_ii = 1 + nugget, symmetry, scipy cdist
 */
#ifndef MPCHOL_ORACLE_H
#define MPCHOL_ORACLE_H
#include <stdint.h>
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

/* precision codes: identical numbering to include/mpchol.h */
enum { ORC_FP32 = 0, ORC_BF16 = 1, ORC_FP16 = 2, ORC_E5M2 = 3, ORC_E4M3 = 4,
       ORC_MXE4M3 = 5, ORC_MXE5M2 = 6 };

/* ---- number formats (lo_float.h:1896-2117 restated; IEEE RNE for 16-bit) */
uint8_t  orc_f32_to_e4m3(float x);
uint8_t  orc_f32_to_e5m2(float x);
float    orc_e4m3_to_f32(uint8_t v);
float    orc_e5m2_to_f32(uint8_t v);
uint16_t orc_f32_to_bf16(float x);
uint16_t orc_f32_to_f16(float x);
float    orc_bf16_to_f32(uint16_t v);
float    orc_f16_to_f32(uint16_t v);
void     orc_f32_to_e4m3_n(const float* x, uint8_t* y, int64_t n);
void     orc_f32_to_e5m2_n(const float* x, uint8_t* y, int64_t n);
/* disagreements between the conversion of float bit patterns [lo, hi) and an
   expected table, among finite inputs with |x| <= max_abs; *other receives
   the disagreements outside that domain */
int64_t  orc_sweep_e4m3(uint64_t lo, uint64_t hi, const uint8_t* expect, double max_abs,
                        int64_t* other);
int64_t  orc_sweep_e5m2(uint64_t lo, uint64_t hi, const uint8_t* expect, double max_abs,
                        int64_t* other);
/* independent frexp/ldexp closed form and the exhaustive self-check against it */
uint8_t  orc_closed_form_e4m3(float x);
uint8_t  orc_closed_form_e5m2(float x);
int64_t  orc_selfcheck_f8(int e4m3, uint64_t lo, uint64_t hi);
/* C (lower) -= Q Q^T with FP32 accumulation (the restated trailing update) */
void     orc_syrk_lower(float* C, int64_t m, int64_t ldc, const float* Q, int64_t ldq, int64_t k);

int   orc_elem_bytes(int prec);
double orc_unit_roundoff(int prec);
/* power-of-two scale exponent: smallest e with amax / 2^e <= max_finite(prec),
   clamped to the E8M0 range [-127, 127]; 0 for amax == 0 or unscaled formats */
int   orc_scale_exp(float amax, int prec);

/* Quantise an m x k panel (column-major, leading dim ld) row-tile by row-tile.
   prec_of_tile[t] is the format of rows [t*tile_rows, (t+1)*tile_rows).
   Output payload is row-major, k-contiguous, row stride row_stride bytes;
   scales is m x (k/32) E8M0 bytes (biased exponent, 127 = 1.0).  deq (optional)
   receives the dequantised FP32 values, column-major with leading dim ld_deq.
   Returns 0, or -1 if a NaN/Inf was found. */
int orc_quantize_panel(const float* P, int64_t m, int64_t k, int64_t ld,
                       int tile_rows, const uint8_t* prec_of_tile,
                       uint8_t* payload, int64_t row_stride, uint8_t* scales,
                       float* deq, int64_t ld_deq);

/* ---- random numbers: Philox4x32-10 (Salmon et al., SC'11) */
void   orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* element idx of stream `stream` under `seed`: uniform in [0,1) with 53 bits */
double orc_uniform(uint64_t seed, uint64_t stream, uint64_t idx);
/* Rademacher trace probes: z_pi = -1 if orc_uniform(seed, 3, p*n + i) < 1/2, else +1 */
void orc_probes(int64_t n, int nprobe, uint64_t seed, double* Z, int64_t ldz);
/* standard normal via Box-Muller on the pair (2*(idx/2), +1) */
double orc_normal(uint64_t seed, uint64_t stream, uint64_t idx);

/* ---- generators (matgen.cuh:188-298 / kernels_macros.cuh:635-675 API;
   the configurations replace the QR-based generator, see oracle.c) */
void orc_spectrum(int dist, int64_t n, double cond, double lam_max,
                  int64_t n_cluster, double cluster_lo, double cluster_hi,
                  uint64_t seed, double* lam);
/* Householder product H = H_1...H_r, H_k = I - 2 w_k w_k^T: returns per-row
   4-vectors (r <= 4, padded with zeros) W, P = W*T and g = G*p with
   G = W^T Lambda W, so that (H Lambda H^T)_ij =
       lam_i [i==j] - lam_j <p_i,w_j> - lam_i <p_j,w_i> + <p_i, g_j>        */
int  orc_hh_setup(int64_t n, int nrefl, uint64_t seed, const double* lam,
                  double* W, double* P, double* g);
double orc_hh_entry(int64_t i, int64_t j, const double* lam, const double* W,
                    const double* P, const double* g);
void orc_hh_matrix(int64_t n, const double* lam, const double* W, const double* P,
                   const double* g, float* A, int64_t lda);
/* RBF kernel K_ij = variance * exp(-0.5 * sum_d ((x_id - x_jd)/ls_d)^2) + nugget*[i==j]
   X row-major n x d (ldx >= d).  n_ls is 1 (isotropic) or d (ARD). */
void orc_rbf_matrix(int64_t n, int d, const float* X, int64_t ldx, const double* ls,
                    int n_ls, double variance, double nugget, float* K, int64_t ldk);
void orc_points_gaussian(int64_t n, int d, uint64_t seed, float* X, int64_t ldx);
void orc_points_sparse_image(int64_t n, int d, double p_zero, uint64_t seed,
                             float* X, int64_t ldx);

/* ---- FP64 oracle */
int  orc_potrf_f64(double* A, int64_t n, int64_t lda);            /* lower, blocked */
void orc_trsv_lower_f64(const double* L, int64_t n, int64_t ldl, double* x, int trans);
void orc_symv_f32_f64(const float* A, int64_t n, int64_t lda, const double* x, double* y);

/* ---- restatement of the reference path (FP32) */
/* micro_chol.hpp:14-44: unblocked right-looking, in place lower, column-major.
   Returns 0 or the 1-based column of the first non-positive pivot. */
int  orc_potrf_f32_micro(float* A, int n, int ld);

typedef struct {
    int    nb;            /* block size */
    int    update_prec;   /* precision code, or -1 = per-tile switching */
    int    ladder[4];     /* switching ladder, lowest precision first */
    int    n_ladder;
    double eps_prime;     /* switching threshold */
    int    tile_rows;     /* rows per precision tile (= nb) */
    int    rule;          /* 0 min pivot, 1 Frobenius, 2 Schur diagonal (D3) */
} orc_factor_opts;

typedef struct {
    int    info;          /* 0 or 1-based failing column */
    double min_pivot;     /* min L_kk^2 */
    double flops[8];      /* trailing-update flops by compute precision */
} orc_factor_info;

/* Right-looking blocked factor in place (lower).  If map_in != NULL the
   per-tile precisions are taken from it (nt x nt, entry [i + j*nt] for panel
   tile (i,j), i>j) instead of the rule, so the CPU and GPU factors use the
   same schedule.  map_out (optional) receives the schedule.  updated_diag
   (optional, length n) receives the Schur-diagonal tracker.
   max_steps < 0 runs all steps (>= 0 stops early; used for bounded timing). */
int orc_factor_mixed(float* A, int64_t n, int64_t lda, const orc_factor_opts* o,
                     const uint8_t* map_in, uint8_t* map_out, float* updated_diag,
                     int max_steps, orc_factor_info* info);
/* the switching rule for one tile (tile_amax: largest magnitude in the tile;
   FP16 is skipped above 65504; orc_rule_minpivot: the rule without that test) */
int orc_rule_minpivot_amax(const int* ladder, int n_ladder, double eps_prime,
                           double max_row_sumsq, double min_pivot, double tile_amax);
int orc_rule_minpivot(const int* ladder, int n_ladder, double eps_prime,
                      double max_row_sumsq, double min_pivot);

typedef struct {
    double tol;
    int    maxit;
    int    stop_rule;     /* 0: ||b-Ax||/||b|| (true residual), 1: sqrt(r.z)/||b|| (reference) */
    int    x0_mode;       /* 0: x0 = M^-1 b (reference), 1: x0 = 0 */
} orc_pcg_opts;

/* mixed_cholesky.cu:145-217.  A FP32 (full, column-major), L FP32 lower.
   Returns iterations (>= 0) if converged, -1 if maxit reached.  relres
   receives the final true relative residual.  alphas/betas (optional, length
   maxit) receive the CG coefficients. */
int orc_pcg(const float* A, int64_t lda, const float* L, int64_t ldl, int64_t n,
            const double* b, double* x, const orc_pcg_opts* o, double* relres,
            double* alphas, double* betas);

/* Iterative refinement x += M^-1 (b - A x) from x0 (x0_mode 0: M^-1 b,
   1: 0) until the FP64 true residual meets tol.  Returns corrections
   applied, -1 at maxit; hist (optional, maxit + 1) gets the residual of
   every iterate. */
int orc_refine(const float* A, int64_t lda, const float* L, int64_t ldl, int64_t n,
               const double* b, double* x, const orc_pcg_opts* o, double* relres,
               double* hist);

/* Block PCG (O'Leary 1980) over s right-hand sides (B, X: n x s, ld n).
   Returns iterations, -1 at maxit, -2 on breakdown of the block recurrence;
   relres (s values) gets the final true relative residuals. */
int orc_block_pcg(const float* A, int64_t lda, const float* L, int64_t ldl, int64_t n, int s,
                  const double* B, double* X, const orc_pcg_opts* o, double* relres);

/* cond estimate of the preconditioned operator from CG coefficients
   (Lanczos tridiagonal eigenvalue extremes, bisection) */
double orc_lanczos_cond(const double* alphas, const double* betas, int k);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
