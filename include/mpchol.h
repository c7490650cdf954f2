/*
 * mpchol.h -- C ABI of libmpchol.so: mixed-precision blocked Cholesky used as
 * a preconditioner for conjugate gradients, on one MI355X (gfx950).
 *
 * Drop-in boundary for the reference's driver API (paths relative to the
 * reference repository):
 *   mpc_matgen         <- generatePSD            src/matgen.cuh:188-194
 *                         (DistType ordinals     src/kernels_macros.cuh:24-29)
 *   mpc_factor         <- switching_chol         src/switching_chol.cuh:19-22
 *   mpc_pcg            <- PCG loop of precond_CG src/mixed_cholesky.cu:145-217
 *   mpc_precond_cg     <- precond_CG             src/mixed_cholesky.cu:76-78
 *   mpc_quantize_panel <- convert_fp32_to_mxe4m3 src/fp8_utils.cuh:27-66 and
 *                         convertFloatToHalf     src/kernels_macros.cuh:115-124
 *   mpc_syrk,          <- LtMxfp8Matmul / LtHSgemm_fp16in_fp32out / LtSgemm
 *   mpc_syrk_masked
 *                         src/fp8_utils.cuh:136,304,231 and the WMMA SYRK
 *                         src/syrk_gemm.cuh:14-124
 *   mpc_potrf_block    <- cusolverDnSpotrf call  src/switching_chol.cuh:86-95
 *                         (CPU form micro_cholesky src/micro_chol.hpp:14-44)
 *   mpc_apply_precond  <- the cusolverDnDpotrs calls of mixed_cholesky.cu:147,166,204
 *   (update_diag, kernels_macros.cuh:177-190, is fused into mpc_factor: its
 *    tracker comes back through d_updated_diag)
 *   mpc_rbf_* / mpc_gp_* : no reference counterpart (the RBF/GP workloads and
 *                         the backward pass are specified by the workload
 *                         definitions only; src/RBF.cu holds no RBF code).
 *
 * Conventions: matrices are column-major with an explicit leading dimension;
 * every pointer named d* is device memory owned by the caller.  All calls are
 * asynchronous on the handle's stream except those that return host values,
 * which synchronise once.  Return value: 0 on success; > 0 LAPACK-style info
 * (1-based column of the first non-positive pivot); < 0 an MPC_ERR_* code.
 * A handle is not thread-safe: use one handle per host thread.
 */
#ifndef MPCHOL_H
#define MPCHOL_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPC_VERSION_MAJOR 0
#define MPC_VERSION_MINOR 1

/* ---- status codes (negative) ------------------------------------------ */
#define MPC_ERR_ARG            (-1)    /* an argument is invalid (see last error) */
#define MPC_ERR_HIP            (-100)  /* HIP runtime failure */
#define MPC_ERR_NOT_CONVERGED  (-101)  /* PCG hit maxit */
#define MPC_ERR_NAN            (-102)  /* NaN/Inf in the matrix or a panel */
#define MPC_ERR_NOMEM          (-103)  /* device allocation failed */
#define MPC_ERR_UNSUPPORTED    (-104)  /* option combination not supported */
#define MPC_ERR_NO_DEVICE      (-105)

/* ---- precision codes (storage / compute format of a trailing-update tile) */
#define MPC_PREC_FP32    0
#define MPC_PREC_BF16    1
#define MPC_PREC_FP16    2
#define MPC_PREC_E5M2    3   /* OCP E5M2, one power-of-two scale per tile      */
#define MPC_PREC_E4M3    4   /* OCP E4M3, one power-of-two scale per tile      */
#define MPC_PREC_MXE4M3  5   /* OCP MX: E4M3 + E8M0 scale per 32 along k       */
#define MPC_PREC_MXE5M2  6   /* OCP MX: E5M2 + E8M0 scale per 32 along k       */
#define MPC_PREC_SWITCH  (-1)/* per-tile switching over opts.ladder           */

/* ---- spectrum distributions: 0..3 keep the reference's DistType ordinals */
#define MPC_DIST_RANDOM_LOG_UNIFORM 0
#define MPC_DIST_CLUSTERED          1
#define MPC_DIST_ARITHMETIC         2
#define MPC_DIST_GEOMETRIC          3
#define MPC_DIST_LOGSPACED          4   /* lambda_i increasing, log-spaced */

/* ---- matrix kinds for mpc_matgen */
#define MPC_MAT_HOUSEHOLDER_SPD 0       /* A = H Lambda H^T, H = prod of reflectors */

typedef struct mpc_handle mpc_handle;

/* ---- handle --------------------------------------------------------------- */
int         mpc_handle_create(mpc_handle** h, int device);
int         mpc_handle_destroy(mpc_handle* h);
/* use a caller-owned hipStream_t (never destroyed by the library); NULL = own */
int         mpc_set_stream(mpc_handle* h, void* hip_stream);
void*       mpc_get_stream(mpc_handle* h);
int         mpc_synchronize(mpc_handle* h);
const char* mpc_last_error_string(void);
const char* mpc_version_string(void);

/* ---- matrix generation (generatePSD, matgen.cuh:188-298) ----------------- */
typedef struct {
    int      kind;          /* MPC_MAT_HOUSEHOLDER_SPD */
    int      dist;          /* MPC_DIST_* for the main part of the spectrum */
    double   cond;          /* ratio of the main part: lambda in [lam_max/cond, lam_max] */
    double   lam_max;
    int      n_reflectors;  /* 1..4 */
    uint64_t seed;
    int64_t  n_cluster;     /* leading eigenvalues log-spaced in [cluster_lo, cluster_hi] */
    double   cluster_lo;
    double   cluster_hi;
} mpc_matgen_opts;
void mpc_matgen_opts_default(mpc_matgen_opts* o);
/* Writes A (n x n, FP32, both triangles).  d_lambda (optional, device, n
   doubles) receives the exact spectrum. */
int mpc_matgen(mpc_handle* h, float* dA, int64_t n, int64_t lda,
               const mpc_matgen_opts* o, double* d_lambda);

/* ---- RBF / GP assembly (no reference counterpart) ----------------------- */
/* points, row-major n x d (ld >= d): kind 0 = N(0, I_d); kind 1 = each
   coordinate 0 with probability p_zero, else uniform in (0, 1]            */
int mpc_rbf_points(mpc_handle* h, float* dX, int64_t n, int d, int64_t ldx,
                   int kind, double p_zero, uint64_t seed);
/* K_ij = variance * exp(-1/2 sum_c ((x_ic - x_jc)/ls_c)^2) + nugget [i==j];
   n_ls = 1 (isotropic) or d (ARD); ls is a host array */
int mpc_rbf_assemble(mpc_handle* h, const float* dX, int64_t n, int d, int64_t ldx,
                     const double* ls, int n_ls, double variance, double nugget,
                     float* dK, int64_t ldk);
/* median pairwise Euclidean distance of nsub points picked without
   replacement by a seeded permutation (host result) */
int mpc_median_pairwise_distance(mpc_handle* h, const float* dX, int64_t n, int d,
                                 int64_t ldx, int64_t nsub, uint64_t seed, double* out);

/* ---- factorisation (switching_chol, switching_chol.cuh:19-227) --------- */
typedef struct {
    int    nb;              /* block size: 64, 128 or 256 (default) */
    int    update_prec;     /* MPC_PREC_* for every tile, or MPC_PREC_SWITCH */
    int    ladder[4];       /* switching ladder, lowest precision first */
    int    n_ladder;
    double eps_prime;       /* switching threshold (rule below) */
    int    fallback;        /* 1: on a non-positive pivot restore A, divide
                               eps_prime by 16 and retry (twice), then all-FP32 */
    int    profile;         /* 1: time the phases with events (adds syncs) */
    int    agg_panels;      /* panels whose trailing updates beyond their group
                               of block columns are applied together
                               (K = agg_panels * nb): 1..4, default 2 */
    int    rule;            /* MPC_RULE_* below (default MPC_RULE_MINPIVOT) */
} mpc_factor_opts;
/* Switching rules.  Each picks, for a panel tile L_ij (i > j) of panel j, the
   lowest-precision ladder entry p (unit roundoff u_p) with
       u_p * stat  <=  eps_prime * ref,
   FP16 skipped for a tile with an entry above 65504; the last ladder entry
   when none qualifies.  A tile pair (i, l) of the trailing update is computed
   in the higher of the two formats.
     MPC_RULE_MINPIVOT (0): stat = max over the tile's rows r of sum_k L_rk^2
       (k over panel j), ref = min_k L_kk^2 over the diagonal blocks factored so
       far.
     MPC_RULE_FROBENIUS (1): stat = ||L_ij||_F^2, ref = ||L_jj||_F^2 (the tile's
       Frobenius norm relative to its diagonal block).
     MPC_RULE_SCHUR (2): stat = max over the rows r below the panel of
       d_r / a_rr, ref = 1, one decision for every tile of the panel, where d_r
       is the Schur-diagonal tracker after panel j (a_rr - sum_{k<=j blocks}
       L_rk^2) and a_rr the original diagonal: the per-step test of can_switch
       (kernels_macros.cuh:191-211) evaluated deterministically. */
#define MPC_RULE_MINPIVOT  0
#define MPC_RULE_FROBENIUS 1
#define MPC_RULE_SCHUR     2

typedef struct {
    int    info;            /* 0, or 1-based column of the first bad pivot */
    int    retries;         /* fallback restarts used */
    double eps_prime_used;
    double min_pivot;       /* min L_kk^2 */
    double flops_total;     /* n^3/3 + n^2/2 + n/6 */
    double syrk_flops[8];   /* trailing-update flops by compute class (index = prec code) */
    int64_t tiles[8];       /* panel tiles by storage precision */
    double ms_total;        /* wall time of the call (device, events) */
    double ms_potrf, ms_trsm, ms_quant, ms_syrk;   /* profile = 1 only */
} mpc_factor_info;
void mpc_factor_opts_default(mpc_factor_opts* o);
/* In place, lower, column-major (upper triangle untouched).  d_prec_map
   (optional, device, nt*nt bytes, nt = ceil(n/nb)) receives the precision of
   each panel tile L_ij at [i + j*nt] (diagonal tiles: FP32; upper: 255).
   d_updated_diag (optional, device, n floats) receives the Schur-diagonal
   tracker of update_diag.  info may be NULL. */
int mpc_factor(mpc_handle* h, float* dA, int64_t n, int64_t lda, const mpc_factor_opts* o,
               uint8_t* d_prec_map, float* d_updated_diag, mpc_factor_info* info);

/* ---- PCG (precond_CG loop, mixed_cholesky.cu:145-217) ------------------ */
typedef struct {
    double tol;             /* target ||b - A x||_2 / ||b||_2 (FP64) */
    int    maxit;
    int    stop_rule;       /* 0: true residual (default); 1: reference sqrt(r.z)/||b|| */
    int    x0_mode;         /* 0: x0 = M^-1 b (reference); 1: x0 = 0; 2: x0 from dX */
    int    graph;           /* 1 (default): replay each iteration as one HIP graph */
    int    check_every;     /* iterations between host convergence checks;
                               0 (default): adaptive, from the observed rate */
} mpc_pcg_opts;
typedef struct {
    int    iters;           /* max over right-hand sides */
    int    converged;       /* number of converged right-hand sides */
    double relres;          /* max true relative residual */
    double kappa_est;       /* Lanczos estimate of cond(M^-1 A) (first rhs) */
    double ms_total;
} mpc_pcg_info;
void mpc_pcg_opts_default(mpc_pcg_opts* o);
/* Solves A X = B for nrhs columns.  A: FP32 n x n (full storage, only read);
   L: FP32 lower factor from mpc_factor with block size nb; B, X: FP64 n x nrhs.
   Returns 0 if all columns converged, MPC_ERR_NOT_CONVERGED otherwise. */
int mpc_pcg(mpc_handle* h, const float* dA, int64_t lda, const float* dL, int64_t ldl,
            int64_t n, int nb, int nrhs, const double* dB, int64_t ldb, double* dX,
            int64_t ldx, const mpc_pcg_opts* o, mpc_pcg_info* info);

/* ---- iterative refinement and block PCG (multi-RHS) --------------------- */
/* Mixed-precision iterative refinement around the factor: x0 (x0_mode 0:
   M^-1 b, 1: 0, 2: from dX), then x += M^-1 (b - A x) until every column's
   FP64 true residual ||b - A x|| / ||b|| <= tol.  The residual and the
   update are FP64, the correction goes through the FP32 factor L.  The
   reference has no such loop; its error trace of one
   (src/e5m2_error_f_cond.csv:1-5, error per refinement step) is what it
   produces.  Converges when the spectral radius of I - M^-1 A is below 1
   (for FP8-update factors of ill-conditioned A it may not be: PCG then
   still converges).  info->iters = corrections applied.  Arguments and
   returns as mpc_pcg (stop_rule, graph and check_every are not used). */
int mpc_refine(mpc_handle* h, const float* dA, int64_t lda, const float* dL, int64_t ldl,
               int64_t n, int nb, int nrhs, const double* dB, int64_t ldb, double* dX,
               int64_t ldx, const mpc_pcg_opts* o, mpc_pcg_info* info);

/* Block PCG (O'Leary 1980) over the nrhs columns together: one block
   Krylov space, s x s = nrhs x nrhs FP64 coefficient solves per step
   instead of the scalar alpha, beta of the single-vector loop
   (mixed_cholesky.cu:180-217).  nrhs <= MPC_BLOCK_PCG_MAX.  If the block
   recurrence breaks down (P^T A P or Z^T R not positive definite in FP64)
   the columns finish as independent PCGs (mpc_pcg) from the current
   iterate.  info->kappa_est is 0 (no Lanczos estimate).  Arguments and
   returns as mpc_pcg (stop_rule, graph and check_every are not used). */
#define MPC_BLOCK_PCG_MAX 16
int mpc_block_pcg(mpc_handle* h, const float* dA, int64_t lda, const float* dL, int64_t ldl,
                  int64_t n, int nb, int nrhs, const double* dB, int64_t ldb, double* dX,
                  int64_t ldx, const mpc_pcg_opts* o, mpc_pcg_info* info);

/* ---- factor + solve (precond_CG, mixed_cholesky.cu:76-229) ------------- */
/* Factors a workspace copy of A (A itself is not modified), then runs PCG.
   Returns 0 (converged), MPC_ERR_NOT_CONVERGED, or factor info > 0. */
int mpc_precond_cg(mpc_handle* h, const float* dA, int64_t lda, int64_t n, int nrhs,
                   const double* dB, int64_t ldb, double* dX, int64_t ldx,
                   const mpc_factor_opts* fo, const mpc_pcg_opts* po,
                   mpc_factor_info* finfo, mpc_pcg_info* pinfo);

/* PCG whose operator is an FP64 matrix: every product A p and the true
   residual use dA (FP64, column-major); the FP32 factor dL preconditions.
   The reference's PCG also multiplies by its FP64 A (the cublasDgemv of
   mixed_cholesky.cu:180-217).  Used where the solution must match an FP64
   solve (the GP gradient).  Arguments and returns as mpc_pcg. */
int mpc_pcg64(mpc_handle* h, const double* dA, int64_t lda, const float* dL, int64_t ldl,
              int64_t n, int nb, int nrhs, const double* dB, int64_t ldb, double* dX,
              int64_t ldx, const mpc_pcg_opts* o, mpc_pcg_info* info);

/* ---- GP loss and hand-written backward (no reference counterpart) ------ */
/* The RBF kernel of mpc_rbf_assemble evaluated in FP64 into dK64 (may be
   NULL), with an FP32 copy rounded from it into dK32 (may be NULL). */
int mpc_rbf_assemble64(mpc_handle* h, const float* dX, int64_t n, int d, int64_t ldx,
                       const double* ls, int n_ls, double variance, double nugget,
                       double* dK64, int64_t ldk64, float* dK32, int64_t ldk32);
/* Trace-estimation probes, n x nprobe column-major: Rademacher z_pi = -1 if
   the Philox uniform of stream 3 at index p*n + i is < 1/2, else +1; or the
   unit vectors e_p when unit != 0 (nprobe <= n). */
int mpc_probes(mpc_handle* h, double* dZ, int64_t ldz, int64_t n, int nprobe, uint64_t seed,
               int unit);
typedef struct {
    int      n_probes;      /* Rademacher probes for the logdet term (default 32) */
    int      probe_kind;    /* 0: Rademacher; 1: unit vectors, exact trace (n probes) */
    uint64_t probe_seed;
    mpc_factor_opts factor; /* preconditioner factor of the FP32 copy of A */
    mpc_pcg_opts    pcg;    /* adjoint solve; x0_mode is forced to 0 */
} mpc_gp_opts;
typedef struct {
    double loss;            /* datafit + logdet / 2 */
    double datafit;         /* b^T A^-1 b / 2 (from the FP64-operator solve) */
    double logdet;          /* 2 sum log L_ii of the preconditioner factor */
    mpc_factor_info factor;
    mpc_pcg_info    pcg;
} mpc_gp_info;
void mpc_gp_opts_default(mpc_gp_opts* o);
/* loss = b^T A^-1 b / 2 + logdet(A) / 2 with
   A = variance * exp(-1/2 sum_c ((x_ic - x_jc)/ls_c)^2) + noise [i==j]
   (n_ls = 1 isotropic or d ARD, d <= 16), and its gradient:
     d_grad_b (device, n)       = A^-1 b
     grad_ls (host, n_ls), grad_variance, grad_noise (host; each may be NULL)
   from one PCG solve of A [alpha, w_1..w_P] = [b, z_1..z_P] on the FP64
   operator, preconditioned by the mixed-precision factor of the FP32 copy of
   A, then a fused FP64 contraction of -alpha alpha^T / 2 + sum_p w_p z_p^T / (2P)
   against dA/dtheta that recomputes kernel tiles.  The logdet gradient is the
   Hutchinson estimate over the probes (exact for probe_kind = 1). */
int mpc_gp_loss_grad(mpc_handle* h, const float* dX, int64_t n, int d, int64_t ldx,
                     const double* ls, int n_ls, double variance, double noise,
                     const double* dB, const mpc_gp_opts* o, double* d_grad_b,
                     double* grad_ls, double* grad_variance, double* grad_noise,
                     mpc_gp_info* info);

/* ---- building blocks, exported for tests and advanced callers --------- */
/* Quantise an m x k panel (column-major, ld) in row tiles of tile_rows rows;
   prec_of_tile is a device array of ceil(m/tile_rows) codes.  FP32 and 16-bit
   tiles: payload row-major k-contiguous with row_stride bytes per row.  8-bit
   tiles: the matrix-core operand order, per 32-row block b and 64-wide chunk
   c a 2048-byte group at byte (32 b) * row_stride + 2048 c, in which lane
   l = n + 32 h holds 32 bytes (bytes 0-15: row 32 b + n, k in [64 c + 16 h,
   +16); bytes 16-31: k in [64 c + 32 + 16 h, +16)); rows of the last block
   beyond m are zero, so the payload needs 32 * ceil(m / 32) rows and
   row_stride >= 64 * ceil(k / 64).  scales is m x (k/32) E8M0 bytes.  k must
   be a multiple of 32. */
int mpc_quantize_panel(mpc_handle* h, const float* dP, int64_t m, int64_t k, int64_t ld,
                       int tile_rows, const uint8_t* d_prec_of_tile, uint8_t* d_payload,
                       int64_t row_stride, uint8_t* d_scales);
/* mpc_quantize_panel without any device allocation (no counterpart in the
   reference): for callers that cannot allocate (code under stream capture,
   serving loops, harnesses that forbid hipMalloc).  Same arguments, outputs
   and errors as mpc_quantize_panel; its scratch (one word per row tile and a
   flag) is a static per-device array of MPC_NOALLOC_MAX_TILES tiles, so
   ceil(m / tile_rows) above that is rejected (MPC_ERR_ARG).  Calls that
   overlap on one device race on that array. */
#define MPC_NOALLOC_MAX_TILES 1024
int mpc_quantize_panel_noalloc(mpc_handle* h, const float* dP, int64_t m, int64_t k, int64_t ld,
                               int tile_rows, const uint8_t* d_prec_of_tile, uint8_t* d_payload,
                               int64_t row_stride, uint8_t* d_scales);
/* C (m x m lower, FP32, ldc) -= Q Q^T for a quantised panel from
   mpc_quantize_panel (k multiple of 64, tile_rows multiple of 128). */
int mpc_syrk(mpc_handle* h, float* dC, int64_t m, int64_t ldc, int64_t k,
             const uint8_t* d_payload, int64_t row_stride, const uint8_t* d_scales,
             int tile_rows, const uint8_t* d_prec_of_tile);
/* mpc_syrk (the same reference calls) with the tile formats present named by
   the caller instead of read back from d_prec_of_tile: no device-to-host copy
   and no synchronisation, so back-to-back calls queue only kernels.
   class_mask: bit c set for every compute class c of a tile pair (the
   higher-precision code of the pair's two formats, MPC_PREC_E4M3 for two 8-bit
   formats); fp8_mask: bit 0 if an E4M3-element format (E4M3, MXE4M3) occurs,
   bit 1 if an E5M2-element one does.  Tile pairs outside the masks are not
   updated. */
int mpc_syrk_masked(mpc_handle* h, float* dC, int64_t m, int64_t ldc, int64_t k,
                    const uint8_t* d_payload, int64_t row_stride, const uint8_t* d_scales,
                    int tile_rows, const uint8_t* d_prec_of_tile, unsigned class_mask,
                    unsigned fp8_mask);
/* z = (L L^T)^{-1} r for nrhs columns (the preconditioner application of the
   PCG loop, the Dpotrs calls of mixed_cholesky.cu:147,166,204), FP64
   vectors, FP32 lower factor L from mpc_factor.  R is not modified. */
int mpc_apply_precond(mpc_handle* h, const float* dL, int64_t ldl, int64_t n, int nrhs,
                      const double* dR, int64_t ldr, double* dZ, int64_t ldz);
/* POTRF of one diagonal block (n <= 256) in place; host info result */
int mpc_potrf_block(mpc_handle* h, float* dA, int n, int64_t lda, int* info);
/* element-wise FP32 -> 8-bit conversion of a device array (prec E4M3 or E5M2),
   used for the exhaustive conversion sweep */
int mpc_convert_f8(mpc_handle* h, const float* dx, uint8_t* dy, int64_t count, int prec);
/* The CG coefficients of the last PCG call on this handle (first right-hand
   side): alphas[i], betas[i] for i < *count (count <= max).  kappa_est of
   mpc_pcg_info is the ratio of the extreme eigenvalues of the Lanczos
   tridiagonal they define (the condition estimate that replaces
   estimateConditionNumber, src/matgen.cuh:9-95, for the preconditioned
   operator).  Host arrays. */
int mpc_pcg_last_coeffs(mpc_handle* h, double* alphas, double* betas, int max, int* count);

/* ---- options file (settings.json, src/settings.json:1-14) -------------- */
typedef struct {
    int64_t n;                  /* matrix_settings.n */
    double  condition_number;   /* matrix_settings.condition_number */
    int     dist_type;          /* matrix_settings.dist_type as MPC_DIST_* */
    int     block_size;         /* factorization_settings.block_size as written */
    double  eps_prime;          /* factorization_settings.eps_prime */
    double  floor;              /* factorization_settings.floor (the reference's
                                   flr argument, unused by its solver) */
    int     diag_pert;          /* factorization_settings.diag_pert (perturb_diag) */
    int     left;               /* factorization_settings.left */
} mpc_settings;
void mpc_settings_default(mpc_settings* s);
/* Parse the JSON text of a settings file (replaces set_matrix_params,
   src/json_utils.hpp:2-32, which reads mismatched keys).  Values may be JSON
   strings holding numbers (as in the reference's file) or numbers; keys are
   accepted with underscores or spaces; absent keys keep the defaults. */
int mpc_settings_parse(const char* text, size_t len, mpc_settings* out);
int mpc_settings_read(const char* path, mpc_settings* out);
/* Options from settings: matgen dist and cond; factor nb, agg_panels and
   eps_prime.  block_size 64/128/256 -> nb; a * 256 (a = 2..4) -> nb 256 with
   a panels per aggregated update (the reference's panel width is the K of its
   update); anything else -> MPC_ERR_UNSUPPORTED.  Either output may be NULL. */
int mpc_settings_apply(const mpc_settings* s, mpc_matgen_opts* mo, mpc_factor_opts* fo);

/* ---- idx files (the MNIST data of src/MNIST_data) ---------------------- */
/* Header of a big-endian idx file: element type code (0x08 unsigned byte,
   ...), ndim (1..4), dims (unused entries 1), byte offset of the data.  The
   file size must match the header exactly. */
int mpc_idx_header(const char* path, int* type_code, int* ndim, int64_t dims[4],
                   int64_t* data_offset);
/* The first count items (all if count < 0) of an unsigned-byte idx file into
   out (host, out_bytes long); *items receives the number read. */
int mpc_idx_read_u8(const char* path, int64_t count, uint8_t* out, int64_t out_bytes,
                    int64_t* items);
/* The first n images (all if n < 0) of an idx3 image file as points: dX
   (device, row-major, ld ldx >= rows*cols) receives pixel / 255 in FP32;
   *n_out and *d_out the point count and dimension. */
int mpc_idx_points(mpc_handle* h, const char* path, int64_t n, float* dX, int64_t ldx,
                   int64_t* n_out, int* d_out);

#ifdef __cplusplus
}
#endif
#endif
