// mpc_internal.hpp -- shared declarations of libmpchol (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <vector>
#include "../../include/mpchol.h"

// ---------------------------------------------------------------- device helpers
// True in every thread of the one workgroup that arrives last (of `total`)
// at counter *t, which that workgroup resets to zero for the next launch.
// The release fence before the increment and the acquire fence after it make
// every workgroup's earlier global writes visible to the last one.  Nothing
// waits: the other workgroups simply end.  The agent-scope release writes
// back the XCD's L2 (about 1 us each, serialised per XCD on gfx950), so this
// is for launches of a few workgroups only.
static __device__ __forceinline__ bool mpc_last_arrival(int* t, int total) {
    __shared__ int flag;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const int last = atomicAdd(t, 1) == total - 1;
        if (last) atomicExch(t, 0);
        flag = last;
    }
    __syncthreads();
    const bool last = flag != 0;
    if (last) __threadfence();
    return last;
}

// ---------------------------------------------------------------- errors
void mpc_set_error(const char* fmt, ...);

#define MPC_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            mpc_set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call,   \
                          hipGetErrorString(e_));                              \
            return MPC_ERR_HIP;                                                \
        }                                                                      \
    } while (0)

#define MPC_CHECK(cond, ...)                                                   \
    do {                                                                       \
        if (!(cond)) {                                                         \
            mpc_set_error(__VA_ARGS__);                                        \
            return MPC_ERR_ARG;                                                \
        }                                                                      \
    } while (0)

#define MPC_LAUNCH_CHECK()                                                     \
    do {                                                                       \
        hipError_t e_ = hipGetLastError();                                     \
        if (e_ != hipSuccess) {                                                \
            mpc_set_error("%s:%d: kernel launch failed: %s", __FILE__, __LINE__,\
                          hipGetErrorString(e_));                              \
            return MPC_ERR_HIP;                                                \
        }                                                                      \
    } while (0)

// ---------------------------------------------------------------- precision
// unit roundoff u_p of each storage/compute format
inline double mpc_unit_roundoff(int p) {
    switch (p) {
    case MPC_PREC_FP32: return 5.9604644775390625e-08;   // 2^-24
    case MPC_PREC_BF16: return 3.90625e-03;              // 2^-8
    case MPC_PREC_FP16: return 4.8828125e-04;            // 2^-11
    case MPC_PREC_E5M2: case MPC_PREC_MXE5M2: return 0.125;
    default: return 0.0625;                              // E4M3: 2^-4
    }
}
inline bool mpc_is_fp8(int p) { return p >= MPC_PREC_E5M2 && p <= MPC_PREC_MXE5M2; }
inline bool mpc_valid_prec(int p) { return p >= 0 && p <= MPC_PREC_MXE5M2; }

// ---------------------------------------------------------------- handle
struct mpc_workspace {
    void*  ptr = nullptr;
    size_t bytes = 0;
};

struct mpc_handle {
    // the last PCG iteration graph and the arguments it was captured with
    hipGraphExec_t        pcg_exec = nullptr;
    std::vector<uint64_t> pcg_key;
    // CG coefficients of the last PCG call (first right-hand side)
    std::vector<double>   last_alpha, last_beta;
    int          device = 0;
    hipStream_t  stream = nullptr;
    bool         own_stream = false;
    // look-ahead: the next block column's update and panel (POTRF, TRSM,
    // quantise) run on pstream while the rest of the current trailing update
    // runs on stream; both at the default priority
    hipStream_t  pstream = nullptr;
    hipEvent_t   ev_col = nullptr;     // factor start (on stream)
    hipEvent_t   ev_panel = nullptr;   // next panel ready (on pstream)
    hipEvent_t   ev_bulk = nullptr;    // bulk update of a step done (on stream)
    // aggregated updates of the block columns beyond the next group's window
    hipStream_t  fstream = nullptr;
    hipEvent_t   ev_far = nullptr;     // an aggregated far update done (on fstream)
    // big-block inverses of L for the PCG sweeps, built on wstream during the
    // factor as each big diagonal block becomes final
    hipStream_t  wstream = nullptr;
    hipEvent_t   ev_wready = nullptr;  // a big block of L is final
    hipEvent_t   ev_wdone = nullptr;   // all inverses built (on wstream)
    hipEvent_t   ev_wpre = nullptr;    // all but the last big block's built (on wstream)
    // the factor returns with the inverses still being built on wstream:
    // users of winv / dinv / the factor buffer wait for ev_wdone first (the
    // first preconditioner apply of a PCG only before its last big block)
    bool         winv_pending = false;
    // factor workspace
    mpc_workspace panel;      // quantised panel payload, n * nb * 4 bytes
    mpc_workspace scales;     // n * nb/32 bytes
    mpc_workspace dinv;       // inverses of 32x32 diagonal sub-blocks, n * 32 floats
    mpc_workspace small;      // per-step scalars and tile stats
    mpc_workspace rowstat;    // row sum of squares (n floats) + tracker (n floats)
    mpc_workspace amat;       // copy of A for precond_cg / fallback
    mpc_workspace pcg;        // PCG vectors
    mpc_workspace scratch;    // generic
    mpc_workspace winv;       // inverses of the big diagonal blocks of L (Bo x Bo each)
    mpc_workspace tlist;      // SYRK work lists, one region per factor stream
    mpc_workspace gp64;       // GP: FP64 operator A (n x n)
    mpc_workspace gp32;       // GP: FP32 copy of A (factor source)
    mpc_workspace gpvec;      // GP: right-hand sides, solutions, U, V, partials
    mpc_workspace tick;       // arrival counters of split reductions (zero between launches)
    // provenance of the dinv blocks: the factor they belong to
    const float* dinv_owner = nullptr;
    int64_t      dinv_n = 0;
    int64_t      dinv_ld = 0;
    int          num_cu = 0;
    // provenance of the big-block inverses
    const float* winv_owner = nullptr;
    int64_t      winv_n = 0, winv_ld = 0;
    int          winv_bo = 0;
};

// Workgroup barrier that orders LDS accesses only: outstanding global loads
// and stores are not waited for (a __syncthreads() fence waits for them),
// for kernels that exchange data through LDS alone.
__device__ __forceinline__ void mpc_lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// grow-only device allocation in a handle workspace
int mpc_ws_reserve(mpc_workspace& w, size_t bytes);
// count zeroed arrival counters (grow-only, cleared when allocated); nullptr
// with the error set on failure
int* mpc_tickets(mpc_handle* h, size_t count);

// noalloc.hip: looks up the current device's static scratch array (once per
// device; mpc_handle_create calls it)
int mpc_noalloc_init_device(int device);

// ---------------------------------------------------------------- small device state
// layout of handle->small (all device):
struct mpc_small {
    int      info;            // 1-based failing column, 0 = ok
    int      nan_flag;        // NaN/Inf seen in a panel
    uint32_t mu_bits;         // min pivot^2 so far (bits of a positive float; atomicMin)
    // mu_bits as the POTRF of a step left it, by step parity: the tile
    // decisions of that step's panel read this copy, which stays fixed while
    // the next step's POTRF lowers mu_bits
    uint32_t mu_step[2];
    // by step parity too: ||L_jj||_F^2 of the step's diagonal block (POTRF)
    // and the largest ratio tracker / original diagonal below it (TRSM;
    // bits of a non-negative float, atomicMax), for the switching rules
    float    diag_f2[2];
    uint32_t schur_max[2];
    uint32_t pad;
    // per-step tile stats (index by tile within the panel), max 1024 tiles,
    // double-buffered by step parity so panel j+1 can be prepared while the
    // trailing update of panel j still reads panel j's buffers
    uint32_t tile_amax[2][1024];   // bits of positive floats (atomicMax)
    uint32_t tile_maxrs[2][1024];  // max row sum of squares
    // decided precision per panel tile, one slot per panel buffer (a panel
    // stays live until the aggregated update of its group has read it)
    uint8_t  tile_prec[2 * 4][1024];
};
constexpr int MPC_AGG_MAX = 4;                    // panels per aggregated update
constexpr int MPC_PANEL_SLOTS = 2 * MPC_AGG_MAX;  // panel buffers in the ring
static_assert(sizeof(((mpc_small*)nullptr)->tile_prec) / 1024 == MPC_PANEL_SLOTS, "slots");

// ---------------------------------------------------------------- kernels (launchers)
// k_potrf.hip
int mpc_launch_potrf(hipStream_t st, float* A, int64_t lda, int bs, int col0,
                     float* dinv, mpc_small* sm, int ntiles_next, uint32_t* stats_amax,
                     uint32_t* stats_maxrs, int mu_slot = 0, int want_f2 = 0);
// k_panel.hip
// rows [row0, row1) of the s-row panel (row1 < 0: s); the panel kernels
// index tiles and statistics by the row within the whole panel
int mpc_launch_trsm_stats(hipStream_t st, float* A21, int64_t lda, int64_t s, int nb,
                          const float* dinv_j, const float* Ljj, float* rowss, float* tracker,
                          int tile_rows, mpc_small* sm, uint32_t* tile_amax, uint32_t* tile_maxrs,
                          int64_t row0 = 0, int64_t row1 = -1, const float* diag0 = nullptr,
                          int slot = 0);
int mpc_launch_quantize_ex(hipStream_t st, const float* P, int64_t m, int k, int64_t ld,
                           int tile_rows, const uint8_t* prec_in, const uint32_t* tile_amax,
                           const uint32_t* tile_maxrs, mpc_small* sm, int update_prec,
                           const int* ladder, int n_ladder, double eps_prime, uint8_t* prec_out,
                           uint8_t* map_col, uint8_t* payload, int64_t row_stride, uint8_t* scales,
                           int* nan_flag, int64_t row0 = 0, int64_t row1 = -1, int mu_slot = 0,
                           int map_rep = 1, int64_t map_len = -1, int rule = MPC_RULE_MINPIVOT,
                           const float* rowss = nullptr);
int mpc_launch_tile_amax(hipStream_t st, const float* P, int64_t m, int k, int64_t ld,
                         int tile_rows, uint32_t* tile_amax);
// k_syrk.hip
int mpc_launch_syrk(hipStream_t st, float* C, int64_t m, int64_t ldc, int k,
                    const uint8_t* payload, int64_t row_stride, const uint8_t* scales,
                    int tile_rows, const uint8_t* prec_of_tile, const mpc_small* sm);
constexpr int MPC_SYRK_TILE = 128;   // output tile edge of the SYRK kernels
// Several quantised panels whose products are summed into one update
// (K = np * k).  The panels sit in a ring of `slots` buffers (payload,
// scales, 1024 tile precisions each); panel q of the update (q < np) is
// panel p0 + q, in slot (p0 + q) % slots, and C's row 0 is its row
// rows0 - q * nb (each panel starts nb rows below the previous one).
// Scalars only: a kernel argument of pointer arrays costs scalar registers.
constexpr int MPC_SYRK_MAXP = 8;
struct mpc_syrk_panels {
    const uint8_t* payload;
    const uint8_t* scales;
    const uint8_t* prec;
    int64_t pstride, sstride;    // bytes between slots (prec: 1024 per slot)
    int64_t rows0;
    int slots, p0, np, nb;
    // optional: C's values are read from csrc (same leading dimension) by
    // the first launch of the class sequence that touches a tile, and written
    // to C; later launches read C.  nullptr: C in place.
    const float* csrc;
};
// C(lower) -= sum_p Q_p Q_p^T over tile columns [tcol0, tcol1), tiles of
// 64 * qa rows: qa = 1 one launch with the class chosen per workgroup and
// panel; qa = 2 one launch per class in class_mask / fp8 pair in fp8_mask.
// trow0 > 0: only the tile rows from trow0 (>= tcol1: a rectangle below the
// diagonal)
// lists (optional, mpc_syrk_list_ints(m, 64 qa) ints on the device, private to
// the stream): with several class launches, the first builds work lists of
// the tile pairs each later launch has work on, and those run over their
// lists only
int mpc_launch_syrk_panels(hipStream_t st, float* C, int64_t m, int64_t ldc, int k,
                           const mpc_syrk_panels& P, int64_t row_stride, int tile_rows,
                           const mpc_small* sm, int64_t tcol0, int64_t tcol1, int qa,
                           unsigned class_mask = 0xffu, unsigned fp8_mask = 3u,
                           int* lists = nullptr, int64_t trow0 = 0);
int64_t mpc_syrk_list_ints(int64_t m, int tile);
int mpc_launch_syrk_classes(hipStream_t st, float* C, int64_t m, int64_t ldc, int k,
                            const uint8_t* payload, int64_t row_stride, const uint8_t* scales,
                            int tile_rows, const uint8_t* prec_of_tile, const mpc_small* sm,
                            unsigned class_mask, unsigned fp8_mask, int64_t tcol0 = 0,
                            int64_t tcol1 = -1, int mixed = 0,    // mixed: 0 or 1
                            int* lists = nullptr, int64_t trow0 = 0);
// compute classes / FP8 element formats that a factor configuration can produce
void mpc_syrk_masks(int update_prec, const int* ladder, int n_ladder, unsigned* class_mask,
                    unsigned* fp8_mask);
// k_pcg.hip
int mpc_launch_symv(hipStream_t st, const float* A, int64_t lda, int64_t n, int nrhs,
                    const double* X, int64_t ldx, double* Y, int64_t ldy, double* part);
int mpc_launch_symv64(hipStream_t st, const double* A, int64_t lda, int64_t n, int nrhs,
                      const double* X, int64_t ldx, double* Y, int64_t ldy, double* part);
int64_t mpc_symv_partial_doubles(int64_t n, int nrhs);
int mpc_launch_dot(hipStream_t st, int64_t n, int nrhs, const double* a, int64_t lda_,
                   const double* b, int64_t ldb, double* out, double* part);
// out_b = a . b and out_c = a . c per column in two launches (part: twice
// the partials of one dot); each sum in the same order as mpc_launch_dot's
int mpc_launch_dot2(hipStream_t st, int64_t n, int nrhs, const double* a, int64_t lda_,
                    const double* b, int64_t ldb, double* out_b, const double* c, int64_t ldc,
                    double* out_c, double* part);
int mpc_launch_pcg_update_xr(hipStream_t st, int64_t n, int nrhs, double* X, int64_t ldx,
                             double* R, int64_t ldr, const double* P, int64_t ldp,
                             const double* Q, int64_t ldq, const double* rz, const double* pq,
                             const int* active);
int mpc_launch_pcg_update_p(hipStream_t st, int64_t n, int nrhs, double* P, int64_t ldp,
                            const double* Z, int64_t ldz, const double* rz_new,
                            const double* rz_old, const int* active);
int mpc_launch_pcg_status(hipStream_t st, int nrhs, double* rz, const double* rz_new,
                          const double* pq, const double* rr, const double* bb, double tol,
                          int stop_rule, int* active, int* cand, int* iters, double* relres,
                          double* ahist, double* bhist, int hmax);
int mpc_launch_axpby(hipStream_t st, int64_t n, int nrhs, double a, const double* X,
                     int64_t ldx, double b, double* Y, int64_t ldy);
int mpc_launch_copy2d(hipStream_t st, int64_t n, int nrhs, const double* X, int64_t ldx,
                      double* Y, int64_t ldy);
int mpc_launch_copy_f32(hipStream_t st, int64_t rows, int64_t cols, const float* A,
                        int64_t lda, float* B, int64_t ldb);
// k_matgen.hip
int mpc_launch_hh_matrix(hipStream_t st, int64_t n, const double* lam, const double* W,
                         const double* P, const double* g, float* A, int64_t lda);
int mpc_launch_points(hipStream_t st, float* X, int64_t n, int d, int64_t ldx, int kind,
                      double p_zero, uint64_t seed);
int mpc_launch_rbf(hipStream_t st, const float* X, int64_t n, int d, int64_t ldx,
                   const float* inv_ls, int n_ls, double variance, double nugget, float* K,
                   int64_t ldk);
int mpc_launch_rbf64(hipStream_t st, const float* X, int64_t n, int d, int64_t ldx,
                     const double* inv_ls, int n_ls, double variance, double nugget, double* K64,
                     int64_t ldk64, float* K32, int64_t ldk32);
constexpr int MPC_GP_MAXD = 16;   // ARD dimensions of the fused gradient contraction
int64_t mpc_gp_partial_doubles(int64_t n, int nout);
int mpc_launch_gp_grad(hipStream_t st, const float* X, int64_t n, int d, int64_t ldx,
                       const double* inv_ls, int n_ls, double variance, const double* U,
                       const double* V, int64_t ldu, int R, double* part, double* out);
int mpc_launch_logdiag(hipStream_t st, const float* L, int64_t ldl, int64_t n, double* out);
int mpc_launch_probes(hipStream_t st, double* Z, int64_t ldz, int64_t n, int nprobe, uint64_t seed,
                      int unit);
int mpc_launch_convert_f8(hipStream_t st, const float* x, uint8_t* y, int64_t count, int prec);
int mpc_launch_dinv_from_L(hipStream_t st, const float* L, int64_t ldl, int64_t n, float* dinv);
// block PCG helpers (k_pcg.hip): s <= MPC_BLOCK_PCG_MAX columns
int64_t mpc_gram_partial_doubles(int64_t n, int s);
// G = X^T Y (s x s); tick: one zeroed counter (left zeroed)
int mpc_launch_gram(hipStream_t st, int64_t n, int s, const double* X, int64_t ldx,
                    const double* Y, int64_t ldy, double* part, double* G, int* tick);
// X = G^-1 H (s x s, G SPD); *flag = 1 and X = 0 on a non-positive pivot
int mpc_launch_spd_solve(hipStream_t st, int s, const double* G, const double* H, double* X,
                         int* flag);
// Y = (Z ? Z : Y) + a M C (n x s blocks, s x s C); Y may be M
int mpc_launch_block_axpy(hipStream_t st, int64_t n, int s, const double* M, int64_t ldm,
                          const double* Cm, double a, const double* Z, int64_t ldz, double* Y,
                          int64_t ldy);
// k_tri.hip
int mpc_build_block_inverses_range(hipStream_t st, const float* L, int64_t ldl, int64_t n, int Bo,
                                   const float* dinv, float* W, float* WT, int64_t I0, int64_t I1);
int mpc_build_block_inverses(hipStream_t st, const float* L, int64_t ldl, int64_t n, int Bo,
                             const float* dinv, float* W, float* WT);
// part: mpc_block_prec_partial_doubles(Bo, nrhs) doubles of scratch
// z = L^-T L^-1 r (R in, Y work, Z out; ld each)
int mpc_apply_block_prec(hipStream_t st, const float* L, int64_t ldl, int64_t n, int Bo,
                         const float* Winv, const float* WinvT, const double* R, double* Y,
                         double* Z, int64_t ld, int nrhs, double* part,
                         hipEvent_t late_ev = nullptr, int64_t I_late = -1);
int64_t mpc_block_prec_partial_doubles(int64_t n, int Bo, int nrhs);
inline int mpc_block_bo(int64_t n) { return n > 8192 ? 2048 : (n > 2048 ? 1024 : 256); }

// host-side generators shared with mpc_matgen (RNG identical to the device one)
void mpc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double mpc_normal_host(uint64_t seed, uint64_t stream, uint64_t idx);
double mpc_uniform_host(uint64_t seed, uint64_t stream, uint64_t idx);
