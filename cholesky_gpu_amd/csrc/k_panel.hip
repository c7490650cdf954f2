// k_panel.hip -- panel TRSM with fused row statistics, per-tile precision
// decision, and the fused tile quantiser.
//
//  trsm_stats: L21 = A21 * L_jj^{-T} (the RIGHT/LOWER/TRANS form of
//    fp8_chol.cuh:73-85; switching_chol.cuh:98-110 calls a LEFT form that is
//    out of bounds).  One wave owns a 32-row slab and substitutes over the
//    32-column blocks of the panel with the inverses of L_jj's 32x32 diagonal
//    sub-blocks (MFMA 32x32x2 f32) plus one refinement step per block.  The epilogue computes each row's sum of
//    squares, subtracts it from the Schur-diagonal tracker (update_diag,
//    kernels_macros.cuh:177-190) and folds the tile amax / max row sum of
//    squares into per-tile slots with order-independent atomics.
//  quantize: the panel becomes the SYRK operand: row-major k-contiguous
//    payload (8-bit tiles in the matrix-core operand order of
//    mpc_f8_offset) plus one E8M0 byte per row per 32 columns (replaces
//    convert_fp32_to_mxe4m3 / convert_transpose_fp32_to_mxe4m3,
//    fp8_utils.cuh:27-130, and convertFloatToHalf, kernels_macros.cuh:115-124).
//    Each workgroup decides its tile's precision from the tile statistics;
//    all workgroups of a tile evaluate the same expression on the same inputs.
#include "mpc_internal.hpp"
#include "mpc_formats.hpp"

namespace {

typedef float v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ int sigma(int s, int h) { return 8 * (s >> 2) + 4 * h + (s & 3); }

// one refinement step after the product with the explicit 32x32 inverse
constexpr bool kTrsmRefine = false;

// One workgroup of NW waves per 32-row slab.  Wave w owns the column blocks
// b = w, w + NW, ...  Step b: the owner finishes X_b = acc_b Dinv_b^T (one
// refinement step against L_bb when REF), publishes it in LDS; after the
// barrier every wave subtracts X_b L_{b'b}^T from its owned blocks b' > b,
// the soonest-needed block first.  The L operands for a step are loaded
// before the barrier so their latency overlaps the owner's solve.
// MFMA roles (32x32x2 f32, k index sig(t, h) on both operands): A = L or
// Dinv rows (lane m = lane % 32), B = X or acc columns (lane n = row of the
// slab); accumulator register i of lane (n, h) holds column sig(i, h), row n.
template <int NBK, int NW, bool REF>
__global__ __launch_bounds__(NW * 64, 2) void trsm_stats_kernel(float* __restrict__ A21, int64_t lda,
                                                         int64_t s, const float* __restrict__ dinv_j,
                                                         const float* __restrict__ Ljj,
                                                         float* __restrict__ rowss,
                                                         float* __restrict__ tracker, int tile_rows,
                                                         mpc_small* __restrict__ sm,
                                                         uint32_t* __restrict__ t_amax,
                                                         uint32_t* __restrict__ t_maxrs,
                                                         int64_t row0,
                                                         const float* __restrict__ diag0,
                                                         int slot) {
    constexpr int NO = NBK / NW;          // blocks per wave
    constexpr int XLD = 33;
    __shared__ float Xs[NBK][32 * XLD];
    __shared__ float red_ss[NW][64], red_am[NW][64];
    if (sm->info != 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 31, h = lane >> 5;
    const int64_t r0 = row0 + (int64_t)blockIdx.x * 32;
    const int64_t row = r0 + n;
    const bool valid = row < s;
    const int64_t rowc = valid ? row : s - 1;     // loads from clamped rows, then select

    v16f acc[NO];
#pragma unroll
    for (int o = 0; o < NO; ++o) {
        const int b = wave + NW * o;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const float x = A21[rowc + (int64_t)(32 * b + sigma(i, h)) * lda];
            acc[o][i] = valid ? x : 0.f;
        }
    }
    float ss = 0.f, am = 0.f;
#pragma unroll
    for (int b = 0; b < NBK; ++b) {
        // prefetch this step's L operands: -L(32 b' + m, 32 b + sig(t, h)) for owned b' > b
        float lv[NO][16];
#pragma unroll
        for (int o = 0; o < NO; ++o) {
            const int bp = wave + NW * o;
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const float x = Ljj[(32 * bp + n) + (int64_t)(32 * b + sigma(t, h)) * lda];
                lv[o][t] = bp > b ? -x : 0.f;
            }
        }
        if (b % NW == wave) {
            const int o = b / NW;
            const float* Db = dinv_j + b * 1024;
            v16f xb = {};
#pragma unroll
            for (int t = 0; t < 16; ++t)
                xb = __builtin_amdgcn_mfma_f32_32x32x2f32(Db[n + sigma(t, h) * 32], acc[o][t], xb, 0,
                                                          0, 0);
            if constexpr (REF) {
                // r = acc - xb L_bb^T (strict upper part of L_jj holds A: masked), xb += r Dinv^T
                v16f rr = acc[o];
#pragma unroll
                for (int t = 0; t < 16; ++t) {
                    const int q = sigma(t, h);
                    const float x = Ljj[(32 * b + n) + (int64_t)(32 * b + q) * lda];
                    const float a = q <= n ? -x : 0.f;
                    rr = __builtin_amdgcn_mfma_f32_32x32x2f32(a, xb[t], rr, 0, 0, 0);
                }
#pragma unroll
                for (int t = 0; t < 16; ++t)
                    xb = __builtin_amdgcn_mfma_f32_32x32x2f32(Db[n + sigma(t, h) * 32], rr[t], xb, 0,
                                                              0, 0);
            }
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                Xs[b][n + sigma(i, h) * XLD] = xb[i];
                ss = fmaf(xb[i], xb[i], ss);
                am = fmaxf(am, fabsf(xb[i]));
            }
            if (valid) {
#pragma unroll
                for (int i = 0; i < 16; ++i) A21[row + (int64_t)(32 * b + sigma(i, h)) * lda] = xb[i];
            }
        }
        if (b + 1 < NBK) {
            mpc_lds_barrier();
#pragma unroll
            for (int o = 0; o < NO; ++o) {
                const int bp = wave + NW * o;
                if (bp > b) {
#pragma unroll
                    for (int t = 0; t < 16; ++t)
                        acc[o] = __builtin_amdgcn_mfma_f32_32x32x2f32(
                            lv[o][t], Xs[b][n + sigma(t, h) * XLD], acc[o], 0, 0, 0);
                }
            }
        }
    }
    // row statistics: lane (n, h) of every wave holds part of row n
    red_ss[wave][lane] = ss;
    red_am[wave][lane] = am;
    mpc_lds_barrier();
    if (wave != 0) return;
    float tss = 0.f, tam = 0.f;
    if (lane < 32) {
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            tss += red_ss[w][lane] + red_ss[w][lane + 32];
            tam = fmaxf(tam, fmaxf(red_am[w][lane], red_am[w][lane + 32]));
        }
    }
    if (!valid || lane >= 32) { tss = 0.f; tam = 0.f; }
    if (valid && lane < 32) {
        rowss[row] = tss;
        tracker[row] -= tss;
    }
    // switching rule SCHUR: largest tracker / original diagonal below the panel
    if (diag0) {
        float q = 0.f;
        if (valid && h == 0 && wave == 0) {
            const float d0 = diag0[row];
            const float qq = tracker[row] / d0;
            q = qq > 0.f ? qq : 0.f;            // NaN and negative ratios count as 0
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) q = fmaxf(q, __shfl_xor(q, o));
        if (lane == 0 && wave == 0) atomicMax(&sm->schur_max[slot], __float_as_uint(q));
    }
    if (!(tss == tss) || !(tam <= 3.4e38f)) atomicOr(&sm->nan_flag, 1);
    // slab max (all 32 rows of the slab are in one tile: tile_rows % 32 == 0)
    float wss = tss, wam = tam;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        wss = fmaxf(wss, __shfl_xor(wss, o));
        wam = fmaxf(wam, __shfl_xor(wam, o));
    }
    if (lane == 0) {
        const int t = (int)(r0 / tile_rows);
        atomicMax(&t_maxrs[t], __float_as_uint(wss));
        atomicMax(&t_amax[t], __float_as_uint(wam));
    }
}

// tile amax for the exported quantiser (no TRSM in front of it)
__global__ __launch_bounds__(256) void tile_amax_kernel(const float* __restrict__ P, int64_t m,
                                                        int k, int64_t ld, int tile_rows,
                                                        uint32_t* __restrict__ tile_amax) {
    const int64_t r0 = (int64_t)blockIdx.x * 32;
    const int r = threadIdx.x & 31, qq = threadIdx.x >> 5;
    const int64_t row = r0 + r;
    float am = 0.f;
    if (row < m)
        for (int c = qq; c < k; c += 8) am = fmaxf(am, fabsf(P[row + (int64_t)c * ld]));
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) am = fmaxf(am, __shfl_xor(am, o));
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = am;
    __syncthreads();
    if (threadIdx.x == 0) {
        float v = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        atomicMax(&tile_amax[r0 / tile_rows], __float_as_uint(v));
    }
}

struct rule_args {
    int update_prec;
    int n_ladder;
    int ladder[4];
    double eps_prime;
    int rule;                 // MPC_RULE_*
};

// The lowest ladder format p with u_p * stat <= eps' * ref (stat and ref as
// the rule defines them, include/mpchol.h).  FP16 is stored unscaled: a tile
// with an entry above its largest finite value (65504) takes the next format.
__device__ __forceinline__ int decide(const rule_args& ra, float stat, float ref, float amax) {
    if (ra.update_prec >= 0) return ra.update_prec;
    for (int q = 0; q < ra.n_ladder; ++q) {
        if (ra.ladder[q] == MPC_PREC_FP16 && !(amax <= 65504.f)) continue;
        if (mpc_dev_unit_roundoff(ra.ladder[q]) * (double)stat <= ra.eps_prime * (double)ref)
            return ra.ladder[q];
    }
    return ra.ladder[ra.n_ladder - 1];
}

// One workgroup = 32 rows x k columns; thread (r, qq) owns row r, MX block qq.
__global__ __launch_bounds__(256) void quantize_kernel(
    const float* __restrict__ P, int64_t m, int k, int64_t ld, int tile_rows,
    const uint8_t* __restrict__ prec_in, const uint32_t* __restrict__ tile_amax,
    const uint32_t* __restrict__ tile_maxrs, const mpc_small* __restrict__ smc, rule_args ra,
    uint8_t* __restrict__ prec_out, uint8_t* __restrict__ map_col, uint8_t* __restrict__ payload,
    int64_t row_stride, uint8_t* __restrict__ scales, int* __restrict__ nan_flag, int64_t row0,
    int mu_slot, int map_rep, int64_t map_len, const float* __restrict__ rowss) {
    if (smc && smc->info != 0) return;
    const int64_t r0 = row0 + (int64_t)blockIdx.x * 32;
    const int r = threadIdx.x & 31, qq0 = threadIdx.x >> 5;
    const int64_t row = r0 + r;
    const int tile = (int)(r0 / tile_rows);
    int prec;
    if (prec_in) {
        prec = prec_in[tile];
    } else {
        float stat, ref;
        if (ra.rule == MPC_RULE_FROBENIUS) {
            // ||tile||_F^2: the tile's row sums of squares added in row order
            // (every workgroup of the tile forms the same sum)
            __shared__ float fr[256];
            const int64_t t0 = (int64_t)tile * tile_rows;
            const int nr = (int)(m - t0 < tile_rows ? m - t0 : tile_rows);
            float v = 0.f;
            for (int q = threadIdx.x; q < nr; q += 256) v += rowss[t0 + q];
            fr[threadIdx.x] = v;
            __syncthreads();
            for (int w = 128; w > 0; w >>= 1) {
                if ((int)threadIdx.x < w) fr[threadIdx.x] += fr[threadIdx.x + w];
                __syncthreads();
            }
            stat = fr[0];
            ref = smc->diag_f2[mu_slot];
        } else if (ra.rule == MPC_RULE_SCHUR) {
            stat = __uint_as_float(smc->schur_max[mu_slot]);
            ref = 1.f;
        } else {
            stat = __uint_as_float(tile_maxrs[tile]);
            ref = __uint_as_float(smc->mu_step[mu_slot]);
        }
        prec = decide(ra, stat, ref, __uint_as_float(tile_amax[tile]));
        if (r0 % tile_rows == 0 && threadIdx.x == 0) {
            if (prec_out) prec_out[tile] = (uint8_t)prec;
            // the map is in nb blocks: a tile of map_rep blocks fills each of them
            for (int u = 0; u < map_rep && map_col; ++u)
                if ((int64_t)tile * map_rep + u < map_len) map_col[tile * map_rep + u] = (uint8_t)prec;
        }
    }
    const int kb = k >> 5;
    const int esz = mpc_dev_elem_bytes(prec);
    if (row >= m) {
        // rows of the last 32-row block beyond m: zero slots in the 8-bit
        // layout, where a wave reads whole blocks
        if (esz == 1) {
            for (int qq = qq0; qq < kb; qq += 8) {
                *reinterpret_cast<uint4*>(payload + mpc_f8_offset(row, 32 * qq, row_stride)) =
                    make_uint4(0u, 0u, 0u, 0u);
                *reinterpret_cast<uint4*>(payload + mpc_f8_offset(row, 32 * qq + 16, row_stride)) =
                    make_uint4(0u, 0u, 0u, 0u);
            }
        }
        return;
    }
    const float tam = __uint_as_float(tile_amax[tile]);
    uint8_t* prow = payload + row * row_stride;
    bool bad = false;
    for (int qq = qq0; qq < kb; qq += 8) {
        float v[32];
        float am = 0.f;
#pragma unroll
        for (int c = 0; c < 32; ++c) {
            v[c] = P[row + (int64_t)(32 * qq + c) * ld];
            am = fmaxf(am, fabsf(v[c]));
        }
        if (!(am <= 3.4e38f)) bad = true;
        int e = 0;
        if (prec == MPC_PREC_MXE4M3 || prec == MPC_PREC_MXE5M2) e = mpc_dev_scale_exp(am, prec);
        else if (prec == MPC_PREC_E4M3 || prec == MPC_PREC_E5M2) e = mpc_dev_scale_exp(tam, prec);
        scales[row * kb + qq] = (uint8_t)(e + 127);
        uint8_t* dst = prow + (int64_t)32 * qq * esz;
        if (prec == MPC_PREC_FP32) {
#pragma unroll
            for (int c = 0; c < 32; c += 4)
                *reinterpret_cast<float4*>(dst + 4 * c) = make_float4(v[c], v[c + 1], v[c + 2], v[c + 3]);
        } else if (prec == MPC_PREC_BF16 || prec == MPC_PREC_FP16) {
            uint32_t w[16];
#pragma unroll
            for (int c = 0; c < 32; c += 2) {
                uint32_t lo = prec == MPC_PREC_BF16 ? mpc_dev_f32_to_bf16(v[c]) : mpc_dev_f32_to_f16(v[c]);
                uint32_t hi = prec == MPC_PREC_BF16 ? mpc_dev_f32_to_bf16(v[c + 1]) : mpc_dev_f32_to_f16(v[c + 1]);
                w[c >> 1] = lo | (hi << 16);
                if (prec == MPC_PREC_FP16 && ((lo & 0x7c00u) == 0x7c00u || (hi & 0x7c00u) == 0x7c00u)) bad = true;
            }
#pragma unroll
            for (int c = 0; c < 16; c += 4)
                *reinterpret_cast<uint4*>(dst + 4 * c) = make_uint4(w[c], w[c + 1], w[c + 2], w[c + 3]);
        } else {
            uint32_t w[8];
#pragma unroll
            for (int c = 0; c < 32; c += 4) {
                float a0 = ldexpf(v[c], -e), a1 = ldexpf(v[c + 1], -e);
                float a2 = ldexpf(v[c + 2], -e), a3 = ldexpf(v[c + 3], -e);
                w[c >> 2] = mpc_dev_pack4_f8(a0, a1, a2, a3, prec);
            }
            *reinterpret_cast<uint4*>(payload + mpc_f8_offset(row, 32 * qq, row_stride)) =
                make_uint4(w[0], w[1], w[2], w[3]);
            *reinterpret_cast<uint4*>(payload + mpc_f8_offset(row, 32 * qq + 16, row_stride)) =
                make_uint4(w[4], w[5], w[6], w[7]);
        }
    }
    if (bad) atomicOr(nan_flag, 1);
}

}  // namespace

int mpc_launch_trsm_stats(hipStream_t st, float* A21, int64_t lda, int64_t s, int nb,
                          const float* dinv_j, const float* Ljj, float* rowss, float* tracker,
                          int tile_rows, mpc_small* sm, uint32_t* tile_amax, uint32_t* tile_maxrs,
                          int64_t row0, int64_t row1, const float* diag0, int slot) {
    if (row1 < 0 || row1 > s) row1 = s;
    if (row1 <= row0) return 0;
    if (tile_rows % 32 || row0 % 32) {
        mpc_set_error("trsm: tile_rows %d / first row %lld not multiples of 32", tile_rows,
                      (long long)row0);
        return MPC_ERR_ARG;
    }
    // rows [row0, row1): the last slab's rows beyond row1 belong to the next
    // launch only when row1 == s (a slab never straddles row1 otherwise)
    if (row1 != s && row1 % 32) {
        mpc_set_error("trsm: row range end %lld not a multiple of 32", (long long)row1);
        return MPC_ERR_ARG;
    }
    dim3 grid((unsigned)((row1 - row0 + 31) / 32));
    switch (nb) {
    case 64:
        hipLaunchKernelGGL((trsm_stats_kernel<2, 2, kTrsmRefine>), grid, dim3(128), 0, st, A21,
                           lda, s, dinv_j, Ljj, rowss, tracker, tile_rows, sm, tile_amax,
                           tile_maxrs, row0, diag0, slot & 1);
        break;
    case 128:
        hipLaunchKernelGGL((trsm_stats_kernel<4, 4, kTrsmRefine>), grid, dim3(256), 0, st, A21,
                           lda, s, dinv_j, Ljj, rowss, tracker, tile_rows, sm, tile_amax,
                           tile_maxrs, row0, diag0, slot & 1);
        break;
    case 256:
        hipLaunchKernelGGL((trsm_stats_kernel<8, 4, kTrsmRefine>), grid, dim3(256), 0, st, A21,
                           lda, s, dinv_j, Ljj, rowss, tracker, tile_rows, sm, tile_amax,
                           tile_maxrs, row0, diag0, slot & 1);
        break;
    default:
        mpc_set_error("trsm: unsupported block size %d", nb);
        return MPC_ERR_ARG;
    }
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_tile_amax(hipStream_t st, const float* P, int64_t m, int k, int64_t ld,
                         int tile_rows, uint32_t* tile_amax) {
    if (m <= 0) return 0;
    hipLaunchKernelGGL(tile_amax_kernel, dim3((unsigned)((m + 31) / 32)), dim3(256), 0, st, P, m,
                       k, ld, tile_rows, tile_amax);
    MPC_LAUNCH_CHECK();
    return 0;
}

// factor path: decision from the tile statistics in sm; exported path: prec_in.
int mpc_launch_quantize_ex(hipStream_t st, const float* P, int64_t m, int k, int64_t ld,
                           int tile_rows, const uint8_t* prec_in, const uint32_t* tile_amax,
                           const uint32_t* tile_maxrs, mpc_small* sm, int update_prec,
                           const int* ladder, int n_ladder, double eps_prime, uint8_t* prec_out,
                           uint8_t* map_col, uint8_t* payload, int64_t row_stride, uint8_t* scales,
                           int* nan_flag, int64_t row0, int64_t row1, int mu_slot, int map_rep,
                           int64_t map_len, int rule, const float* rowss) {
    if (map_len < 0) map_len = ((m + tile_rows - 1) / tile_rows) * (int64_t)map_rep;
    if (row1 < 0 || row1 > m) row1 = m;
    if (row1 <= row0) return 0;
    if (k % 32 || tile_rows % 32 || row0 % 32 || (row1 != m && row1 % 32)) {
        mpc_set_error("quantize: k (%d), tile_rows (%d) and the row range [%lld, %lld) must be "
                      "multiples of 32", k, tile_rows, (long long)row0, (long long)row1);
        return MPC_ERR_ARG;
    }
    rule_args ra;
    ra.update_prec = update_prec;
    ra.n_ladder = n_ladder;
    for (int i = 0; i < 4; ++i) ra.ladder[i] = i < n_ladder ? ladder[i] : 0;
    ra.eps_prime = eps_prime;
    ra.rule = rule;
    if (!prec_in && update_prec < 0 && rule == MPC_RULE_FROBENIUS && !rowss) {
        mpc_set_error("quantize: the Frobenius rule needs the row sums of squares");
        return MPC_ERR_ARG;
    }
    hipLaunchKernelGGL(quantize_kernel, dim3((unsigned)((row1 - row0 + 31) / 32)), dim3(256), 0, st,
                       P, m, k, ld, tile_rows, prec_in, tile_amax, tile_maxrs, sm, ra, prec_out,
                       map_col, payload, row_stride, scales, nan_flag, row0, mu_slot & 1, map_rep,
                       map_len, rowss);
    MPC_LAUNCH_CHECK();
    return 0;
}
