// k_syrk.hip -- trailing-update SYRK, C(lower) -= Q Q^T, FP32 accumulation.
//
// Replaces the cuBLASLt launches of switching_chol.cuh:145-219
// (LtMxfp8Matmul fp8_utils.cuh:136-227, LtHSgemm_fp16in_fp32out :304-412,
// LtSgemm :231-302) and the WMMA SYRK of syrk_gemm.cuh:14-124, with the
// operand/target fix of fp8_chol.cuh:125-152 (C = A22, m = s) and only the
// lower triangle computed.
//
// Q is the quantised panel: row-major, k-contiguous payload plus one E8M0
// scale byte per row per 32 columns.  Each precision tile of Q (tile_rows
// rows) has its own storage format; an output tile pair is computed in the
// higher-precision class of its two operand tiles:
//   FP8 class  v_mfma_scale_f32_32x32x64_f8f6f4 (E4M3/E5M2 per operand, the
//              E8M0 scales applied by the matrix core: MX per 32, or the tile
//              scale replicated per 32)
//   16-bit     v_mfma_f32_32x32x16_bf16 / _f16 (FP8 operands converted exactly)
//   FP32       v_mfma_f32_32x32x2f32 (lower formats expanded exactly)
//
// Launch: one workgroup (4 waves) per 128x128 lower tile of C, wave (wr, wc)
// owns a 64x64 quadrant as 2x2 MFMA blocks.  The MFMA "row" index runs along
// C's columns and the "column" index along C's rows, so each accumulator
// register maps to 32 consecutive rows of the column-major C (128 B per
// half-wave).  Block ids are remapped so consecutive tiles share an XCD.
#include "mpc_internal.hpp"
#include "mpc_formats.hpp"

namespace {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef float v2f __attribute__((ext_vector_type(2)));

constexpr int TILE = 128;

__device__ __forceinline__ int pair_class(int p, int q) {
    if (p == MPC_PREC_FP32 || q == MPC_PREC_FP32) return MPC_PREC_FP32;
    if (p == MPC_PREC_BF16 || q == MPC_PREC_BF16) return MPC_PREC_BF16;
    if (p == MPC_PREC_FP16 || q == MPC_PREC_FP16) return MPC_PREC_FP16;
    return MPC_PREC_E4M3;   // both 8-bit
}

struct operand_src {
    const uint8_t* payload;
    const uint8_t* scales;
    int64_t row_stride;
    int kb;          // k / 32
    int prec;        // storage format of this operand tile
};

// 8 consecutive k values of one row as floats (stored format -> f32, exact)
__device__ __forceinline__ void load8_f32(const operand_src& o, int64_t row, int k0, float* f) {
    const uint8_t* p = o.payload + row * o.row_stride;
    if (o.prec == MPC_PREC_FP32) {
        float4 a = *reinterpret_cast<const float4*>(p + 4 * k0);
        float4 b = *reinterpret_cast<const float4*>(p + 4 * k0 + 16);
        f[0] = a.x; f[1] = a.y; f[2] = a.z; f[3] = a.w;
        f[4] = b.x; f[5] = b.y; f[6] = b.z; f[7] = b.w;
    } else if (o.prec == MPC_PREC_BF16 || o.prec == MPC_PREC_FP16) {
        uint4 a = *reinterpret_cast<const uint4*>(p + 2 * k0);
        uint32_t w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (o.prec == MPC_PREC_BF16) {
                f[2 * i] = __uint_as_float(w[i] << 16);
                f[2 * i + 1] = __uint_as_float(w[i] & 0xffff0000u);
            } else {
                f[2 * i] = (float)__builtin_bit_cast(_Float16, (unsigned short)(w[i] & 0xffffu));
                f[2 * i + 1] = (float)__builtin_bit_cast(_Float16, (unsigned short)(w[i] >> 16));
            }
        }
    } else {
        uint2 a = *reinterpret_cast<const uint2*>(p + k0);
        float sc = mpc_dev_exp2i((int)o.scales[row * o.kb + (k0 >> 5)] - 127);
        v2f t0, t1, t2, t3;
        if (mpc_dev_is_e4m3(o.prec)) {
            t0 = __builtin_amdgcn_cvt_pk_f32_fp8((int)a.x, false);
            t1 = __builtin_amdgcn_cvt_pk_f32_fp8((int)a.x, true);
            t2 = __builtin_amdgcn_cvt_pk_f32_fp8((int)a.y, false);
            t3 = __builtin_amdgcn_cvt_pk_f32_fp8((int)a.y, true);
        } else {
            t0 = __builtin_amdgcn_cvt_pk_f32_bf8((int)a.x, false);
            t1 = __builtin_amdgcn_cvt_pk_f32_bf8((int)a.x, true);
            t2 = __builtin_amdgcn_cvt_pk_f32_bf8((int)a.y, false);
            t3 = __builtin_amdgcn_cvt_pk_f32_bf8((int)a.y, true);
        }
        f[0] = t0.x * sc; f[1] = t0.y * sc; f[2] = t1.x * sc; f[3] = t1.y * sc;
        f[4] = t2.x * sc; f[5] = t2.y * sc; f[6] = t3.x * sc; f[7] = t3.y * sc;
    }
}

__device__ __forceinline__ v8bf load8_bf16(const operand_src& o, int64_t row, int k0) {
    if (o.prec == MPC_PREC_BF16) {
        uint4 a = *reinterpret_cast<const uint4*>(o.payload + row * o.row_stride + 2 * k0);
        return __builtin_bit_cast(v8bf, a);
    }
    float f[8];
    load8_f32(o, row, k0, f);      // FP8 * 2^e is exact in bf16 range
    uint4 w;
    w.x = (__float_as_uint(f[0]) >> 16) | (__float_as_uint(f[1]) & 0xffff0000u);
    w.y = (__float_as_uint(f[2]) >> 16) | (__float_as_uint(f[3]) & 0xffff0000u);
    w.z = (__float_as_uint(f[4]) >> 16) | (__float_as_uint(f[5]) & 0xffff0000u);
    w.w = (__float_as_uint(f[6]) >> 16) | (__float_as_uint(f[7]) & 0xffff0000u);
    return __builtin_bit_cast(v8bf, w);
}

__device__ __forceinline__ v8h load8_f16(const operand_src& o, int64_t row, int k0) {
    uint4 a = *reinterpret_cast<const uint4*>(o.payload + row * o.row_stride + 2 * k0);
    return __builtin_bit_cast(v8h, a);
}

struct tile_ctx {
    float* C;
    int64_t ldc, m;
    int64_t rowbase, colbase;   // first C row / col of this wave's 64x64 quadrant
    bool diag;                  // this quadrant straddles the diagonal (rowbase == colbase)
};

// Element (a, b, i) of the wave's quadrant sits at wave base + 32-bit offset
//   (32a + n) + (32b + 8(i/4) + 4h + i%4) * ldc,
// which stays below 2^31 for ldc <= 2^24; only the base pointer is 64-bit.
// The product Q Q^T is accumulated from zero and subtracted from C once, as a
// BLAS GEMM with beta = 1 does: seeding the accumulator with -C instead would
// round every one of the k FMA steps at the scale of |C| rather than of the
// partial sum, about sqrt(k) times the error for a trailing update.
// Addresses: element (a, b, i) = wave base + column offset (32b + 8(i/4) +
// i%4) * ldc, which is wave-uniform (scalar registers), + the lane's byte
// offset (32a + n + 4h ldc) * 4, one 32-bit VGPR per a.
__device__ __forceinline__ void update_c(const tile_ctx& t, const v16f acc[2][2]) {
    const int lane = threadIdx.x & 63, h = lane >> 5, n = lane & 31;
    const char* cbase = reinterpret_cast<const char*>(t.C + t.rowbase + t.colbase * t.ldc);
    char* wbase = reinterpret_cast<char*>(t.C + t.rowbase + t.colbase * t.ldc);
    const int64_t ld = t.ldc;
    const int rlim = (int)(t.m - t.rowbase), clim = (int)(t.m - t.colbase);
    uint32_t loff[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) loff[a] = (uint32_t)((32 * a + n) + 4 * h * ld) * 4u;
    // all 64 loads first (one HBM latency), then subtract and store.  Loads
    // are unconditional (a per-lane conditional load becomes a branch with its
    // own wait): interior quadrants read every element (the diagonal
    // quadrant's upper part is a valid address, masked at the store); edge
    // quadrants read from a clamped row and column.
    float v[2][2][16];
    const bool interior = rlim >= 64 && clim >= 64;
    if (interior) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const char* col = cbase + (int64_t)(32 * b + 8 * (i >> 2) + (i & 3)) * ld * 4;
                    v[a][b][i] = *reinterpret_cast<const float*>(col + loff[a]);
                }
    } else {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int r = 32 * a + n, c = 32 * b + 8 * (i >> 2) + 4 * h + (i & 3);
                    const int rc = r < rlim ? r : rlim - 1, cc = c < clim ? c : clim - 1;
                    v[a][b][i] = reinterpret_cast<const float*>(cbase)[rc + (int64_t)cc * ld];
                }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int r = 32 * a + n, c = 32 * b + 8 * (i >> 2) + 4 * h + (i & 3);
                const bool ok = r < rlim && c < clim && (!t.diag || r >= c);
                char* col = wbase + (int64_t)(32 * b + 8 * (i >> 2) + (i & 3)) * ld * 4;
                if (ok) *reinterpret_cast<float*>(col + loff[a]) = v[a][b][i] - acc[a][b][i];
            }
}

// FP8 class: A operand = column-side rows, B operand = row-side rows.
template <int FA, int FB>
__device__ __forceinline__ void mm_fp8(const operand_src& oc, const operand_src& orow,
                                       const tile_ctx& t, int k, v16f acc[2][2]) {
    const int lane = threadIdx.x & 63, h = lane >> 5, n = lane & 31;
    int64_t ra[2], rb[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        ra[q] = t.colbase + 32 * q + n;
        if (ra[q] >= t.m) ra[q] = t.m - 1;
        rb[q] = t.rowbase + 32 * q + n;
        if (rb[q] >= t.m) rb[q] = t.m - 1;
    }
    // Operand layout of the 8-bit 32x32x64 matrix-core op (measured on gfx950):
    // the scale of lane half h (E8M0 block h of the 64-wide K chunk) multiplies
    // bytes 0-15 of BOTH lane halves when h = 0 and bytes 16-31 of both when
    // h = 1.  So lane half h carries K [kc + 16h, +16) in bytes 0-15 and
    // K [kc + 32 + 16h, +16) in bytes 16-31; K block h then spans exactly the
    // 32 elements the hardware scales with lane half h's scale byte.
    for (int kc = 0; kc < k; kc += 64) {
        v8i A[2], B[2];
        int sa[2], sb[2];
        const int koff = kc + 32 * h;          // scale block of this lane half
        const int kl = kc + 16 * h;            // first K of bytes 0-15
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const uint8_t* pa = oc.payload + ra[q] * oc.row_stride + kl;
            const uint8_t* pb = orow.payload + rb[q] * orow.row_stride + kl;
            uint4 a0 = *reinterpret_cast<const uint4*>(pa), a1 = *reinterpret_cast<const uint4*>(pa + 32);
            uint4 b0 = *reinterpret_cast<const uint4*>(pb), b1 = *reinterpret_cast<const uint4*>(pb + 32);
            A[q] = v8i{(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w, (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
            B[q] = v8i{(int)b0.x, (int)b0.y, (int)b0.z, (int)b0.w, (int)b1.x, (int)b1.y, (int)b1.z, (int)b1.w};
            sa[q] = oc.scales[ra[q] * oc.kb + (koff >> 5)];
            sb[q] = orow.scales[rb[q] * orow.kb + (koff >> 5)];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                    A[b], B[a], acc[a][b], FA, FB, 0, sa[b], 0, sb[a]);
    }
}

template <bool BF>
__device__ __forceinline__ void mm_16(const operand_src& oc, const operand_src& orow,
                                      const tile_ctx& t, int k, v16f acc[2][2]) {
    const int lane = threadIdx.x & 63, h = lane >> 5, n = lane & 31;
    int64_t ra[2], rb[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        ra[q] = t.colbase + 32 * q + n;
        if (ra[q] >= t.m) ra[q] = t.m - 1;
        rb[q] = t.rowbase + 32 * q + n;
        if (rb[q] >= t.m) rb[q] = t.m - 1;
    }
    for (int kc = 0; kc < k; kc += 16) {
        const int k0 = kc + 8 * h;
        if constexpr (BF) {
            v8bf A[2], B[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) { A[q] = load8_bf16(oc, ra[q], k0); B[q] = load8_bf16(orow, rb[q], k0); }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[b], B[a], acc[a][b], 0, 0, 0);
        } else {
            v8h A[2], B[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) { A[q] = load8_f16(oc, ra[q], k0); B[q] = load8_f16(orow, rb[q], k0); }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(A[b], B[a], acc[a][b], 0, 0, 0);
        }
    }
}

__device__ __forceinline__ void mm_f32(const operand_src& oc, const operand_src& orow,
                                       const tile_ctx& t, int k, v16f acc[2][2]) {
    const int lane = threadIdx.x & 63, h = lane >> 5, n = lane & 31;
    int64_t ra[2], rb[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        ra[q] = t.colbase + 32 * q + n;
        if (ra[q] >= t.m) ra[q] = t.m - 1;
        rb[q] = t.rowbase + 32 * q + n;
        if (rb[q] >= t.m) rb[q] = t.m - 1;
    }
    // k permutation: half h takes the 8-element groups k = 16c + 8h .. +8
    for (int kc = 0; kc < k; kc += 16) {
        const int k0 = kc + 8 * h;
        float A[2][8], B[2][8];
#pragma unroll
        for (int q = 0; q < 2; ++q) { load8_f32(oc, ra[q], k0, A[q]); load8_f32(orow, rb[q], k0, B[q]); }
#pragma unroll
        for (int e = 0; e < 8; ++e)
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(A[b][e], B[a][e], acc[a][b], 0, 0, 0);
    }
}

template <int CLS, int FA, int FB>
__global__ __launch_bounds__(256) void syrk_kernel(float* __restrict__ C, int64_t m, int64_t ldc,
                                                   int k, const uint8_t* __restrict__ payload,
                                                   int64_t row_stride,
                                                   const uint8_t* __restrict__ scales,
                                                   int tile_rows,
                                                   const uint8_t* __restrict__ prec_of_tile,
                                                   const mpc_small* __restrict__ sm, int64_t npair,
                                                   int64_t per_xcd, int64_t mt, int64_t tcol0) {
    if (sm && sm->info != 0) return;
    // XCD-aware remap: hardware places block b on XCD b % 8
    const int64_t b = blockIdx.x;
    const int64_t pid = (b & 7) * per_xcd + (b >> 3);
    if (pid >= npair) return;
    // column-major over the lower triangle from tile column tcol0: column c
    // holds tiles ti = c .. mt-1; start(c) = c*mt - c*(c-1)/2
    auto start = [&](int64_t c) { return c * mt - c * (c - 1) / 2; };
    const int64_t gid = pid + start(tcol0);
    const double tw = 2.0 * (double)mt + 1.0;
    int64_t tj = (int64_t)((tw - sqrt(tw * tw - 8.0 * (double)gid)) * 0.5);
    if (tj < 0) tj = 0;
    while (tj > 0 && start(tj) > gid) --tj;
    while (start(tj + 1) <= gid) ++tj;
    const int64_t ti = tj + (gid - start(tj));

    operand_src orow, ocol;
    orow.prec = prec_of_tile[(ti * TILE) / tile_rows];
    ocol.prec = prec_of_tile[(tj * TILE) / tile_rows];
    if (pair_class(orow.prec, ocol.prec) != CLS) return;   // another class's launch
    if constexpr (CLS == MPC_PREC_E4M3) {                  // another format pair's launch
        if ((mpc_dev_is_e4m3(ocol.prec) ? 0 : 1) != FA || (mpc_dev_is_e4m3(orow.prec) ? 0 : 1) != FB)
            return;
    }

    const int wave = threadIdx.x >> 6, wr = wave >> 1, wc = wave & 1;
    const bool diag = ti == tj;
    if (diag && wr == 0 && wc == 1) return;   // strictly upper quadrant

    tile_ctx t;
    t.C = C;
    t.ldc = ldc;
    t.m = m;
    t.rowbase = ti * TILE + 64 * wr;
    t.colbase = tj * TILE + 64 * wc;
    t.diag = diag && wr == wc;     // only the diagonal quadrants need the r >= c mask
    if (t.rowbase >= m || t.colbase >= m) return;
    orow.payload = ocol.payload = payload;
    orow.scales = ocol.scales = scales;
    orow.row_stride = ocol.row_stride = row_stride;
    orow.kb = ocol.kb = k >> 5;

    v16f acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = v16f{};
    if constexpr (CLS == MPC_PREC_E4M3) {
        mm_fp8<FA, FB>(ocol, orow, t, k, acc);
    } else if constexpr (CLS == MPC_PREC_BF16) {
        mm_16<true>(ocol, orow, t, k, acc);
    } else if constexpr (CLS == MPC_PREC_FP16) {
        mm_16<false>(ocol, orow, t, k, acc);
    } else {
        mm_f32(ocol, orow, t, k, acc);
    }
    update_c(t, acc);
}

}  // namespace

// class_mask: bit c set if pairs of compute class c (MPC_PREC_FP32/BF16/FP16/E4M3)
// can occur; fp8_mask: bit 0 if E4M3-element tiles can occur, bit 1 if E5M2.
// One launch per class (and FP8 element-format pair); pairs belonging to
// another launch exit at once.
int mpc_launch_syrk_classes(hipStream_t st, float* C, int64_t m, int64_t ldc, int k,
                            const uint8_t* payload, int64_t row_stride, const uint8_t* scales,
                            int tile_rows, const uint8_t* prec_of_tile, const mpc_small* sm,
                            unsigned class_mask, unsigned fp8_mask, int64_t tcol0, int64_t tcol1) {
    if (m <= 0) return 0;
    if (k % 64 || tile_rows % TILE) {
        mpc_set_error("syrk: k (%d) must be a multiple of 64 and tile_rows (%d) of 128", k,
                      tile_rows);
        return MPC_ERR_ARG;
    }
    const int64_t mt = (m + TILE - 1) / TILE;
    if (tcol1 < 0 || tcol1 > mt) tcol1 = mt;
    if (tcol0 >= tcol1) return 0;
    auto start = [&](int64_t c) { return c * mt - c * (c - 1) / 2; };
    const int64_t npair = start(tcol1) - start(tcol0);
    const int64_t per_xcd = (npair + 7) / 8;
    dim3 grid((unsigned)(per_xcd * 8));
#define MPC_SYRK_LAUNCH(CLS, FA, FB)                                                             \
    hipLaunchKernelGGL((syrk_kernel<CLS, FA, FB>), grid, dim3(256), 0, st, C, m, ldc, k, payload, \
                       row_stride, scales, tile_rows, prec_of_tile, sm, npair, per_xcd, mt, tcol0)
    if (class_mask & (1u << MPC_PREC_E4M3)) {
        const bool e4 = fp8_mask & 1u, e5 = fp8_mask & 2u;
        if (e4) MPC_SYRK_LAUNCH(MPC_PREC_E4M3, 0, 0);
        if (e4 && e5) { MPC_SYRK_LAUNCH(MPC_PREC_E4M3, 0, 1); MPC_SYRK_LAUNCH(MPC_PREC_E4M3, 1, 0); }
        if (e5) MPC_SYRK_LAUNCH(MPC_PREC_E4M3, 1, 1);
    }
    if (class_mask & (1u << MPC_PREC_BF16)) MPC_SYRK_LAUNCH(MPC_PREC_BF16, 0, 0);
    if (class_mask & (1u << MPC_PREC_FP16)) MPC_SYRK_LAUNCH(MPC_PREC_FP16, 0, 0);
    if (class_mask & (1u << MPC_PREC_FP32)) MPC_SYRK_LAUNCH(MPC_PREC_FP32, 0, 0);
#undef MPC_SYRK_LAUNCH
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_syrk(hipStream_t st, float* C, int64_t m, int64_t ldc, int k,
                    const uint8_t* payload, int64_t row_stride, const uint8_t* scales,
                    int tile_rows, const uint8_t* prec_of_tile, const mpc_small* sm) {
    const unsigned all = (1u << MPC_PREC_E4M3) | (1u << MPC_PREC_BF16) | (1u << MPC_PREC_FP16) |
                         (1u << MPC_PREC_FP32);
    return mpc_launch_syrk_classes(st, C, m, ldc, k, payload, row_stride, scales, tile_rows,
                                   prec_of_tile, sm, all, 3u, 0, -1);
}
