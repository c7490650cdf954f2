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
//   FP32       v_mfma_f32_32x32x2f32 for FP32 x FP32 and FP32 x FP16 pairs;
//              FP32 x FP8/BF16 pairs as three exact bf16 products per chunk
//              (the FP32 operand split as hi + mid + lo)
//
// Launch: one workgroup (4 waves) per 128x128 lower tile of C, wave (wr, wc)
// owns a 64x64 quadrant as 2x2 MFMA blocks (64x64 tiles of 1x1 blocks for the
// look-ahead's block column).  The MFMA "row" index runs along
// C's columns and the "column" index along C's rows, so each accumulator
// register maps to 32 consecutive rows of the column-major C (128 B per
// half-wave).  Block ids are remapped so consecutive tiles share an XCD.
#include "mpc_internal.hpp"
#include "mpc_formats.hpp"
#include <type_traits>

namespace {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef float v2f __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int pair_class(int p, int q) {
    if (p == MPC_PREC_FP32 || q == MPC_PREC_FP32) return MPC_PREC_FP32;
    if (p == MPC_PREC_BF16 || q == MPC_PREC_BF16) return MPC_PREC_BF16;
    if (p == MPC_PREC_FP16 || q == MPC_PREC_FP16) return MPC_PREC_FP16;
    return MPC_PREC_E4M3;   // both 8-bit
}

// Position of a tile pair's launch in mpc_launch_syrk_panels' per-class
// sequence (E4M3 element pairs 00, 01, 10, 11; BF16; FP16; FP32 against
// FP32, against an FP8/BF16 format, against FP16).
__device__ __forceinline__ int launch_rank_of(int cls, int fa, int fb) {
    if (cls == MPC_PREC_E4M3) return 2 * fa + fb;
    if (cls == MPC_PREC_BF16) return 4;
    if (cls == MPC_PREC_FP16) return 5;
    return fa == 1 ? 6 : (fa == 2 ? 7 : 8);      // FP32: fa is the partner kind
}
__device__ __forceinline__ bool is_e4m3_elem(int p) {
    return p == MPC_PREC_E4M3 || p == MPC_PREC_MXE4M3;
}
__device__ __forceinline__ int launch_rank(int prow, int pcol) {
    const int cls = pair_class(prow, pcol);
    if (cls == MPC_PREC_E4M3)
        return launch_rank_of(cls, is_e4m3_elem(pcol) ? 0 : 1, is_e4m3_elem(prow) ? 0 : 1);
    if (cls == MPC_PREC_FP32) {
        const bool f32pair = prow == MPC_PREC_FP32 && pcol == MPC_PREC_FP32;
        const bool f16 = prow == MPC_PREC_FP16 || pcol == MPC_PREC_FP16;
        return launch_rank_of(cls, f32pair ? 1 : (f16 ? 0 : 2), 0);
    }
    return launch_rank_of(cls, 0, 0);
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
        uint2 a = *reinterpret_cast<const uint2*>(o.payload + mpc_f8_offset(row, k0, o.row_stride));
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

// x = hi + mid + lo exactly, each term a bf16 (8 significant bits, taken by
// truncation): the remainders x - hi and (x - hi) - mid are exact in FP32 and
// the last one has at most 8 significant bits.  Eight values, packed as the
// 16-bit matrix-core operand.
__device__ __forceinline__ void split3_bf16(const float* f, v8bf& hi, v8bf& mid, v8bf& lo) {
    uint32_t uh[8], um[8], ul[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        uh[e] = __float_as_uint(f[e]) & 0xffff0000u;
        const float r1 = f[e] - __uint_as_float(uh[e]);
        um[e] = __float_as_uint(r1) & 0xffff0000u;
        const float r2 = r1 - __uint_as_float(um[e]);
        ul[e] = __float_as_uint(r2) & 0xffff0000u;
    }
    uint4 wh, wm, wl;
    wh.x = (uh[0] >> 16) | uh[1]; wh.y = (uh[2] >> 16) | uh[3]; wh.z = (uh[4] >> 16) | uh[5]; wh.w = (uh[6] >> 16) | uh[7];
    wm.x = (um[0] >> 16) | um[1]; wm.y = (um[2] >> 16) | um[3]; wm.z = (um[4] >> 16) | um[5]; wm.w = (um[6] >> 16) | um[7];
    wl.x = (ul[0] >> 16) | ul[1]; wl.y = (ul[2] >> 16) | ul[3]; wl.z = (ul[4] >> 16) | ul[5]; wl.w = (ul[6] >> 16) | ul[7];
    hi = __builtin_bit_cast(v8bf, wh);
    mid = __builtin_bit_cast(v8bf, wm);
    lo = __builtin_bit_cast(v8bf, wl);
}

struct tile_ctx {
    float* C;
    const float* Cs;            // where C's old values are read (C, or a source copy)
    int64_t ldc, m;
    int64_t rowbase, colbase;   // first C row / col of this wave's quadrant
    bool diag;                  // this quadrant straddles the diagonal (rowbase == colbase)
};

// Each wave owns a (32 QA) x (32 QA) quadrant of C as QA x QA MFMA blocks.
// Element (a, b, i) of the wave's quadrant sits at wave base + byte offset
//   ((32a + n) + (32b + 8(i/4) + 4h + i%4) * ldc) * 4,
// which stays below 2^31 for ldc <= 2^23; only the base pointer is 64-bit.
// The product Q Q^T is accumulated from zero and subtracted from C once, as a
// BLAS GEMM with beta = 1 does: seeding the accumulator with -C instead would
// round every one of the k FMA steps at the scale of |C| rather than of the
// partial sum, about sqrt(k) times the error for a trailing update.
//
// C's old values are loaded (load_c) before the K loop, so their HBM latency
// overlaps the operand loads and the products; store_c subtracts and
// stores.  C is streamed: its loads and stores carry the non-temporal cache
// policy (CPOL), so the lines of C, touched once per update, do not evict
// the panel rows that the other tiles of the XCD re-read from L2.
constexpr int kCpolStream = 2;     // nt: non-temporal (streaming) policy bit

template <int QA>
struct c_frag {
    float v[QA][QA][16];
};

// Buffer addressing: the wave's base in the resource (scalar), the column
// offset (32b + 8(i/4) + i%4) * ldc * 4 as the scalar soffset and the lane's
// (32a + n + 4h ldc) * 4 as the one 32-bit VGPR per a.  With global
// addressing every load would keep its own 64-bit VGPR address alive until
// the matching store.
template <int QA>
__device__ __forceinline__ void load_c(const tile_ctx& t, c_frag<QA>& c) {
    constexpr int QE = 32 * QA;
    const int lane = threadIdx.x & 63, h = lane >> 5, n = lane & 31;
    const __amdgpu_buffer_rsrc_t rsl = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float*>(t.Cs) + t.rowbase + t.colbase * t.ldc, 0, 0x7fffffff, 0x00020000);
    const int ld4 = (int)t.ldc * 4;
    const int rlim = (int)(t.m - t.rowbase), clim = (int)(t.m - t.colbase);
    // Loads are unconditional (a per-lane conditional load becomes a branch
    // with its own wait): interior quadrants read every element (the
    // diagonal quadrant's upper part is a valid address, masked at the
    // store); edge quadrants read from a clamped row and column.
    const bool interior = rlim >= QE && clim >= QE;
    if (interior) {
#pragma unroll
        for (int a = 0; a < QA; ++a) {
            const int loff = ((32 * a + n) + 4 * h * (int)t.ldc) * 4;
#pragma unroll
            for (int b = 0; b < QA; ++b)
#pragma unroll
                for (int i = 0; i < 16; ++i)
                    c.v[a][b][i] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                        rsl, loff, (32 * b + 8 * (i >> 2) + (i & 3)) * ld4, kCpolStream));
        }
    } else {
#pragma unroll
        for (int a = 0; a < QA; ++a)
#pragma unroll
            for (int b = 0; b < QA; ++b)
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int r = 32 * a + n, cc0 = 32 * b + 8 * (i >> 2) + 4 * h + (i & 3);
                    const int rc = r < rlim ? r : rlim - 1, cc = cc0 < clim ? cc0 : clim - 1;
                    c.v[a][b][i] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                        rsl, rc * 4 + cc * ld4, 0, kCpolStream));
                }
    }
}

template <int QA>
__device__ __forceinline__ void store_c(const tile_ctx& t, const c_frag<QA>& c,
                                        const v16f acc[QA][QA]) {
    const int lane = threadIdx.x & 63, h = lane >> 5, n = lane & 31;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(t.C + t.rowbase + t.colbase * t.ldc, 0, 0x7fffffff,
                                          0x00020000);
    const int ld4 = (int)t.ldc * 4;
    const int rlim = (int)(t.m - t.rowbase), clim = (int)(t.m - t.colbase);
#pragma unroll
    for (int a = 0; a < QA; ++a) {
        const int loff = ((32 * a + n) + 4 * h * (int)t.ldc) * 4;
#pragma unroll
        for (int b = 0; b < QA; ++b)
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int r = 32 * a + n, cc = 32 * b + 8 * (i >> 2) + 4 * h + (i & 3);
                const bool ok = r < rlim && cc < clim && (!t.diag || r >= cc);
                if (ok)
                    __builtin_amdgcn_raw_buffer_store_b32(
                        __float_as_uint(c.v[a][b][i] - acc[a][b][i]), rs, loff,
                        (32 * b + 8 * (i >> 2) + (i & 3)) * ld4, kCpolStream);
            }
    }
}

// operand rows of the wave: column side (ra) and row side (rb), clamped to m
template <int QA>
__device__ __forceinline__ void operand_rows(const tile_ctx& t, int64_t ra[QA], int64_t rb[QA]) {
    const int n = threadIdx.x & 31;
#pragma unroll
    for (int q = 0; q < QA; ++q) {
        ra[q] = t.colbase + 32 * q + n;
        if (ra[q] >= t.m) ra[q] = t.m - 1;
        rb[q] = t.rowbase + 32 * q + n;
        if (rb[q] >= t.m) rb[q] = t.m - 1;
    }
}

// body(kc) for kc = 0, STEP, 2 STEP, ... < k: fully unrolled when the extent
// KC is a compile-time constant (every chunk's loads can then be issued
// before the first product), a plain loop otherwise
template <int KC, int STEP, typename F>
__device__ __forceinline__ void k_loop(int k, F&& body) {
    if constexpr (KC > 0) {
#pragma unroll
        for (int kc = 0; kc < KC; kc += STEP) body(kc);
    } else {
        for (int kc = 0; kc < k; kc += STEP) body(kc);
    }
}

// K chunks of the 8-bit products kept in flight ahead of the chunk being
// multiplied (register buffers; the runtime-K loop is otherwise bound by one
// operand-load latency per 64-wide chunk)
constexpr int kFp8Ahead = 0;

template <int QA>
struct f8_frag {
    v8i A[QA], B[QA];
    int sa[QA], sb[QA];
};

// FP8 class: A operand = column-side rows, B operand = row-side rows.
template <int QA, int FA, int FB, int KC>
__device__ __forceinline__ void mm_fp8(const operand_src& oc, const operand_src& orow,
                                       const tile_ctx& t, int k, v16f acc[QA][QA]) {
    const int h = (threadIdx.x & 63) >> 5;
    int64_t ra[QA], rb[QA];
    operand_rows<QA>(t, ra, rb);
    // Operand layout of the 8-bit 32x32x64 matrix-core op (measured on gfx950):
    // the scale of lane half h (E8M0 block h of the 64-wide K chunk) multiplies
    // bytes 0-15 of BOTH lane halves when h = 0 and bytes 16-31 of both when
    // h = 1.  So lane half h carries K [kc + 16h, +16) in bytes 0-15 and
    // K [kc + 32 + 16h, +16) in bytes 16-31; K block h then spans exactly the
    // 32 elements the hardware scales with lane half h's scale byte.
    // the 32-row operand blocks of the wave (the lane's own 32 bytes of a
    // chunk: mpc_f8_offset's layout)
    const int64_t lane32 = (int64_t)(threadIdx.x & 63) * 32;
    const uint8_t* ba[QA];
    const uint8_t* bb[QA];
#pragma unroll
    for (int q = 0; q < QA; ++q) {
        ba[q] = oc.payload + (t.colbase + 32 * q) * oc.row_stride + lane32;
        bb[q] = orow.payload + (t.rowbase + 32 * q) * orow.row_stride + lane32;
    }
    auto load = [&](int kc, f8_frag<QA>& f) {
        const int koff = kc + 32 * h;          // scale block of this lane half
        const int64_t co = (int64_t)(kc >> 6) * 2048;
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const uint8_t* pa = ba[q] + co;
            const uint8_t* pb = bb[q] + co;
            uint4 a0 = *reinterpret_cast<const uint4*>(pa), a1 = *reinterpret_cast<const uint4*>(pa + 16);
            uint4 b0 = *reinterpret_cast<const uint4*>(pb), b1 = *reinterpret_cast<const uint4*>(pb + 16);
            f.A[q] = v8i{(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w, (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
            f.B[q] = v8i{(int)b0.x, (int)b0.y, (int)b0.z, (int)b0.w, (int)b1.x, (int)b1.y, (int)b1.z, (int)b1.w};
            f.sa[q] = oc.scales[ra[q] * oc.kb + (koff >> 5)];
            f.sb[q] = orow.scales[rb[q] * orow.kb + (koff >> 5)];
        }
    };
    auto mma = [&](const f8_frag<QA>& f) {
#pragma unroll
        for (int a = 0; a < QA; ++a)
#pragma unroll
            for (int b = 0; b < QA; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                    f.A[b], f.B[a], acc[a][b], FA, FB, 0, f.sa[b], 0, f.sb[a]);
    };
    if constexpr (KC > 0) {
        k_loop<KC, 64>(k, [&](int kc) {
            f8_frag<QA> f;
            load(kc, f);
            mma(f);
        });
    } else {
        // rolling buffer of kFp8Ahead + 1 chunks; past the end, the last
        // chunk is loaded again (branch-free) and never multiplied
        constexpr int NB = kFp8Ahead + 1;
        f8_frag<QA> buf[NB];
        const int klast = k - 64;
#pragma unroll
        for (int i = 0; i < kFp8Ahead; ++i) load(64 * i < klast ? 64 * i : klast, buf[i]);
        for (int kc = 0; kc < k; kc += 64 * NB) {
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const int c = kc + 64 * i;                       // chunk multiplied now
                const int cn = c + 64 * kFp8Ahead;               // chunk loaded now
                load(cn < klast ? cn : klast, buf[(i + kFp8Ahead) % NB]);
                if (c < k) mma(buf[i]);
            }
        }
    }
}

template <int QA, bool BF, int KC>
__device__ __forceinline__ void mm_16(const operand_src& oc, const operand_src& orow,
                                      const tile_ctx& t, int k, v16f acc[QA][QA]) {
    const int h = (threadIdx.x & 63) >> 5;
    int64_t ra[QA], rb[QA];
    operand_rows<QA>(t, ra, rb);
    const int native = BF ? MPC_PREC_BF16 : MPC_PREC_FP16;
    auto body = [&](int kc, auto both_native) {
        const int k0 = kc + 8 * h;
        if constexpr (BF) {
            v8bf A[QA], B[QA];
#pragma unroll
            for (int q = 0; q < QA; ++q) {
                if constexpr (decltype(both_native)::value) {
                    A[q] = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4*>(
                                                        oc.payload + ra[q] * oc.row_stride + 2 * k0));
                    B[q] = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4*>(
                                                        orow.payload + rb[q] * orow.row_stride + 2 * k0));
                } else {
                    A[q] = load8_bf16(oc, ra[q], k0);
                    B[q] = load8_bf16(orow, rb[q], k0);
                }
            }
#pragma unroll
            for (int a = 0; a < QA; ++a)
#pragma unroll
                for (int b = 0; b < QA; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[b], B[a], acc[a][b], 0, 0, 0);
        } else {
            v8h A[QA], B[QA];
#pragma unroll
            for (int q = 0; q < QA; ++q) { A[q] = load8_f16(oc, ra[q], k0); B[q] = load8_f16(orow, rb[q], k0); }
#pragma unroll
            for (int a = 0; a < QA; ++a)
#pragma unroll
                for (int b = 0; b < QA; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(A[b], B[a], acc[a][b], 0, 0, 0);
        }
    };
    // both operands stored in the class's own format: no per-load format branch
    if (KC > 0 && oc.prec == native && orow.prec == native)
        k_loop<KC, 16>(k, [&](int kc) { body(kc, std::true_type{}); });
    else
        k_loop<0, 16>(k, [&](int kc) { body(kc, std::false_type{}); });
}

// NATIVE (runtime-K form only): both operands stored in FP32
// SPLIT: 0 never, 1 always (the launch holds only such pairs), 2 per pair
template <int QA, int KC, bool NATIVE, int SPLIT>
__device__ __forceinline__ void mm_f32(const operand_src& oc, const operand_src& orow,
                                       const tile_ctx& t, int k, v16f acc[QA][QA]) {
    const int h = (threadIdx.x & 63) >> 5;
    int64_t ra[QA], rb[QA];
    operand_rows<QA>(t, ra, rb);
    // FP32 against a format exact in bf16 (FP8 with its scale, or BF16): the
    // FP32 operand as hi + mid + lo (exact), three bf16 products per 16-wide
    // chunk; the products are exact and the sums FP32, as in the FP32
    // matrix-core op, at a sixth of its cost per k
    if (SPLIT == 1 || (SPLIT == 2 && (oc.prec == MPC_PREC_FP32) != (orow.prec == MPC_PREC_FP32) &&
                       oc.prec != MPC_PREC_FP16 && orow.prec != MPC_PREC_FP16)) {
        const bool col32 = oc.prec == MPC_PREC_FP32;
        k_loop<0, 16>(k, [&](int kc) {
            const int k0 = kc + 8 * h;
            v8bf F[QA][3], Lw[QA];
#pragma unroll
            for (int q = 0; q < QA; ++q) {
                float f[8];
                load8_f32(col32 ? oc : orow, col32 ? ra[q] : rb[q], k0, f);
                split3_bf16(f, F[q][0], F[q][1], F[q][2]);
                Lw[q] = load8_bf16(col32 ? orow : oc, col32 ? rb[q] : ra[q], k0);
            }
#pragma unroll
            for (int s3 = 0; s3 < 3; ++s3)
#pragma unroll
                for (int a = 0; a < QA; ++a)
#pragma unroll
                    for (int b = 0; b < QA; ++b)
                        acc[a][b] = col32
                            ? __builtin_amdgcn_mfma_f32_32x32x16_bf16(F[b][s3], Lw[a], acc[a][b], 0, 0, 0)
                            : __builtin_amdgcn_mfma_f32_32x32x16_bf16(Lw[b], F[a][s3], acc[a][b], 0, 0, 0);
        });
        return;
    }
    // k permutation: half h takes the 8-element groups k = 16c + 8h .. +8
    if constexpr (KC > 0) {
        // unrolled (small launches): branch-free loads for the all-FP32 pair
        auto body = [&](int kc, auto both_native) {
            const int k0 = kc + 8 * h;
            float A[QA][8], B[QA][8];
#pragma unroll
            for (int q = 0; q < QA; ++q) {
                if constexpr (decltype(both_native)::value) {
                    const float* pa = reinterpret_cast<const float*>(oc.payload + ra[q] * oc.row_stride) + k0;
                    const float* pb = reinterpret_cast<const float*>(orow.payload + rb[q] * orow.row_stride) + k0;
                    const float4 a0 = *reinterpret_cast<const float4*>(pa), a1 = *reinterpret_cast<const float4*>(pa + 4);
                    const float4 b0 = *reinterpret_cast<const float4*>(pb), b1 = *reinterpret_cast<const float4*>(pb + 4);
                    A[q][0] = a0.x; A[q][1] = a0.y; A[q][2] = a0.z; A[q][3] = a0.w;
                    A[q][4] = a1.x; A[q][5] = a1.y; A[q][6] = a1.z; A[q][7] = a1.w;
                    B[q][0] = b0.x; B[q][1] = b0.y; B[q][2] = b0.z; B[q][3] = b0.w;
                    B[q][4] = b1.x; B[q][5] = b1.y; B[q][6] = b1.z; B[q][7] = b1.w;
                } else {
                    load8_f32(oc, ra[q], k0, A[q]);
                    load8_f32(orow, rb[q], k0, B[q]);
                }
            }
#pragma unroll
            for (int e = 0; e < 8; ++e)
#pragma unroll
                for (int a = 0; a < QA; ++a)
#pragma unroll
                    for (int b = 0; b < QA; ++b)
                        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(A[b][e], B[a][e], acc[a][b], 0, 0, 0);
        };
        if (oc.prec == MPC_PREC_FP32 && orow.prec == MPC_PREC_FP32)
            k_loop<KC, 16>(k, [&](int kc) { body(kc, std::true_type{}); });
        else
            k_loop<KC, 16>(k, [&](int kc) { body(kc, std::false_type{}); });
    } else if constexpr (NATIVE) {
        // both operands FP32: one branch-free loop, the next chunk's operands
        // loaded before this chunk's products (the FP32 class is MFMA-bound
        // only if its loads overlap the products)
        const float* pa[QA];
        const float* pb[QA];
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            pa[q] = reinterpret_cast<const float*>(oc.payload + ra[q] * oc.row_stride) + 8 * h;
            pb[q] = reinterpret_cast<const float*>(orow.payload + rb[q] * orow.row_stride) + 8 * h;
        }
        float4 A[QA][2], B[QA][2];
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            A[q][0] = *reinterpret_cast<const float4*>(pa[q]);
            A[q][1] = *reinterpret_cast<const float4*>(pa[q] + 4);
            B[q][0] = *reinterpret_cast<const float4*>(pb[q]);
            B[q][1] = *reinterpret_cast<const float4*>(pb[q] + 4);
        }
        for (int kc = 0; kc < k; kc += 16) {
            const int kn = kc + 16 < k ? kc + 16 : kc;       // the last chunk reloads itself
            float4 An[QA][2], Bn[QA][2];
#pragma unroll
            for (int q = 0; q < QA; ++q) {
                An[q][0] = *reinterpret_cast<const float4*>(pa[q] + kn);
                An[q][1] = *reinterpret_cast<const float4*>(pa[q] + kn + 4);
                Bn[q][0] = *reinterpret_cast<const float4*>(pb[q] + kn);
                Bn[q][1] = *reinterpret_cast<const float4*>(pb[q] + kn + 4);
            }
#pragma unroll
            for (int e = 0; e < 8; ++e)
#pragma unroll
                for (int a = 0; a < QA; ++a)
#pragma unroll
                    for (int b = 0; b < QA; ++b) {
                        const float av = e < 4 ? A[b][0][e] : A[b][1][e - 4];
                        const float bv = e < 4 ? B[a][0][e] : B[a][1][e - 4];
                        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc[a][b], 0, 0, 0);
                    }
#pragma unroll
            for (int q = 0; q < QA; ++q) {
                A[q][0] = An[q][0]; A[q][1] = An[q][1];
                B[q][0] = Bn[q][0]; B[q][1] = Bn[q][1];
            }
        }
    } else {
        // an FP32 operand against a lower format (rare): expanded per chunk
        for (int kc = 0; kc < k; kc += 16) {
            const int k0 = kc + 8 * h;
            float A[QA][8], B[QA][8];
#pragma unroll
            for (int q = 0; q < QA; ++q) { load8_f32(oc, ra[q], k0, A[q]); load8_f32(orow, rb[q], k0, B[q]); }
#pragma unroll
            for (int e = 0; e < 8; ++e)
#pragma unroll
                for (int a = 0; a < QA; ++a)
#pragma unroll
                    for (int b = 0; b < QA; ++b)
                        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(A[b][e], B[a][e], acc[a][b], 0, 0, 0);
        }
    }
}

// The products of the P.np panels are summed in the accumulators (K = np*k)
// and subtracted from C once.  CLS >= 0: only the panels whose pair on this
// tile has that compute class (and, for the FP8 class, that element-format
// pair FA/FB) are used, and a tile with none exits without touching C; a
// tile mixing classes across panels is updated once per class launch.
// CLS < 0: every panel, the class chosen per workgroup and panel.
// Work lists of a launch sequence (int32): 16 counters indexed by launch
// rank (launch_rank_of), then one list of tile pair ids per rank at
// 16 + rank * cap.  The first launch of a sequence runs over every tile pair
// and appends each pair to the lists of the later launches that have work on
// it; those launches then run over their lists only, instead of a full grid
// of workgroups of which most would find nothing of their class.
struct syrk_lists {
    int* base;       // nullptr: every launch runs over the full grid
    int  cap;        // tile pairs per list
    int  mode;       // 0 full grid, 1 full grid + build the later lists, 2 from the list
};
constexpr int kListHead = 16;

// The tile pairs of a launch: the lower triangle of tile columns
// [tcol0, tcol0 + ...) (column-major), or, when rect_ncol > 0, the rectangle
// of tile rows [rect_tr0, mt) and tile columns [tcol0, tcol0 + rect_ncol)
// (row-major; every pair strictly below the diagonal).
struct syrk_shape {
    int64_t mt, tcol0, tcol1, rect_tr0, rect_ncol;
};

template <int QA, int CLS, int FA, int FB, int KC>
__device__ __forceinline__ void syrk_tile(float* __restrict__ C, int64_t m, int64_t ldc, int k,
                                          const mpc_syrk_panels& P, int64_t row_stride,
                                          int tile_rows, const syrk_shape& G, int64_t pid,
                                          const syrk_lists& L);

// LISTED: a later launch of a sequence, over its work list (L.mode 2);
// otherwise one workgroup per tile pair (L.mode 0, or 1 to build the lists)
template <int QA, int CLS, int FA, int FB, int KC, bool LISTED>
__global__ __launch_bounds__(256, 2) void syrk_kernel(float* __restrict__ C, int64_t m, int64_t ldc,
                                                   int k, const mpc_syrk_panels P,
                                                   int64_t row_stride, int tile_rows,
                                                   const mpc_small* __restrict__ sm, int64_t npair,
                                                   int64_t per_xcd, const syrk_shape G,
                                                   const syrk_lists L) {
    if (sm && sm->info != 0) return;
    if constexpr (LISTED) {
        const int rank = launch_rank_of(CLS, FA, FB);
        const int count = L.base[rank];
        const int* list = L.base + kListHead + (int64_t)rank * L.cap;
        for (int it = blockIdx.x; it < count; it += gridDim.x)
            syrk_tile<QA, CLS, FA, FB, KC>(C, m, ldc, k, P, row_stride, tile_rows, G, list[it], L);
        return;
    }
    // XCD-aware remap: hardware places block b on XCD b % 8
    const int64_t b = blockIdx.x;
    const int64_t pid = (b & 7) * per_xcd + (b >> 3);
    if (pid >= npair) return;
    syrk_tile<QA, CLS, FA, FB, KC>(C, m, ldc, k, P, row_stride, tile_rows, G, pid, L);
}

template <int QA, int CLS, int FA, int FB, int KC>
__device__ __forceinline__ void syrk_tile(float* __restrict__ C, int64_t m, int64_t ldc, int k,
                                          const mpc_syrk_panels& P, int64_t row_stride,
                                          int tile_rows, const syrk_shape& G, int64_t pid,
                                          const syrk_lists& L) {
    constexpr int TILE = 64 * QA;
    int64_t ti, tj;
    if (G.rect_ncol > 0) {
        ti = G.rect_tr0 + pid / G.rect_ncol;
        tj = G.tcol0 + pid % G.rect_ncol;
    } else {
        // The lower triangle from tile column tcol0 in super-blocks of SB x SB
        // tiles (2048 rows and columns): super-columns in order, in each the
        // diagonal super-block and then the super-blocks below it, each
        // column-major.  Consecutive pair ids go to one XCD, whose L2 then
        // keeps the panel rows of the few super-blocks it works on (a plain
        // column-major order streams the whole panel past every tile column).
        // start(c) = c*mt - c*(c-1)/2 counts the pairs left of tile column c.
        constexpr int64_t SB = 2048 / TILE;
        const int64_t mt = G.mt;
        auto start = [&](int64_t c) { return c * mt - c * (c - 1) / 2; };
        const int64_t gid = pid + start(G.tcol0);
        const double tw = 2.0 * (double)mt + 1.0;
        int64_t tc = (int64_t)((tw - sqrt(tw * tw - 8.0 * (double)gid)) * 0.5);
        if (tc < 0) tc = 0;
        while (tc > 0 && start(tc) > gid) --tc;
        while (start(tc + 1) <= gid) ++tc;
        // the super-column holding tile column tc: columns [c0, c1)
        const int64_t c0 = G.tcol0 + ((tc - G.tcol0) / SB) * SB;
        const int64_t c1 = c0 + SB < G.tcol1 ? c0 + SB : G.tcol1;
        const int64_t rb = c0 + SB < mt ? c0 + SB : mt;       // rows of its diagonal block
        const int64_t loc = gid - start(c0);
        // pairs in the diagonal block: sum over tj in [c0, c1) of (rb - tj)
        const int64_t w = c1 - c0;
        const int64_t D = w * rb - (c0 + c1 - 1) * w / 2;
        if (loc < D) {
            // column-major over the triangle: column c0 + u holds the rows
            // [c0 + u, rb), M - u of them (M = rb - c0); dstart(u) = u M - u(u-1)/2
            const int64_t M = rb - c0;
            auto dstart = [&](int64_t u) { return u * M - u * (u - 1) / 2; };
            const double tw2 = 2.0 * (double)M + 1.0;
            int64_t u = (int64_t)((tw2 - sqrt(fmax(tw2 * tw2 - 8.0 * (double)loc, 0.0))) * 0.5);
            if (u < 0) u = 0;
            if (u > w - 1) u = w - 1;
            while (u > 0 && dstart(u) > loc) --u;
            while (u + 1 < w && dstart(u + 1) <= loc) ++u;
            tj = c0 + u;
            ti = tj + (loc - dstart(u));
        } else {
            // rows [rb, mt) as SB-row super-blocks, each column-major (the
            // tiles of one column, consecutive in memory, on one XCD together)
            const int64_t l2 = loc - D;
            const int64_t sbi = l2 / (SB * w), inner = l2 - sbi * SB * w;
            const int64_t rbase = rb + sbi * SB;
            const int64_t rows = mt - rbase < SB ? mt - rbase : SB;
            tj = c0 + inner / rows;
            ti = rbase + inner % rows;
        }
    }
    const int64_t trow = (ti * TILE) / tile_rows, tcol = (tj * TILE) / tile_rows;

    if (L.mode == 1 && threadIdx.x == 0) {
        // the later launches of the sequence with work on this tile pair
        unsigned need = 0;
        for (int q = 0; q < P.np; ++q) {
            const int slot = (P.p0 + q) % P.slots;
            const uint8_t* prec = P.prec + slot * 1024 + (P.rows0 - (int64_t)q * P.nb) / tile_rows;
            need |= 1u << launch_rank(prec[trow], prec[tcol]);
        }
        need &= ~((2u << launch_rank_of(CLS, FA, FB)) - 1u);
        for (int r = 0; r < kListHead; ++r)
            if ((need >> r) & 1u) {
                const int slot = atomicAdd(&L.base[r], 1);
                L.base[kListHead + (int64_t)r * L.cap + slot] = (int)pid;
            }
    }

    const int wave = threadIdx.x >> 6, wr = wave >> 1, wc = wave & 1;
    const bool diag = ti == tj;
    if (diag && wr == 0 && wc == 1) return;   // strictly upper quadrant

    tile_ctx t;
    t.C = C;
    t.Cs = C;
    t.ldc = ldc;
    t.m = m;
    t.rowbase = ti * TILE + 32 * QA * wr;
    t.colbase = tj * TILE + 32 * QA * wc;
    t.diag = diag && wr == wc;     // only the diagonal quadrants need the r >= c mask
    if (t.rowbase >= m || t.colbase >= m) return;

    v16f acc[QA][QA];
#pragma unroll
    for (int a = 0; a < QA; ++a)
#pragma unroll
        for (int b2 = 0; b2 < QA; ++b2) acc[a][b2] = v16f{};
    bool any = false;
    if (P.csrc) {
        // C's old values come from csrc in the first launch, in the host's
        // launch order, whose class covers one of this tile's panels
        int first = 0;
        if (CLS >= 0) {
            int lo = 99;
            for (int q = 0; q < P.np; ++q) {
                const int slot = (P.p0 + q) % P.slots;
                const uint8_t* prec = P.prec + slot * 1024 + (P.rows0 - (int64_t)q * P.nb) / tile_rows;
                const int r = launch_rank(prec[trow], prec[tcol]);
                lo = r < lo ? r : lo;
            }
            first = lo == launch_rank_of(CLS, FA, FB);
        } else {
            first = 1;                 // one launch for every class
        }
        if (first) t.Cs = P.csrc;
    }
    // the panels this launch applies to the tile (bit q: panel q)
    auto operands = [&](int q, operand_src& orow, operand_src& ocol) {
        const int slot = (P.p0 + q) % P.slots;
        const int64_t roff = P.rows0 - (int64_t)q * P.nb;   // C row 0 within this panel
        const uint8_t* prec = P.prec + slot * 1024 + roff / tile_rows;
        orow.prec = prec[trow];
        ocol.prec = prec[tcol];
        orow.payload = ocol.payload = P.payload + slot * P.pstride + roff * row_stride;
        orow.scales = ocol.scales = P.scales + slot * P.sstride + roff * (k >> 5);
        orow.row_stride = ocol.row_stride = row_stride;
        orow.kb = ocol.kb = k >> 5;
    };
    unsigned use = 0;
    for (int q = 0; q < P.np; ++q) {
        operand_src orow, ocol;
        operands(q, orow, ocol);
        if (CLS >= 0) {            // this launch computes one class (and format pair) only
            // FP32 class: FA 1 both FP32, 2 FP32 against a bf16-exact
            // format (split products), 0 FP32 against FP16
            const int cls = pair_class(orow.prec, ocol.prec);
            const int fa = mpc_dev_is_e4m3(ocol.prec) ? 0 : 1, fb = mpc_dev_is_e4m3(orow.prec) ? 0 : 1;
            const bool f32pair = orow.prec == MPC_PREC_FP32 && ocol.prec == MPC_PREC_FP32;
            const bool f16 = orow.prec == MPC_PREC_FP16 || ocol.prec == MPC_PREC_FP16;
            const int f32kind = f32pair ? 1 : (f16 ? 0 : 2);
            if (cls != CLS || (CLS == MPC_PREC_E4M3 && (fa != FA || fb != FB)) ||
                (CLS == MPC_PREC_FP32 && f32kind != FA))
                continue;
        }
        use |= 1u << q;
    }
    if (!use) return;              // no panel of this launch's class on this tile
    // C's old values in flight during the products (64-row tiles only: at
    // 128-row tiles the K loop's operand buffers need those registers)
    constexpr bool kPrefetchC = QA == 1;
    c_frag<QA> cv;
    if constexpr (kPrefetchC) load_c<QA>(t, cv);
    for (int q = 0; q < P.np; ++q) {
        if (!((use >> q) & 1u)) continue;
        operand_src orow, ocol;
        operands(q, orow, ocol);
        const int cls = pair_class(orow.prec, ocol.prec);
        const int fa = mpc_dev_is_e4m3(ocol.prec) ? 0 : 1, fb = mpc_dev_is_e4m3(orow.prec) ? 0 : 1;
        const int c = CLS >= 0 ? CLS : cls;
        if (c == MPC_PREC_E4M3) {
            if (CLS >= 0) mm_fp8<QA, FA, FB, KC>(ocol, orow, t, k, acc);
            else if (fa == 0 && fb == 0) mm_fp8<QA, 0, 0, KC>(ocol, orow, t, k, acc);
            else if (fa == 0) mm_fp8<QA, 0, 1, KC>(ocol, orow, t, k, acc);
            else if (fb == 0) mm_fp8<QA, 1, 0, KC>(ocol, orow, t, k, acc);
            else mm_fp8<QA, 1, 1, KC>(ocol, orow, t, k, acc);
        } else if (c == MPC_PREC_BF16) {
            mm_16<QA, true, KC>(ocol, orow, t, k, acc);
        } else if (c == MPC_PREC_FP16) {
            mm_16<QA, false, KC>(ocol, orow, t, k, acc);
        } else if (CLS == MPC_PREC_FP32) {
            if constexpr (FA == 1) mm_f32<QA, KC, true, 0>(ocol, orow, t, k, acc);
            else if constexpr (FA == 2) mm_f32<QA, KC, false, 1>(ocol, orow, t, k, acc);
            else mm_f32<QA, KC, false, 0>(ocol, orow, t, k, acc);
        } else {                   // mixed-class launch: FP32 against an 8-bit or BF16 tile
            mm_f32<QA, KC, false, 2>(ocol, orow, t, k, acc);
        }
    }
    if constexpr (!kPrefetchC) load_c<QA>(t, cv);
    store_c<QA>(t, cv, acc);
}

}  // namespace

static int syrk_check(int64_t ldc, int k, int tile_rows) {
    if (ldc > (int64_t(1) << 22)) {       // 32-bit buffer offsets of a 128-column tile
        mpc_set_error("syrk: ldc %lld exceeds 2^22", (long long)ldc);
        return MPC_ERR_ARG;
    }
    if (k % 64 || tile_rows % MPC_SYRK_TILE) {
        mpc_set_error("syrk: k (%d) must be a multiple of 64 and tile_rows (%d) of 128", k,
                      tile_rows);
        return MPC_ERR_ARG;
    }
    return 0;
}

// launch geometry over tile columns [tcol0, tcol1) of the lower triangle
struct syrk_grid {
    int64_t mt, npair, per_xcd, tcol0;
    syrk_shape shape;
    dim3 grid;
};
// trow0 > 0: the rectangle of tile rows [trow0, mt) x tile columns [tcol0,
// tcol1) (trow0 >= tcol1: strictly below the diagonal)
static bool syrk_geometry(int64_t m, int tile, int64_t tcol0, int64_t tcol1, syrk_grid& g,
                          int64_t trow0 = 0) {
    g.mt = (m + tile - 1) / tile;
    if (tcol1 < 0 || tcol1 > g.mt) tcol1 = g.mt;
    if (tcol0 >= tcol1) return false;
    if (trow0 > 0) {
        if (trow0 >= g.mt) return false;
        g.npair = (g.mt - trow0) * (tcol1 - tcol0);
        g.shape = syrk_shape{g.mt, tcol0, tcol1, trow0, tcol1 - tcol0};
    } else {
        auto start = [&](int64_t c) { return c * g.mt - c * (c - 1) / 2; };
        g.npair = start(tcol1) - start(tcol0);
        g.shape = syrk_shape{g.mt, tcol0, tcol1, 0, 0};
    }
    g.per_xcd = (g.npair + 7) / 8;
    g.tcol0 = tcol0;
    g.grid = dim3((unsigned)(g.per_xcd * 8));
    return true;
}

#define MPC_SYRK_LAUNCH(QA, CLS, FA, FB, KC)                                                   \
    do {                                                                                       \
        const syrk_lists L_ = next_lists();                                                    \
        if (L_.mode == 2)                                                                      \
            hipLaunchKernelGGL((syrk_kernel<QA, CLS, FA, FB, KC, true>), list_grid, dim3(256), \
                               0, st, C, m, ldc, k, P, row_stride, tile_rows, sm, g.npair,     \
                               g.per_xcd, g.shape, L_);                                        \
        else                                                                                   \
            hipLaunchKernelGGL((syrk_kernel<QA, CLS, FA, FB, KC, false>), g.grid, dim3(256), 0, \
                               st, C, m, ldc, k, P, row_stride, tile_rows, sm, g.npair,        \
                               g.per_xcd, g.shape, L_);                                        \
    } while (0)

// C(lower) -= Q Q^T restricted to the tile columns [tcol0, tcol1) of C, one
// panel.  mixed = 0: tiles of MPC_SYRK_TILE, one launch per compute class
// present in class_mask (and per FP8 element-format pair in fp8_mask; bit 0
// E4M3 elements, bit 1 E5M2).  mixed = 1: 64-row tiles in ONE launch in which
// each workgroup takes its own pair's class (the block column on the
// critical path of the look-ahead: a launch per class would serialise, and a
// 128-row tile per workgroup leaves most of the chip idle there).
int mpc_launch_syrk_classes(hipStream_t st, float* C, int64_t m, int64_t ldc, int k,
                            const uint8_t* payload, int64_t row_stride, const uint8_t* scales,
                            int tile_rows, const uint8_t* prec_of_tile, const mpc_small* sm,
                            unsigned class_mask, unsigned fp8_mask, int64_t tcol0, int64_t tcol1,
                            int mixed, int* lists, int64_t trow0) {
    if (mixed != 0 && mixed != 1) {
        mpc_set_error("syrk: mixed must be 0 or 1 (got %d)", mixed);
        return MPC_ERR_ARG;
    }
    mpc_syrk_panels P{};
    P.payload = payload;
    P.scales = scales;
    P.prec = prec_of_tile;
    P.slots = 1;
    P.np = 1;
    P.nb = tile_rows;
    return mpc_launch_syrk_panels(st, C, m, ldc, k, P, row_stride, tile_rows, sm, tcol0, tcol1,
                                  mixed ? 1 : 2, class_mask, fp8_mask, lists, trow0);
}

__global__ void clear_list_heads_kernel(int* __restrict__ heads) {
    if (threadIdx.x < kListHead) heads[threadIdx.x] = 0;
}

int64_t mpc_syrk_list_ints(int64_t m, int tile) {
    const int64_t mt = (m + tile - 1) / tile;
    return kListHead + 9 * (mt * (mt + 1) / 2);
}

int mpc_launch_syrk_panels(hipStream_t st, float* C, int64_t m, int64_t ldc, int k,
                           const mpc_syrk_panels& P, int64_t row_stride, int tile_rows,
                           const mpc_small* sm, int64_t tcol0, int64_t tcol1, int qa,
                           unsigned class_mask, unsigned fp8_mask, int* lists, int64_t trow0) {
    if (m <= 0) return 0;
    if ((qa != 1 && qa != 2) || P.np < 1 || P.np > MPC_SYRK_MAXP || P.slots < 1 ||
        P.rows0 - (int64_t)(P.np - 1) * P.nb < 0 || P.rows0 % tile_rows || P.nb % tile_rows) {
        mpc_set_error("syrk: bad panel set (qa %d, np %d of at most %d, rows0 %lld, nb %d)", qa,
                      P.np, MPC_SYRK_MAXP, (long long)P.rows0, P.nb);
        return MPC_ERR_ARG;
    }
    int rc = syrk_check(ldc, k, tile_rows);
    if (rc) return rc;
    syrk_grid g;
    if (trow0 > 0 && tcol1 >= 0 && trow0 < tcol1) {
        mpc_set_error("syrk: rectangle rows from tile %lld overlap its columns up to %lld",
                      (long long)trow0, (long long)tcol1);
        return MPC_ERR_ARG;
    }
    if (!syrk_geometry(m, 64 * qa, tcol0, tcol1, g, trow0)) return 0;
    syrk_lists none{nullptr, 0, 0};
    if (qa == 1) {
        // latency-bound: as many resident workgroups as possible.  The K loop
        // unrolled for k = 256 took 240 VGPRs (two workgroups per CU, two
        // rounds of workgroups at the first steps); the runtime-k loop takes
        // 152 (three per CU) and made the solve 0.3 ms faster.
        const dim3 list_grid = g.grid;
        auto next_lists = [&]() { return none; };
        MPC_SYRK_LAUNCH(1, -1, 0, 0, 0);
        MPC_LAUNCH_CHECK();
        return 0;
    }
    // one launch per compute class (and FP8 element pair) in the masks, in
    // launch-rank order; with lists, the first one builds the others' lists
    const bool e4 = (class_mask & (1u << MPC_PREC_E4M3)) && (fp8_mask & 1u);
    const bool e5 = (class_mask & (1u << MPC_PREC_E4M3)) && (fp8_mask & 2u);
    const bool f32 = class_mask & (1u << MPC_PREC_FP32);
    const bool f32s = f32 && (class_mask & ((1u << MPC_PREC_E4M3) | (1u << MPC_PREC_BF16)));
    const bool f32h = f32 && (class_mask & (1u << MPC_PREC_FP16));
    const int nlaunch = (int)e4 + 2 * (int)(e4 && e5) + (int)e5 +
                        (int)(bool)(class_mask & (1u << MPC_PREC_BF16)) +
                        (int)(bool)(class_mask & (1u << MPC_PREC_FP16)) + (int)f32 + (int)f32s +
                        (int)f32h;
    const bool use_lists = lists && nlaunch > 1;
    int launched = 0;
    if (use_lists) hipLaunchKernelGGL(clear_list_heads_kernel, dim3(1), dim3(64), 0, st, lists);
    // later launches: a grid that strides over the list (at most the tile count)
    const dim3 list_grid((unsigned)std::min<int64_t>(g.per_xcd * 8, 2048));
    auto next_lists = [&]() {
        syrk_lists L{use_lists ? lists : nullptr, (int)g.npair,
                     use_lists ? (launched == 0 ? 1 : 2) : 0};
        ++launched;
        return L;
    };
    if (e4) MPC_SYRK_LAUNCH(2, MPC_PREC_E4M3, 0, 0, 0);
    if (e4 && e5) { MPC_SYRK_LAUNCH(2, MPC_PREC_E4M3, 0, 1, 0); MPC_SYRK_LAUNCH(2, MPC_PREC_E4M3, 1, 0, 0); }
    if (e5) MPC_SYRK_LAUNCH(2, MPC_PREC_E4M3, 1, 1, 0);
    if (class_mask & (1u << MPC_PREC_BF16)) MPC_SYRK_LAUNCH(2, MPC_PREC_BF16, 0, 0, 0);
    if (class_mask & (1u << MPC_PREC_FP16)) MPC_SYRK_LAUNCH(2, MPC_PREC_FP16, 0, 0, 0);
    if (f32) MPC_SYRK_LAUNCH(2, MPC_PREC_FP32, 1, 0, 0);         // FP32 x FP32
    if (f32s) MPC_SYRK_LAUNCH(2, MPC_PREC_FP32, 2, 0, 0);        // FP32 x FP8 / BF16 (split)
    if (f32h) MPC_SYRK_LAUNCH(2, MPC_PREC_FP32, 0, 0, 0);        // FP32 x FP16
    MPC_LAUNCH_CHECK();
    return 0;
}
#undef MPC_SYRK_LAUNCH

int mpc_launch_syrk(hipStream_t st, float* C, int64_t m, int64_t ldc, int k,
                    const uint8_t* payload, int64_t row_stride, const uint8_t* scales,
                    int tile_rows, const uint8_t* prec_of_tile, const mpc_small* sm) {
    const unsigned all = (1u << MPC_PREC_E4M3) | (1u << MPC_PREC_BF16) | (1u << MPC_PREC_FP16) |
                         (1u << MPC_PREC_FP32);
    return mpc_launch_syrk_classes(st, C, m, ldc, k, payload, row_stride, scales, tile_rows,
                                   prec_of_tile, sm, all, 3u, 0, -1, 0);
}
