// mpc_formats.hpp -- device-side number-format helpers (gfx950).
//
// Element rounding follows lo_float's ConvertImpl (lo_float.h:1927-2117, RNE,
// non-saturating) on every value the quantiser feeds it: scales guarantee
// |x / 2^e| <= max finite and NaN/Inf are screened, and on that domain the
// hardware OCP conversions (v_cvt_pk_fp8_f32 / v_cvt_pk_bf8_f32, RNE) give the
// same bytes.  The exhaustive sweep test checks this over all 2^32 inputs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/mpchol.h"

__device__ __forceinline__ double mpc_dev_unit_roundoff(int p) {
    switch (p) {
    case MPC_PREC_FP32: return 5.9604644775390625e-08;
    case MPC_PREC_BF16: return 3.90625e-03;
    case MPC_PREC_FP16: return 4.8828125e-04;
    case MPC_PREC_E5M2: case MPC_PREC_MXE5M2: return 0.125;
    default: return 0.0625;
    }
}

__device__ __forceinline__ int mpc_dev_elem_bytes(int p) {
    return p == MPC_PREC_FP32 ? 4 : (p == MPC_PREC_BF16 || p == MPC_PREC_FP16) ? 2 : 1;
}

__device__ __forceinline__ bool mpc_dev_is_e4m3(int p) {
    return p == MPC_PREC_E4M3 || p == MPC_PREC_MXE4M3;
}

// smallest e with amax / 2^e <= max finite (E4M3 448 = 1.75*2^8, E5M2 57344 =
// 1.75*2^15), clamped to the E8M0 range
__device__ __forceinline__ int mpc_dev_scale_exp(float amax, int prec) {
    uint32_t u = __float_as_uint(amax) & 0x7fffffffu;
    if (u == 0) return 0;
    int E;
    uint32_t mant;
    if ((u >> 23) == 0) {
        int lz = __clz(u) - 8;
        E = -126 - lz;
        mant = (u << lz) & 0x7fffffu;
    } else {
        E = (int)(u >> 23) - 127;
        mant = u & 0x7fffffu;
    }
    int emax = mpc_dev_is_e4m3(prec) ? 8 : 15;
    int e = E - emax + (mant > 0x600000u ? 1 : 0);
    return e < -127 ? -127 : (e > 127 ? 127 : e);
}

__device__ __forceinline__ uint32_t mpc_dev_f32_to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}

__device__ __forceinline__ uint32_t mpc_dev_f32_to_f16(float x) {
    _Float16 h = (_Float16)x;
    return (uint32_t)__builtin_bit_cast(unsigned short, h);
}

// four floats -> four 8-bit codes (byte 0 = a0)
__device__ __forceinline__ uint32_t mpc_dev_pack4_f8(float a0, float a1, float a2, float a3,
                                                     int prec) {
    int w = 0;
    if (mpc_dev_is_e4m3(prec)) {
        w = __builtin_amdgcn_cvt_pk_fp8_f32(a0, a1, w, false);
        w = __builtin_amdgcn_cvt_pk_fp8_f32(a2, a3, w, true);
    } else {
        w = __builtin_amdgcn_cvt_pk_bf8_f32(a0, a1, w, false);
        w = __builtin_amdgcn_cvt_pk_bf8_f32(a2, a3, w, true);
    }
    return (uint32_t)w;
}

// Payload layout of an 8-bit tile: per 32-row block and 64-wide K chunk,
// 2 KB in the order the 32x32x64 matrix-core op takes one operand, so that a
// wave reads a block's chunk with two 16-byte loads per lane that together
// cover 2 KB of consecutive bytes.  Lane l = n + 32 h holds 32 bytes: bytes
// 0-15 = row n, K [64c + 16h, +16); bytes 16-31 = row n, K [64c + 32 + 16h,
// +16).  Block b's chunks start at its first row's offset (32 b) * row_stride
// (row_stride >= 64 * ceil(k / 64) bytes); FP32 and 16-bit tiles stay
// row-major.
__device__ __forceinline__ int64_t mpc_f8_offset(int64_t row, int kk, int64_t row_stride) {
    const int c = kk >> 6, kp = kk & 63;
    const int lane = (int)(row & 31) + 32 * ((kp >> 4) & 1);
    return (row & ~(int64_t)31) * row_stride + c * 2048 + lane * 32 + (kp >> 5) * 16 + (kp & 15);
}

// 2^e as a float for e in [-127, 127] (exact)
__device__ __forceinline__ float mpc_dev_exp2i(int e) { return ldexpf(1.0f, e); }
