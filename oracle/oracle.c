/*
 * oracle.c -- CPU restatement of the reference's mixed-precision Cholesky-PCG
 * path.  TEST INFRASTRUCTURE ONLY (see oracle.h): the product never links it.
 *
 * What is restated, and from where (paths relative to the reference repo):
 *   - element rounding FP32 -> E4M3/E5M2: src/lo_float.h:1896-1910
 *     (RoundBitsToNearestEven) and :1927-2117 (ConvertImpl::run with
 *     kSaturate=false, kTruncate=false, STOCHASTIC_ROUND off per :6-7);
 *     special encodings :789-847 (e4m3fn) and :962-1025 (e5m2); negation
 *     flips bit 7 (:360-362).
 *   - diagonal-block POTRF: src/micro_chol.hpp:14-44 (same loop order).
 *   - blocked factor: src/switching_chol.cuh:76-224 with the operand forms of
 *     src/fp8_chol.cuh:73-85 (TRSM RIGHT/LOWER/T) and :125-152 (trailing
 *     update of A22 by the panel), Schur-diagonal tracking
 *     src/kernels_macros.cuh:177-190 (update_diag).
 *   - PCG recurrence and stopping rule: src/mixed_cholesky.cu:145-217.
 *
 * Deviations from the reference (its code cannot run; these follow intent):
 *   D1 TRSM uses the RIGHT/LOWER/TRANS form on the s x b panel below the
 *      diagonal block (switching_chol.cuh:98-110 uses LEFT/N, out of bounds).
 *   D2 The trailing update targets A22 = A[e:,e:] with m = s and only its lower
 *      triangle (switching_chol.cuh:145-219 points C at the diagonal block with
 *      m = n - i*b and computes the full square).
 *   D3 Precision is chosen per panel tile by the min-pivot rule
 *      (orc_rule_minpivot) instead of can_switch's racy per-step atomicExch
 *      (kernels_macros.cuh:191-211, read on the host at switching_chol.cuh:124).
 *      Two other rules are options: the tile's Frobenius norm against its
 *      diagonal block's (rule 1), and can_switch's own test evaluated
 *      deterministically for the whole panel (rule 2): the lowest format p
 *      with u_p * max_i d_i / a_ii <= eps' over the rows below the panel,
 *      d_i the Schur-diagonal tracker, a_ii the original diagonal.
 *      FP16, stored unscaled, is skipped for a tile whose largest magnitude
 *      exceeds FP16's largest finite value 65504 (the next ladder format is
 *      taken instead).
 *   D4 FP8 scales are powers of two, the smallest with amax/s <= max finite
 *      (no saturation); MX blocks are 32 consecutive elements along k of one
 *      row (fp8_utils.cuh:42-58 reduces the wrong lanes and scales to [-1,1]).
 *   D5 A non-positive pivot stops the factor and returns the 1-based column
 *      (micro_chol.hpp:24,29 prints and continues with NaN).
 *   D6 PCG stops on the FP64 true residual ||b-Ax||/||b|| by default
 *      (mixed_cholesky.cu:81 hard-codes rel_tol = 0 so it never stops early);
 *      the reference's sqrt(r.z)/||b|| rule is stop_rule = 1.
 */
#include "oracle.h"
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static inline uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

/* ------------------------------------------------------------------------ */
/* lo_float ConvertImpl<float, float8_*, false, false>::run, restated for the
   two 8-bit targets.  WideBits is uint32 for both (31 significant bits). */
static inline uint32_t rne_bits(uint32_t bits, int roundoff) {
    /* lo_float.h:1896-1910 */
    uint32_t bias = roundoff == 0 ? 0u
        : ((bits >> roundoff) & 1u) + (1u << (roundoff - 1)) - 1u;
    return bits + bias;
}

static inline uint8_t to_f8(float from, int to_mbits, int to_bias, uint8_t highest,
                            uint8_t inf_code, uint8_t nan_code) {
    const int from_mbits = 23, from_bias = 127;
    const int digit_shift = to_mbits - from_mbits;          /* negative */
    const int exp_offset = to_bias - from_bias;             /* negative */
    uint32_t u = f2u(from);
    int sign = (int)(u >> 31);
    uint32_t from_bits = u & 0x7fffffffu;
    if (from_bits == 0x7f800000u) return sign ? (uint8_t)(inf_code ^ 0x80) : inf_code;
    if (from_bits > 0x7f800000u) return nan_code;            /* quiet_NaN, unsigned */
    if (from_bits == 0) return sign ? 0x80 : 0x00;
    int biased_from_exp = (int)(from_bits >> from_mbits);
    /* To has a larger min_exponent than float: values may become subnormal */
    int unbiased = biased_from_exp - from_bias;
    int biased_to = unbiased + to_bias;
    if (biased_to <= 0) {
        uint32_t has_one = biased_from_exp > 0 ? 1u : 0u;
        int exponent_shift = -digit_shift - biased_to + (int)has_one;
        uint32_t rounded = (from_bits & 0x7fffffu) | (has_one << from_mbits);
        uint32_t bits = 0;
        if (exponent_shift <= from_mbits + 1) {
            rounded = rne_bits(rounded, exponent_shift);
            bits = rounded >> exponent_shift;
        }
        uint8_t to = (uint8_t)bits;
        return sign ? (uint8_t)(to ^ 0x80) : to;
    }
    uint32_t rounded = rne_bits(from_bits, -digit_shift);
    rounded &= ~((1u << (-digit_shift)) - 1u);
    rounded += (uint32_t)exp_offset << from_mbits;          /* wraps like WideBits */
    uint32_t aligned_highest = (uint32_t)highest << (-digit_shift);
    uint8_t to = (uint8_t)(rounded >> (-digit_shift));
    if (rounded > aligned_highest) to = inf_code;
    return sign ? (uint8_t)(to ^ 0x80) : to;
}

uint8_t orc_f32_to_e4m3(float x) { return to_f8(x, 3, 7, 0x7e, 0x7f, 0x80); }
uint8_t orc_f32_to_e5m2(float x) { return to_f8(x, 2, 15, 0x7b, 0x7c, 0x7e); }

void orc_f32_to_e4m3_n(const float* x, uint8_t* y, int64_t n) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) y[i] = orc_f32_to_e4m3(x[i]);
}
void orc_f32_to_e5m2_n(const float* x, uint8_t* y, int64_t n) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) y[i] = orc_f32_to_e5m2(x[i]);
}

/* Compare conversions of float bit patterns [lo, hi) with expect[].  Returns
   mismatches among finite inputs with |x| <= max_abs (the quantiser's domain);
   *other receives mismatches outside it (NaN/Inf/overflow encodings). */
static int64_t sweep(int e4m3, uint64_t lo, uint64_t hi, const uint8_t* expect, double max_abs,
                     int64_t* other) {
    int64_t bad = 0, bad_other = 0;
    #pragma omp parallel for schedule(static) reduction(+:bad, bad_other)
    for (int64_t i = (int64_t)lo; i < (int64_t)hi; ++i) {
        float x = u2f((uint32_t)i);
        uint8_t got = e4m3 ? orc_f32_to_e4m3(x) : orc_f32_to_e5m2(x);
        if (got != expect[i - (int64_t)lo]) {
            if (isfinite(x) && fabs((double)x) <= max_abs) bad++;
            else bad_other++;
        }
    }
    if (other) *other = bad_other;
    return bad;
}
int64_t orc_sweep_e4m3(uint64_t lo, uint64_t hi, const uint8_t* expect, double max_abs,
                       int64_t* other) {
    return sweep(1, lo, hi, expect, max_abs, other);
}
int64_t orc_sweep_e5m2(uint64_t lo, uint64_t hi, const uint8_t* expect, double max_abs,
                       int64_t* other) {
    return sweep(0, lo, hi, expect, max_abs, other);
}

/* Independent closed form for the exhaustive self-check: round |x| to the
   grid of the target format with frexp/ldexp arithmetic (ties to even), then
   encode.  Shares no code with to_f8.  Non-saturating like lo_float. */
static uint8_t closed_form_f8(float x, int mbits, int bias, double maxfin, uint8_t inf_code,
                              uint8_t nan_code) {
    if (isnan(x)) return nan_code;
    int neg = signbit(x) != 0;
    double a = fabs((double)x);
    uint8_t code;
    if (isinf(x)) {
        code = inf_code;
    } else {
        int e;
        (void)frexp(a, &e);                 /* a = f * 2^e, f in [0.5, 1) */
        int ex = e - 1;                     /* unbiased exponent of a */
        int emin = 1 - bias;
        if (ex < emin) ex = emin;           /* subnormal grid */
        double q = ldexp(1.0, ex - mbits);  /* grid spacing */
        double t = a / q;                   /* exact: power-of-two division */
        double fl = floor(t);
        double rem = t - fl;
        double r = fl;
        if (rem > 0.5 || (rem == 0.5 && fmod(fl, 2.0) != 0.0)) r = fl + 1.0;
        double v = r * q;
        if (v > maxfin) {
            code = inf_code;
        } else if (v == 0.0) {
            code = 0;
        } else {
            int e2;
            (void)frexp(v, &e2);
            int ev = e2 - 1;
            if (ev < emin) {                /* subnormal: mantissa = v / 2^(emin - mbits) */
                code = (uint8_t)(v / ldexp(1.0, emin - mbits));
            } else {
                int m = (int)(v / ldexp(1.0, ev - mbits)) - (1 << mbits);
                code = (uint8_t)(((ev + bias) << mbits) | m);
            }
        }
    }
    return neg ? (uint8_t)(code | 0x80) : code;
}
uint8_t orc_closed_form_e4m3(float x) { return closed_form_f8(x, 3, 7, 448.0, 0x7f, 0x80); }
uint8_t orc_closed_form_e5m2(float x) { return closed_form_f8(x, 2, 15, 57344.0, 0x7c, 0x7e); }

/* exhaustive self-check over all 2^32 float bit patterns in [lo, hi): number
   of disagreements between the lo_float restatement and the closed form */
int64_t orc_selfcheck_f8(int e4m3, uint64_t lo, uint64_t hi) {
    int64_t bad = 0;
    #pragma omp parallel for schedule(static) reduction(+:bad)
    for (int64_t i = (int64_t)lo; i < (int64_t)hi; ++i) {
        float x = u2f((uint32_t)i);
        uint8_t a = e4m3 ? orc_f32_to_e4m3(x) : orc_f32_to_e5m2(x);
        uint8_t b = e4m3 ? orc_closed_form_e4m3(x) : orc_closed_form_e5m2(x);
        bad += a != b;
    }
    return bad;
}

/* C (lower, m x m, ldc) -= Q Q^T, Q column-major (ldq) FP32; exported for tests */
static void syrk_lower_f32(float* C, int64_t m, int64_t ldc, const float* Q, int64_t ldq,
                           int64_t k);
void orc_syrk_lower(float* C, int64_t m, int64_t ldc, const float* Q, int64_t ldq, int64_t k) {
    syrk_lower_f32(C, m, ldc, Q, ldq, k);
}

/* decode: value = (-1)^s * 2^(e-bias) * 1.m, subnormal 2^(1-bias) * 0.m */
static float f8_decode(uint8_t v, int mbits, int ebits, int bias) {
    int s = v >> 7, e = (v >> mbits) & ((1 << ebits) - 1), m = v & ((1 << mbits) - 1);
    double mag = e == 0 ? ldexp((double)m, 1 - bias - mbits)
                        : ldexp((double)((1 << mbits) | m), e - bias - mbits);
    return (float)(s ? -mag : mag);
}
float orc_e4m3_to_f32(uint8_t v) {
    if ((v & 0x7f) == 0x7f) return (v & 0x80) ? -INFINITY : INFINITY;  /* lo_float "inf" */
    if (v == 0x80) return NAN;                                         /* lo_float NaN */
    return f8_decode(v, 3, 4, 7);
}
float orc_e5m2_to_f32(uint8_t v) {
    int e = (v >> 2) & 31, m = v & 3;
    if (e == 31) return m == 0 ? ((v & 0x80) ? -INFINITY : INFINITY) : NAN;
    return f8_decode(v, 2, 5, 15);
}

/* IEEE binary16 / bfloat16, round to nearest even */
uint16_t orc_f32_to_bf16(float x) {
    uint32_t u = f2u(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
float orc_bf16_to_f32(uint16_t v) { return u2f((uint32_t)v << 16); }

uint16_t orc_f32_to_f16(float x) {
    uint32_t u = f2u(x);
    uint16_t sign = (uint16_t)((u >> 16) & 0x8000u);
    uint32_t a = u & 0x7fffffffu;
    if (a > 0x7f800000u) return sign | 0x7e00;                 /* NaN */
    if (a >= 0x477ff000u) return sign | 0x7c00;                /* >= 65520 rounds to inf */
    if (a < 0x38800000u) {                                     /* below 2^-14: subnormal */
        if (a < 0x33000000u) return sign;                      /* < 2^-25: rounds to 0 */
        int e = (int)(a >> 23);
        uint32_t mant = (a & 0x7fffffu) | 0x800000u;
        int shift = 126 - e;                                   /* 14..24 */
        /* value = mant * 2^(e-150); subnormal unit 2^-24 -> q = mant >> (shift) with rne */
        uint32_t q = mant >> shift;
        uint32_t rem = mant & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1);
        if (rem > half || (rem == half && (q & 1u))) q++;
        return sign | (uint16_t)q;
    }
    uint32_t r = a + 0xfffu + ((a >> 13) & 1u);
    r -= 0x38000000u;                                          /* rebias 127 -> 15 */
    return sign | (uint16_t)(r >> 13);
}
float orc_f16_to_f32(uint16_t v) {
    int s = v >> 15, e = (v >> 10) & 31, m = v & 1023;
    double mag;
    if (e == 31) return m ? NAN : (s ? -INFINITY : INFINITY);
    mag = e == 0 ? ldexp((double)m, -24) : ldexp((double)(1024 | m), e - 25);
    return (float)(s ? -mag : mag);
}

int orc_elem_bytes(int prec) {
    switch (prec) {
    case ORC_FP32: return 4;
    case ORC_BF16: case ORC_FP16: return 2;
    default: return 1;
    }
}

double orc_unit_roundoff(int prec) {
    switch (prec) {
    case ORC_FP32: return ldexp(1.0, -24);
    case ORC_BF16: return ldexp(1.0, -8);
    case ORC_FP16: return ldexp(1.0, -11);
    case ORC_E5M2: case ORC_MXE5M2: return ldexp(1.0, -3);
    default: return ldexp(1.0, -4);
    }
}

/* max finite value of the element format as (exponent, mantissa-field of FP32) */
static int fmt_emax(int prec) {
    return (prec == ORC_E4M3 || prec == ORC_MXE4M3) ? 8 : 15;
}
static uint32_t fmt_maxmant(int prec) {
    (void)prec;
    return 0x600000u;   /* 1.75 for both E4M3 (448) and E5M2 (57344) */
}

int orc_scale_exp(float amax, int prec) {
    if (!(prec == ORC_E4M3 || prec == ORC_E5M2 || prec == ORC_MXE4M3 || prec == ORC_MXE5M2))
        return 0;
    uint32_t u = f2u(amax) & 0x7fffffffu;
    if (u == 0) return 0;
    int E;
    uint32_t mant;
    if ((u >> 23) == 0) {             /* subnormal amax: normalise */
        int lz = __builtin_clz(u) - 8; /* leading zeros within the 24-bit field */
        E = -126 - lz;
        mant = (u << lz) & 0x7fffffu;
    } else {
        E = (int)(u >> 23) - 127;
        mant = u & 0x7fffffu;
    }
    int e = E - fmt_emax(prec) + (mant > fmt_maxmant(prec) ? 1 : 0);
    if (e < -127) e = -127;
    if (e > 127) e = 127;
    return e;
}

static inline uint8_t enc_elem(float v, int prec) {
    return (prec == ORC_E4M3 || prec == ORC_MXE4M3) ? orc_f32_to_e4m3(v) : orc_f32_to_e5m2(v);
}
/* Arithmetic meaning of a payload byte.  The bytes are lo_float-exact, but
   lo_float writes -0 (and negatives that round to zero) as 0x80, its own NaN
   code; the matrix cores read OCP E4M3, where 0x80 is -0.  NaN/Inf never
   reach the quantiser (screened), so OCP decoding is the right meaning. */
static inline float dec_elem(uint8_t v, int prec) {
    if (prec == ORC_E4M3 || prec == ORC_MXE4M3)
        return v == 0x80 ? -0.0f : orc_e4m3_to_f32(v);
    return orc_e5m2_to_f32(v);
}

int orc_quantize_panel(const float* P, int64_t m, int64_t k, int64_t ld,
                       int tile_rows, const uint8_t* prec_of_tile,
                       uint8_t* payload, int64_t row_stride, uint8_t* scales,
                       float* deq, int64_t ld_deq) {
    int64_t ntile = (m + tile_rows - 1) / tile_rows;
    int64_t kb = k / 32;
    int bad = 0;
    #pragma omp parallel for schedule(dynamic, 1) reduction(|:bad)
    for (int64_t t = 0; t < ntile; ++t) {
        int prec = prec_of_tile[t];
        int64_t r0 = t * tile_rows, r1 = r0 + tile_rows < m ? r0 + tile_rows : m;
        int tile_scaled = prec == ORC_E4M3 || prec == ORC_E5M2;
        int mx = prec == ORC_MXE4M3 || prec == ORC_MXE5M2;
        int tile_e = 0;
        for (int64_t r = r0; r < r1; ++r)
            for (int64_t c = 0; c < k; ++c)
                if (!isfinite(P[r + c * ld])) bad = 1;
        if (tile_scaled) {
            float am = 0.f;
            for (int64_t c = 0; c < k; ++c)
                for (int64_t r = r0; r < r1; ++r) {
                    float a = fabsf(P[r + c * ld]);
                    if (a > am) am = a;
                }
            tile_e = orc_scale_exp(am, prec);
        }
        for (int64_t r = r0; r < r1; ++r) {
            uint8_t* row = payload + r * row_stride;
            for (int64_t b = 0; b < kb; ++b) {
                int e = tile_e;
                if (mx) {
                    float am = 0.f;
                    for (int64_t c = b * 32; c < b * 32 + 32; ++c) {
                        float a = fabsf(P[r + c * ld]);
                        if (a > am) am = a;
                    }
                    e = orc_scale_exp(am, prec);
                }
                if (scales) scales[r * kb + b] = (uint8_t)(e + 127);
                for (int64_t c = b * 32; c < b * 32 + 32; ++c) {
                    float x = P[r + c * ld];
                    float q;
                    if (prec == ORC_FP32) {
                        memcpy(row + 4 * c, &x, 4);
                        q = x;
                    } else if (prec == ORC_BF16) {
                        uint16_t h = orc_f32_to_bf16(x);
                        memcpy(row + 2 * c, &h, 2);
                        q = orc_bf16_to_f32(h);
                    } else if (prec == ORC_FP16) {
                        uint16_t h = orc_f32_to_f16(x);
                        memcpy(row + 2 * c, &h, 2);
                        q = orc_f16_to_f32(h);
                    } else {
                        /* x * 2^-e is exact (power of two, no underflow below
                           the e4m3/e5m2 subnormal range matters: RNE then applies) */
                        float xs = ldexpf(x, -e);
                        uint8_t v = enc_elem(xs, prec);
                        row[c] = v;
                        q = ldexpf(dec_elem(v, prec), e);
                    }
                    if (deq) deq[r + c * ld_deq] = q;
                }
            }
        }
    }
    return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 */
static inline void mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    uint64_t p = (uint64_t)a * b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(0xD2511F53u, c0, &hi0, &lo0);
        mulhilo(0xCD9E8D57u, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
static inline double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}
double orc_uniform(uint64_t seed, uint64_t stream, uint64_t idx) {
    uint32_t ctr[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)stream, 0x554e4946u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)}, o[4];
    orc_philox4x32(ctr, key, o);
    return u53(o[0], o[1]);
}
/* trace-estimation probes (no reference counterpart): Z n x nprobe column-major,
   z_pi = -1 if uniform(seed, 3, p*n + i) < 1/2 else +1 */
void orc_probes(int64_t n, int nprobe, uint64_t seed, double* Z, int64_t ldz) {
    for (int p = 0; p < nprobe; ++p)
        for (int64_t i = 0; i < n; ++i)
            Z[(int64_t)p * ldz + i] = orc_uniform(seed, 3, (uint64_t)p * (uint64_t)n + (uint64_t)i) < 0.5
                                          ? -1.0 : 1.0;
}
double orc_normal(uint64_t seed, uint64_t stream, uint64_t idx) {
    uint64_t q = idx >> 1;
    uint32_t ctr[4] = {(uint32_t)q, (uint32_t)(q >> 32), (uint32_t)stream, 0x4e4f524du};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)}, o[4];
    orc_philox4x32(ctr, key, o);
    double u1 = 1.0 - u53(o[0], o[1]);          /* (0, 1] */
    double u2 = u53(o[2], o[3]);                /* [0, 1) */
    double r = sqrt(-2.0 * log(u1));
    double th = 6.283185307179586 * u2;
    return (idx & 1) ? r * sin(th) : r * cos(th);
}

/* ------------------------------------------------------------------------ */
/* generators */
void orc_spectrum(int dist, int64_t n, double cond, double lam_max,
                  int64_t n_cluster, double cluster_lo, double cluster_hi,
                  uint64_t seed, double* lam) {
    if (n_cluster > n) n_cluster = n;
    for (int64_t i = 0; i < n_cluster; ++i) {
        double t = n_cluster > 1 ? (double)i / (double)(n_cluster - 1) : 0.0;
        lam[i] = exp(log(cluster_lo) + t * (log(cluster_hi) - log(cluster_lo)));
    }
    int64_t m = n - n_cluster;
    double* l = lam + n_cluster;
    for (int64_t i = 0; i < m; ++i) {
        double t = m > 1 ? (double)i / (double)(m - 1) : 0.0;
        double v;
        switch (dist) {
        case 0: {   /* RandomLogUniform, kernels_macros.cuh:635-647 (Philox, not curand) */
            double u = 1.0 - orc_uniform(seed, 7, (uint64_t)i);
            v = pow(10.0, log10(1.0 / cond) * (1.0 - u));
            break; }
        case 1:     /* Clustered, kernels_macros.cuh:650-656 */
            v = i < m - 1 ? 1.0 : 1.0 / cond;
            break;
        case 2:     /* Arithmetic, kernels_macros.cuh:659-666 */
            v = 1.0 - (1.0 - t) * (1.0 - 1.0 / cond);
            break;
        case 3:     /* Geometric, kernels_macros.cuh:669-675 */
            v = pow(cond, -t);
            break;
        default:    /* LogSpaced increasing from lam_max/cond to lam_max */
            v = exp(log(1.0 / cond) * (1.0 - t));
            break;
        }
        l[i] = v * lam_max;
    }
}

static inline double dot4(const double* a, const double* b) {
    return fma(a[0], b[0], fma(a[1], b[1], fma(a[2], b[2], a[3] * b[3])));
}

int orc_hh_setup(int64_t n, int nrefl, uint64_t seed, const double* lam,
                 double* W, double* P, double* g) {
    if (nrefl < 1 || nrefl > 4) return -1;
    memset(W, 0, sizeof(double) * 4 * n);
    for (int k = 0; k < nrefl; ++k) {
        double s = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            double v = orc_normal(seed, (uint64_t)k, (uint64_t)i);
            W[4 * i + k] = v;
            s += v * v;
        }
        double inv = 1.0 / sqrt(s);
        for (int64_t i = 0; i < n; ++i) W[4 * i + k] *= inv;
    }
    /* compact WY: H_1..H_r = I - W T W^T, T upper triangular (LAPACK dlarft) */
    double T[4][4] = {{0}};
    for (int k = 0; k < nrefl; ++k) {
        double y[4] = {0};
        for (int a = 0; a < k; ++a) {
            double s = 0.0;
            for (int64_t i = 0; i < n; ++i) s += W[4 * i + a] * W[4 * i + k];
            y[a] = s;
        }
        for (int a = 0; a < k; ++a) {
            double s = 0.0;
            for (int b = a; b < k; ++b) s += T[a][b] * y[b];
            T[a][k] = -2.0 * s;
        }
        T[k][k] = 2.0;
    }
    double G[4][4] = {{0}};
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
            double s = 0.0;
            for (int64_t i = 0; i < n; ++i) s += W[4 * i + a] * lam[i] * W[4 * i + b];
            G[a][b] = s;
        }
    for (int64_t i = 0; i < n; ++i) {
        for (int b = 0; b < 4; ++b) {
            double s = 0.0;
            for (int a = 0; a <= b; ++a) s += W[4 * i + a] * T[a][b];
            P[4 * i + b] = s;
        }
    }
    for (int64_t i = 0; i < n; ++i)
        for (int a = 0; a < 4; ++a) {
            double s = 0.0;
            for (int b = 0; b < 4; ++b) s += G[a][b] * P[4 * i + b];
            g[4 * i + a] = s;
        }
    return 0;
}

double orc_hh_entry(int64_t i, int64_t j, const double* lam, const double* W,
                    const double* P, const double* g) {
    int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    double t = dot4(P + 4 * lo, g + 4 * hi);
    t = fma(-lam[hi], dot4(P + 4 * lo, W + 4 * hi), t);
    t = fma(-lam[lo], dot4(P + 4 * hi, W + 4 * lo), t);
    if (lo == hi) t += lam[lo];
    return t;
}

void orc_hh_matrix(int64_t n, const double* lam, const double* W, const double* P,
                   const double* g, float* A, int64_t lda) {
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i)
            A[i + j * lda] = (float)orc_hh_entry(i, j, lam, W, P, g);
}

void orc_rbf_matrix(int64_t n, int d, const float* X, int64_t ldx, const double* ls,
                    int n_ls, double variance, double nugget, float* K, int64_t ldk) {
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) {
            double s = 0.0;
            for (int c = 0; c < d; ++c) {
                double l = ls[n_ls == 1 ? 0 : c];
                double t = ((double)X[i * ldx + c] - (double)X[j * ldx + c]) / l;
                s += t * t;
            }
            double v = variance * exp(-0.5 * s);
            if (i == j) v += nugget;
            K[i + j * ldk] = (float)v;
        }
}

void orc_points_gaussian(int64_t n, int d, uint64_t seed, float* X, int64_t ldx) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < d; ++c)
            X[i * ldx + c] = (float)orc_normal(seed, 0, (uint64_t)(i * d + c));
}

void orc_points_sparse_image(int64_t n, int d, double p_zero, uint64_t seed,
                             float* X, int64_t ldx) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < d; ++c) {
            uint64_t idx = (uint64_t)(i * d + c);
            double u = orc_uniform(seed, 1, idx);
            X[i * ldx + c] = u < p_zero ? 0.0f : (float)(1.0 - orc_uniform(seed, 2, idx));
        }
}

/* ------------------------------------------------------------------------ */
/* FP64 oracle */
int orc_potrf_f64(double* A, int64_t n, int64_t lda) {
    const int64_t nb = 64;
    for (int64_t j = 0; j < n; j += nb) {
        int64_t b = j + nb < n ? nb : n - j;
        /* unblocked on the diagonal block */
        for (int64_t k = j; k < j + b; ++k) {
            double d = A[k + k * lda];
            if (!(d > 0.0)) return (int)(k + 1);
            d = sqrt(d);
            A[k + k * lda] = d;
            for (int64_t i = k + 1; i < j + b; ++i) A[i + k * lda] /= d;
            for (int64_t c = k + 1; c < j + b; ++c) {
                double lck = A[c + k * lda];
                for (int64_t i = c; i < j + b; ++i) A[i + c * lda] -= A[i + k * lda] * lck;
            }
        }
        int64_t e = j + b;
        if (e >= n) break;
        /* panel: rows below, X L^T = A21 */
        #pragma omp parallel for schedule(static)
        for (int64_t i = e; i < n; ++i)
            for (int64_t c = j; c < e; ++c) {
                double s = A[i + c * lda];
                for (int64_t p = j; p < c; ++p) s -= A[i + p * lda] * A[c + p * lda];
                A[i + c * lda] = s / A[c + c * lda];
            }
        /* trailing lower update */
        #pragma omp parallel for schedule(dynamic, 8)
        for (int64_t c = e; c < n; ++c)
            for (int64_t p = j; p < e; ++p) {
                double lcp = A[c + p * lda];
                double* dst = A + c * lda;
                const double* src = A + p * lda;
                for (int64_t i = c; i < n; ++i) dst[i] -= src[i] * lcp;
            }
    }
    return 0;
}

void orc_trsv_lower_f64(const double* L, int64_t n, int64_t ldl, double* x, int trans) {
    if (!trans) {
        for (int64_t c = 0; c < n; ++c) {
            x[c] /= L[c + c * ldl];
            double xc = x[c];
            for (int64_t i = c + 1; i < n; ++i) x[i] -= L[i + c * ldl] * xc;
        }
    } else {
        for (int64_t c = n - 1; c >= 0; --c) {
            double s = x[c];
            for (int64_t i = c + 1; i < n; ++i) s -= L[i + c * ldl] * x[i];
            x[c] = s / L[c + c * ldl];
        }
    }
}

void orc_symv_f32_f64(const float* A, int64_t n, int64_t lda, const double* x, double* y) {
    int nt = orc_num_threads();
    int64_t chunk = (n + nt - 1) / nt;
    #pragma omp parallel for schedule(static, 1)
    for (int t = 0; t < nt; ++t) {
        int64_t r0 = t * chunk, r1 = r0 + chunk < n ? r0 + chunk : n;
        for (int64_t i = r0; i < r1; ++i) y[i] = 0.0;
        for (int64_t c = 0; c < n; ++c) {
            double xc = x[c];
            const float* col = A + c * lda;
            for (int64_t i = r0; i < r1; ++i) y[i] += (double)col[i] * xc;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* restatement of micro_chol.hpp:14-44 (compile with -ffp-contract=off) */
int orc_potrf_f32_micro(float* A, int n, int ld) {
    if (n == 0) return 0;
    for (int k = 0; k < n; ++k) {
        float* akk = &A[k * ld + k];
        if (!(*akk > 0.0f)) return k + 1;              /* D5 */
        *akk = sqrtf(*akk);
        float d = *akk;
        for (int i = k + 1; i < n; ++i) A[k * ld + i] /= d;
        for (int j = k + 1; j < n; ++j)
            for (int i = j; i < n; ++i)
                A[j * ld + i] -= A[k * ld + i] * A[k * ld + j];
    }
    return 0;
}

int orc_rule_minpivot_amax(const int* ladder, int n_ladder, double eps_prime,
                           double max_row_sumsq, double min_pivot, double tile_amax) {
    for (int q = 0; q < n_ladder; ++q) {
        if (ladder[q] == ORC_FP16 && !(tile_amax <= 65504.0)) continue;
        if (orc_unit_roundoff(ladder[q]) * max_row_sumsq <= eps_prime * min_pivot)
            return ladder[q];
    }
    return ladder[n_ladder - 1];
}

int orc_rule_minpivot(const int* ladder, int n_ladder, double eps_prime,
                      double max_row_sumsq, double min_pivot) {
    return orc_rule_minpivot_amax(ladder, n_ladder, eps_prime, max_row_sumsq, min_pivot, 0.0);
}

/* compute class of a tile pair: FP32 if either is FP32, else the 16-bit class,
   else FP8 (reported under ORC_E4M3) */
static int pair_class(int p, int q) {
    if (p == ORC_FP32 || q == ORC_FP32) return ORC_FP32;
    if (p == ORC_BF16 || q == ORC_BF16) return ORC_BF16;
    if (p == ORC_FP16 || q == ORC_FP16) return ORC_FP16;
    return ORC_E4M3;
}

/* C(lower, m x m, ldc) -= Q Q^T where Q is m x k column-major (ldq), FP32
   accumulation over k per output element (explicit fmaf).  Tiles of 64 x 64. */
static void syrk_lower_f32(float* C, int64_t m, int64_t ldc, const float* Q, int64_t ldq,
                           int64_t k) {
    const int64_t T = 64;
    int64_t nt = (m + T - 1) / T;
    int64_t npair = nt * (nt + 1) / 2;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t pi = 0; pi < npair; ++pi) {
        /* decode pair index -> (ti, tj), tj <= ti */
        int64_t ti = (int64_t)((sqrt(8.0 * (double)pi + 1.0) - 1.0) / 2.0);
        while (ti * (ti + 1) / 2 > pi) --ti;
        while ((ti + 1) * (ti + 2) / 2 <= pi) ++ti;
        int64_t tj = pi - ti * (ti + 1) / 2;
        int64_t r0 = ti * T, r1 = r0 + T < m ? r0 + T : m;
        int64_t c0 = tj * T, c1 = c0 + T < m ? c0 + T : m;
        float acc[64][16];
        for (int64_t cb = c0; cb < c1; cb += 4) {
            int64_t cn = cb + 4 < c1 ? 4 : c1 - cb;
            for (int64_t rb = r0; rb < r1; rb += 16) {
                int64_t rn = rb + 16 < r1 ? 16 : r1 - rb;
                float a[4][16];
                for (int c = 0; c < 4; ++c)
                    for (int r = 0; r < 16; ++r) a[c][r] = 0.f;
                for (int64_t p = 0; p < k; ++p) {
                    const float* qc = Q + p * ldq;
                    float qr[16];
                    for (int r = 0; r < 16; ++r) qr[r] = r < rn ? qc[rb + r] : 0.f;
                    for (int c = 0; c < 4; ++c) {
                        float b = c < cn ? qc[cb + c] : 0.f;
                        for (int r = 0; r < 16; ++r) a[c][r] = fmaf(qr[r], b, a[c][r]);
                    }
                }
                for (int c = 0; c < cn; ++c)
                    for (int r = 0; r < rn; ++r)
                        if (rb + r >= cb + c) C[(rb + r) + (cb + c) * ldc] -= a[c][r];
            }
        }
        (void)acc;
    }
}

int orc_factor_mixed(float* A, int64_t n, int64_t lda, const orc_factor_opts* o,
                     const uint8_t* map_in, uint8_t* map_out, float* updated_diag,
                     int max_steps, orc_factor_info* info) {
    const int64_t nb = o->nb;
    const int64_t nt = (n + nb - 1) / nb;
    const int tr = o->tile_rows > 0 ? o->tile_rows : (int)nb;
    if (tr < nb || tr % nb) return -3;      /* tiles are whole nb blocks */
    orc_factor_info inf;
    memset(&inf, 0, sizeof inf);
    double mu = INFINITY;
    float* d = updated_diag;
    float* dloc = NULL;
    if (!d) d = dloc = (float*)malloc(sizeof(float) * n);
    for (int64_t i = 0; i < n; ++i) d[i] = A[i + i * lda];
    float* d0 = (float*)malloc(sizeof(float) * n);       /* original diagonal (rule 2) */
    memcpy(d0, d, sizeof(float) * n);
    float* Q = (float*)malloc(sizeof(float) * n * nb);
    uint8_t* pay = (uint8_t*)malloc((size_t)n * nb * 4);
    uint8_t* sc = (uint8_t*)malloc((size_t)n * (nb / 32 > 0 ? nb / 32 : 1));
    uint8_t* ptile = (uint8_t*)malloc(nt);
    double* rs = (double*)malloc(sizeof(double) * n);
    int ret = 0;
    if (map_out) memset(map_out, 0xff, (size_t)(nt * nt));
    for (int64_t j = 0; j < nt; ++j) {
        if (max_steps >= 0 && j >= max_steps) break;
        int64_t c0 = j * nb, bs = c0 + nb < n ? nb : n - c0, e = c0 + bs, s = n - e;
        float* Ajj = A + c0 + c0 * lda;
        int f = orc_potrf_f32_micro(Ajj, (int)bs, (int)lda);
        if (f) { ret = (int)(c0 + f); break; }
        for (int64_t k = 0; k < bs; ++k) {
            double p = (double)Ajj[k + k * lda];
            p *= p;
            if (p < mu) mu = p;
        }
        if (s == 0) break;
        /* D1: panel TRSM, X L_jj^T = A21, per row forward substitution */
        #pragma omp parallel for schedule(static)
        for (int64_t i = e; i < n; ++i) {
            float* row = A + i + c0 * lda;
            for (int64_t c = 0; c < bs; ++c) {
                float acc = row[c * lda];
                for (int64_t p = 0; p < c; ++p) acc -= row[p * lda] * Ajj[c + p * lda];
                row[c * lda] = acc / Ajj[c + c * lda];
            }
        }
        /* update_diag (kernels_macros.cuh:177-190) and row sums of squares */
        #pragma omp parallel for schedule(static)
        for (int64_t i = e; i < n; ++i) {
            double acc = 0.0;
            float di = d[i];
            for (int64_t c = 0; c < bs; ++c) {
                float l = A[i + (c0 + c) * lda];
                di -= l * l;
                acc += (double)l * (double)l;
            }
            d[i] = di;
            rs[i] = acc;
        }
        int64_t ntl = (s + tr - 1) / tr;
        /* reference values of the switching rules (D3) */
        double f2jj = 0.0, schur = 0.0;
        if (o->rule == 1) {
            for (int64_t c = 0; c < bs; ++c)
                for (int64_t r = c; r < bs; ++r) f2jj += (double)Ajj[r + c * lda] * (double)Ajj[r + c * lda];
        } else if (o->rule == 2) {
            for (int64_t i = e; i < n; ++i) {
                const float q = d[i] / d0[i];
                if (q > schur) schur = q;               /* NaN and negative ratios count as 0 */
            }
        }
        for (int64_t t = 0; t < ntl; ++t) {
            int p;
            if (map_in) {
                p = map_in[(j + 1 + t * (tr / nb)) + j * nt];
            } else if (o->update_prec >= 0) {
                p = o->update_prec;
            } else {
                double mx = 0.0, am = 0.0;
                for (int64_t i = e + t * tr; i < e + (t + 1) * tr && i < n; ++i) {
                    if (rs[i] > mx) mx = rs[i];
                    for (int64_t c = 0; c < bs; ++c) {
                        const double v = fabs((double)A[i + (c0 + c) * lda]);
                        if (v > am) am = v;
                    }
                }
                if (o->rule == 1) {                     /* ||tile||_F^2 vs ||L_jj||_F^2 */
                    double f2 = 0.0;
                    for (int64_t i = e + t * tr; i < e + (t + 1) * tr && i < n; ++i) f2 += rs[i];
                    p = orc_rule_minpivot_amax(o->ladder, o->n_ladder, o->eps_prime, f2, f2jj, am);
                } else if (o->rule == 2) {              /* max d_i / a_ii vs 1 */
                    p = orc_rule_minpivot_amax(o->ladder, o->n_ladder, o->eps_prime, schur, 1.0, am);
                } else {
                    p = orc_rule_minpivot_amax(o->ladder, o->n_ladder, o->eps_prime, mx, mu, am);
                }
            }
            ptile[t] = (uint8_t)p;
            /* the map is in nb blocks: a tile of tr / nb blocks fills each */
            for (int64_t u = 0; map_out && u < tr / nb; ++u)
                if (j + 1 + t * (tr / nb) + u < nt) map_out[(j + 1 + t * (tr / nb) + u) + j * nt] = (uint8_t)p;
        }
        if (map_out) map_out[j + j * nt] = ORC_FP32;
        /* quantise the panel (bs must be a multiple of 32 when s > 0) */
        if (orc_quantize_panel(A + e + c0 * lda, s, bs, lda, tr, ptile, pay, nb * 4, sc,
                               Q, s) != 0) {
            ret = -2;
            break;
        }
        for (int64_t tt = 0; tt < ntl; ++tt)
            for (int64_t tl = 0; tl <= tt; ++tl) {
                int64_t rr = (tt + 1) * tr < s ? tr : s - tt * tr;
                int64_t cc = (tl + 1) * tr < s ? tr : s - tl * tr;
                double fl = 2.0 * (double)rr * (double)cc * (double)bs;
                if (tt == tl) fl = (double)rr * (double)(rr + 1) * (double)bs;
                inf.flops[pair_class(ptile[tt], ptile[tl])] += fl;
            }
        syrk_lower_f32(A + e + e * lda, s, lda, Q, s, bs);
    }
    inf.info = ret > 0 ? ret : 0;
    inf.min_pivot = mu;
    if (info) *info = inf;
    free(Q); free(pay); free(sc); free(ptile); free(rs); free(dloc); free(d0);
    return ret;
}

/* ------------------------------------------------------------------------ */
/* PCG (mixed_cholesky.cu:145-217) */
static void trsv_f32L_f64(const float* L, int64_t n, int64_t ldl, double* x, int trans) {
    const int64_t B = 256;
    if (!trans) {
        for (int64_t j = 0; j < n; j += B) {
            int64_t e = j + B < n ? j + B : n;
            for (int64_t c = j; c < e; ++c) {
                x[c] /= (double)L[c + c * ldl];
                double xc = x[c];
                for (int64_t i = c + 1; i < e; ++i) x[i] -= (double)L[i + c * ldl] * xc;
            }
            if (e < n) {
                #pragma omp parallel for schedule(static)
                for (int64_t i = e; i < n; ++i) {
                    double s = 0.0;
                    for (int64_t c = j; c < e; ++c) s += (double)L[i + c * ldl] * x[c];
                    x[i] -= s;
                }
            }
        }
    } else {
        int64_t nblk = (n + B - 1) / B;
        for (int64_t jb = nblk - 1; jb >= 0; --jb) {
            int64_t j = jb * B, e = j + B < n ? j + B : n;
            if (e < n) {
                #pragma omp parallel for schedule(static)
                for (int64_t c = j; c < e; ++c) {
                    double s = 0.0;
                    const float* col = L + c * ldl;
                    for (int64_t i = e; i < n; ++i) s += (double)col[i] * x[i];
                    x[c] -= s;
                }
            }
            for (int64_t c = e - 1; c >= j; --c) {
                double s = x[c];
                for (int64_t i = c + 1; i < e; ++i) s -= (double)L[i + c * ldl] * x[i];
                x[c] = s / (double)L[c + c * ldl];
            }
        }
    }
}

static double dotd(const double* a, const double* b, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

int orc_pcg(const float* A, int64_t lda, const float* L, int64_t ldl, int64_t n,
            const double* b, double* x, const orc_pcg_opts* o, double* relres,
            double* alphas, double* betas) {
    double* r = (double*)malloc(sizeof(double) * n);
    double* z = (double*)malloc(sizeof(double) * n);
    double* p = (double*)malloc(sizeof(double) * n);
    double* q = (double*)malloc(sizeof(double) * n);
    double nb = sqrt(dotd(b, b, n));
    if (nb == 0.0) nb = 1.0;
    if (o->x0_mode == 0) {
        memcpy(x, b, sizeof(double) * n);
        trsv_f32L_f64(L, n, ldl, x, 0);
        trsv_f32L_f64(L, n, ldl, x, 1);
    } else {
        memset(x, 0, sizeof(double) * n);
    }
    orc_symv_f32_f64(A, n, lda, x, q);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
    memcpy(z, r, sizeof(double) * n);
    trsv_f32L_f64(L, n, ldl, z, 0);
    trsv_f32L_f64(L, n, ldl, z, 1);
    memcpy(p, z, sizeof(double) * n);
    double rho = dotd(r, z, n);
    double res = o->stop_rule == 1 ? sqrt(fabs(rho)) / nb : sqrt(dotd(r, r, n)) / nb;
    int k = 0, done = res <= o->tol;
    while (!done && k < o->maxit) {
        ++k;
        orc_symv_f32_f64(A, n, lda, p, q);
        double alpha = rho / dotd(p, q, n);
        for (int64_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
        memcpy(z, r, sizeof(double) * n);
        trsv_f32L_f64(L, n, ldl, z, 0);
        trsv_f32L_f64(L, n, ldl, z, 1);
        double rho_new = dotd(r, z, n);
        double beta = rho_new / rho;
        if (alphas) alphas[k - 1] = alpha;
        if (betas) betas[k - 1] = beta;
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        rho = rho_new;
        if (o->stop_rule == 1) {
            res = sqrt(fabs(rho)) / nb;
            done = res <= o->tol;
        } else {
            res = sqrt(dotd(r, r, n)) / nb;
            if (res <= o->tol) {
                /* confirm on the true residual; replace r if it lags */
                orc_symv_f32_f64(A, n, lda, x, q);
                for (int64_t i = 0; i < n; ++i) q[i] = b[i] - q[i];
                double tr = sqrt(dotd(q, q, n)) / nb;
                done = tr <= o->tol;
                if (!done) memcpy(r, q, sizeof(double) * n);
                res = tr;
            }
        }
    }
    orc_symv_f32_f64(A, n, lda, x, q);
    for (int64_t i = 0; i < n; ++i) q[i] = b[i] - q[i];
    if (relres) *relres = sqrt(dotd(q, q, n)) / nb;
    free(r); free(z); free(p); free(q);
    return done ? k : -1;
}

/* Mixed-precision iterative refinement around the low-precision factor (no
   reference routine: the reference's error trace of such a scheme is
   src/e5m2_error_f_cond.csv:1-5, one row per refinement step).  FP64 residual
   and update, correction solve through L:
     x0 = M^-1 b (x0_mode 0) or 0 (x0_mode 1)
     repeat: r = b - A x;  stop if ||r|| / ||b|| <= tol;  x += M^-1 r
   Returns the number of corrections applied if converged, -1 at maxit.
   hist (optional, maxit + 1) receives ||r_k|| / ||b|| for k = 0 .. steps. */
int orc_refine(const float* A, int64_t lda, const float* L, int64_t ldl, int64_t n,
               const double* b, double* x, const orc_pcg_opts* o, double* relres,
               double* hist) {
    double* r = (double*)malloc(sizeof(double) * n);
    double nb = sqrt(dotd(b, b, n));
    if (nb == 0.0) nb = 1.0;
    if (o->x0_mode == 0) {
        memcpy(x, b, sizeof(double) * n);
        trsv_f32L_f64(L, n, ldl, x, 0);
        trsv_f32L_f64(L, n, ldl, x, 1);
    } else {
        memset(x, 0, sizeof(double) * n);
    }
    int k = 0, ret = -1;
    for (;;) {
        orc_symv_f32_f64(A, n, lda, x, r);
        for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
        const double res = sqrt(dotd(r, r, n)) / nb;
        if (hist) hist[k] = res;
        if (relres) *relres = res;
        if (res <= o->tol) { ret = k; break; }
        if (k == o->maxit) break;
        trsv_f32L_f64(L, n, ldl, r, 0);
        trsv_f32L_f64(L, n, ldl, r, 1);
        for (int64_t i = 0; i < n; ++i) x[i] += r[i];
        ++k;
    }
    free(r);
    return ret;
}

/* s x s helpers of the block PCG (column-major, FP64).  spd_solve:
   X = G^-1 H through the Cholesky factor of G's lower triangle; returns 1
   (and X = 0) if a pivot is not above 1e-13 times its diagonal entry: G is
   singular to working precision and the block recurrence has broken down. */
static int spd_solve(const double* G, const double* H, double* X, int s) {
    double* C = (double*)malloc(sizeof(double) * s * s);
    int bad = 0;
    for (int j = 0; j < s; ++j) {
        const double g = G[j + j * s];
        double d = g;
        for (int k = 0; k < j; ++k) d -= C[j + k * s] * C[j + k * s];
        if (!(d > 1e-13 * g) || !(d < INFINITY)) { bad = 1; break; }
        d = sqrt(d);
        C[j + j * s] = d;
        for (int i = j + 1; i < s; ++i) {
            double v = G[i + j * s];
            for (int k = 0; k < j; ++k) v -= C[i + k * s] * C[j + k * s];
            C[i + j * s] = v / d;
        }
    }
    for (int c = 0; c < s; ++c) {
        double* x = X + c * s;
        if (bad) { for (int i = 0; i < s; ++i) x[i] = 0.0; continue; }
        for (int i = 0; i < s; ++i) {               /* C y = h */
            double v = H[i + c * s];
            for (int k = 0; k < i; ++k) v -= C[i + k * s] * x[k];
            x[i] = v / C[i + i * s];
        }
        for (int i = s - 1; i >= 0; --i) {          /* C^T x = y */
            double v = x[i];
            for (int k = i + 1; k < s; ++k) v -= C[k + i * s] * x[k];
            x[i] = v / C[i + i * s];
        }
    }
    free(C);
    return bad;
}

/* G = X^T Y for n x s blocks (leading dimension n) */
static void gram(const double* X, const double* Y, int64_t n, int s, double* G) {
    for (int b = 0; b < s; ++b)
        for (int a = 0; a < s; ++a) G[a + b * s] = dotd(X + a * n, Y + b * n, n);
}

/* Y (+)= M C for n x s blocks M, Y and an s x s C; Z (optional) replaces Y's
   old value: Y = Z + M C.  In place on Y == M (each row read first). */
static void block_axpy(const double* M, const double* Cm, double a, const double* Z, double* Y,
                       int64_t n, int s) {
    double* row = (double*)malloc(sizeof(double) * s);
    for (int64_t r = 0; r < n; ++r) {
        for (int k = 0; k < s; ++k) row[k] = M[r + k * n];
        for (int j = 0; j < s; ++j) {
            double v = 0.0;
            for (int k = 0; k < s; ++k) v += row[k] * Cm[k + j * s];
            const double base = Z ? Z[r + j * n] : Y[r + j * n];
            Y[r + j * n] = base + a * v;
        }
    }
    free(row);
}

static void prec_block(const float* L, int64_t ldl, int64_t n, int s, double* Z) {
    for (int j = 0; j < s; ++j) {
        trsv_f32L_f64(L, n, ldl, Z + j * n, 0);
        trsv_f32L_f64(L, n, ldl, Z + j * n, 1);
    }
}

/* Block PCG (O'Leary, "The block conjugate gradient algorithm and related
   methods", Linear Algebra Appl. 29, 1980, preconditioned form) over the s
   right-hand sides together, as the block counterpart of the single-vector
   loop of mixed_cholesky.cu:180-217:
     R = B - A X, Z = M^-1 R, P = Z, G = Z^T R
     repeat: Q = A P; alpha = (P^T Q)^-1 G; X += P alpha; R -= Q alpha;
             Z = M^-1 R; Gn = Z^T R; beta = G^-1 Gn; P = Z + P beta; G = Gn
   Stops when every column's recursive residual ||R_j|| / ||B_j|| <= tol and
   the true residual confirms it (a lagging true residual restarts the
   recurrence from X).  B, X: n x s, leading dimension n.  Returns iterations
   if converged, -1 at maxit, -2 if a Cholesky of P^T Q or G fails (the block
   recurrence broke down; X holds the last iterate).  relres (s values)
   receives the final true relative residuals. */
int orc_block_pcg(const float* A, int64_t lda, const float* L, int64_t ldl, int64_t n, int s,
                  const double* B, double* X, const orc_pcg_opts* o, double* relres) {
    const int64_t ns = n * (int64_t)s;
    double* R = (double*)malloc(sizeof(double) * ns);
    double* Z = (double*)malloc(sizeof(double) * ns);
    double* P = (double*)malloc(sizeof(double) * ns);
    double* Q = (double*)malloc(sizeof(double) * ns);
    double* G = (double*)malloc(sizeof(double) * s * s);
    double* Gn = (double*)malloc(sizeof(double) * s * s);
    double* PQ = (double*)malloc(sizeof(double) * s * s);
    double* al = (double*)malloc(sizeof(double) * s * s);
    double* nb = (double*)malloc(sizeof(double) * s);
    for (int j = 0; j < s; ++j) {
        nb[j] = sqrt(dotd(B + j * n, B + j * n, n));
        if (nb[j] == 0.0) nb[j] = 1.0;
    }
    if (o->x0_mode == 0) {
        memcpy(X, B, sizeof(double) * ns);
        prec_block(L, ldl, n, s, X);
    } else {
        memset(X, 0, sizeof(double) * ns);
    }
    int k = 0, ret = -1, restart = 1;
    for (;;) {
        if (restart) {
            for (int j = 0; j < s; ++j) orc_symv_f32_f64(A, n, lda, X + j * n, R + j * n);
            for (int64_t i = 0; i < ns; ++i) R[i] = B[i] - R[i];
            memcpy(Z, R, sizeof(double) * ns);
            prec_block(L, ldl, n, s, Z);
            memcpy(P, Z, sizeof(double) * ns);
            gram(Z, R, n, s, G);
            restart = 0;
        }
        double res = 0.0;
        for (int j = 0; j < s; ++j) {
            const double rj = sqrt(dotd(R + j * n, R + j * n, n)) / nb[j];
            res = rj > res ? rj : res;
        }
        if (res <= o->tol) {                     /* confirm on the true residual */
            double tr = 0.0;
            for (int j = 0; j < s; ++j) orc_symv_f32_f64(A, n, lda, X + j * n, Q + j * n);
            for (int j = 0; j < s; ++j) {
                double q = 0.0;
                for (int64_t i = 0; i < n; ++i) {
                    const double d = B[i + j * n] - Q[i + j * n];
                    q += d * d;
                }
                q = sqrt(q) / nb[j];
                tr = q > tr ? q : tr;
            }
            if (tr <= o->tol) { ret = k; break; }
            restart = 1;
            continue;
        }
        if (k == o->maxit) break;
        ++k;
        for (int j = 0; j < s; ++j) orc_symv_f32_f64(A, n, lda, P + j * n, Q + j * n);
        gram(P, Q, n, s, PQ);
        if (spd_solve(PQ, G, al, s)) { ret = -2; break; }
        block_axpy(P, al, 1.0, NULL, X, n, s);
        block_axpy(Q, al, -1.0, NULL, R, n, s);
        memcpy(Z, R, sizeof(double) * ns);
        prec_block(L, ldl, n, s, Z);
        gram(Z, R, n, s, Gn);
        if (spd_solve(G, Gn, al, s)) { ret = -2; break; }
        block_axpy(P, al, 1.0, Z, P, n, s);
        memcpy(G, Gn, sizeof(double) * s * s);
    }
    if (relres) {
        for (int j = 0; j < s; ++j) orc_symv_f32_f64(A, n, lda, X + j * n, Q + j * n);
        for (int j = 0; j < s; ++j) {
            double q = 0.0;
            for (int64_t i = 0; i < n; ++i) {
                const double d = B[i + j * n] - Q[i + j * n];
                q += d * d;
            }
            relres[j] = sqrt(q) / nb[j];
        }
    }
    free(R); free(Z); free(P); free(Q); free(G); free(Gn); free(PQ); free(al); free(nb);
    return ret;
}

/* Sturm-sequence bisection for the extreme eigenvalues of the Lanczos
   tridiagonal built from CG coefficients (Golub & Van Loan 10.2). */
static int sturm_count(const double* dg, const double* off2, int k, double x) {
    int c = 0;
    double q = dg[0] - x;
    if (q < 0) c++;
    for (int i = 1; i < k; ++i) {
        if (q == 0.0) q = 1e-300;
        q = dg[i] - x - off2[i - 1] / q;
        if (q < 0) c++;
    }
    return c;
}
double orc_lanczos_cond(const double* alphas, const double* betas, int k) {
    if (k < 1) return 1.0;
    double* dg = (double*)malloc(sizeof(double) * k);
    double* off2 = (double*)malloc(sizeof(double) * (k > 1 ? k - 1 : 1));
    for (int i = 0; i < k; ++i) {
        dg[i] = 1.0 / alphas[i] + (i > 0 ? betas[i - 1] / alphas[i - 1] : 0.0);
        if (i > 0) off2[i - 1] = betas[i - 1] / (alphas[i - 1] * alphas[i - 1]);
    }
    double lo = INFINITY, hi = -INFINITY;
    for (int i = 0; i < k; ++i) {
        double rad = (i > 0 ? sqrt(off2[i - 1]) : 0) + (i < k - 1 ? sqrt(off2[i]) : 0);
        if (dg[i] - rad < lo) lo = dg[i] - rad;
        if (dg[i] + rad > hi) hi = dg[i] + rad;
    }
    double a = lo, b = hi;       /* smallest: count(x) >= 1 */
    for (int it = 0; it < 200; ++it) {
        double m = 0.5 * (a + b);
        if (sturm_count(dg, off2, k, m) >= 1) b = m; else a = m;
    }
    double emin = 0.5 * (a + b);
    a = lo; b = hi;              /* largest: count(x) >= k */
    for (int it = 0; it < 200; ++it) {
        double m = 0.5 * (a + b);
        if (sturm_count(dg, off2, k, m) >= k) b = m; else a = m;
    }
    double emax = 0.5 * (a + b);
    free(dg); free(off2);
    return emin > 0 ? emax / emin : INFINITY;
}
