// k_potrf.hip -- FP32 Cholesky of one diagonal block (bs <= 256) in a single
// workgroup, plus the inverses of its 32x32 diagonal sub-blocks.
//
// Replaces the cusolverDnSpotrf call of switching_chol.cuh:86-95 (CPU form:
// micro_cholesky, micro_chol.hpp:14-44).  The block lives in LDS as 32x32
// sub-tiles (lower triangle only, 36 sub-tiles x 32x33 floats = 152 KB for
// bs = 256); the algorithm is right-looking over 32-column sub-panels:
//   1. wave 0 factors the diagonal sub-tile (pipelined right-looking, row
//      per lane) and inverts it (column per lane, rows of L broadcast from
//      LDS);
//   2. the sub-tiles below are multiplied by that inverse (MFMA 32x32x2 f32);
//   3. the trailing sub-tiles receive the rank-32 update (MFMA 32x32x2 f32):
//      sub-column kb+1 at once, the others while wave 0 factors sub-block
//      kb+1 (they are disjoint from it).
// Outputs: L in place (lower; the strict upper triangle of A is untouched),
// dinv (column-major 32x32 inverse per sub-block), the running minimum pivot
// squared and, on a non-positive pivot, the 1-based failing column.
#include "mpc_internal.hpp"

namespace {

constexpr int SUBLD = 33;                 // padded sub-tile column stride
constexpr int SUBSZ = 32 * SUBLD;         // floats per sub-tile
constexpr int MAXSB = 8;                  // 256 / 32
constexpr int NSUB = MAXSB * (MAXSB + 1) / 2;
constexpr int LRLD = 36;                  // row stride of Lr (16-byte aligned rows)
// Four waves (one per SIMD) at about 200 registers each: the factor then
// fits on a CU beside one workgroup of the concurrent trailing update.
// Eight waves are faster alone but need a whole CU, which the update's
// stream of workgroups rarely leaves free (the look-ahead's critical path
// then waits for one).
constexpr int NW = 4;
constexpr int NT = 64 * NW;
// one refinement step after the product with the explicit 32x32 inverse
constexpr bool kPotrfRefine = false;

typedef float v16f __attribute__((ext_vector_type(16)));
// LDS (address space 3) pointers: an opaque copy stays a 32-bit LDS address
typedef const __attribute__((address_space(3))) float* lds_cfloat_p;
typedef float f4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(3))) f4* lds_cf4_p;

__device__ __forceinline__ int sidx(int bi, int bj) { return bi * (bi + 1) / 2 + bj; }

// row of accumulator register i in lane half h (32x32 MFMA output layout)
__device__ __forceinline__ int sig(int i, int h) { return 8 * (i >> 2) + 4 * h + (i & 3); }

__device__ __forceinline__ float rdlane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// D[m][n] (32x32) += A[m][k] * B[k][n] over k = 0..31 with both operands in LDS:
// A(m,k) = a[m + k*lda_], B(k,n) = b[n + k*ldb_] (i.e. B stored as rows n).
__device__ __forceinline__ v16f mfma_32x32x32_lds(v16f acc, const float* a, int lda_,
                                                 const float* b, int ldb_, float asign) {
    const int lane = threadIdx.x & 63;
    const int m = lane & 31, h = lane >> 5;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        const int k = 2 * s + h;
        float av = asign * a[m + k * lda_];
        float bv = b[m + k * ldb_];
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

// S(ib, jb) -= S(ib, kb) * S(jb, kb)^T for one 32x32 sub-tile (one wave, MFMA)
__device__ __forceinline__ void update_subtile(float* T, int ib, int jb, int kb) {
    const int lane = threadIdx.x & 63;
    float* Cs = T + sidx(ib, jb) * SUBSZ;
    const float* Ai = T + sidx(ib, kb) * SUBSZ;
    const float* Bj = T + sidx(jb, kb) * SUBSZ;
    const int n = lane & 31, h = lane >> 5;
    v16f acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = Cs[sig(i, h) + n * SUBLD];
    // A(m=r,k=q) = -Ai(r,q); B(k=q,n=c) = Bj(c,q)
    acc = mfma_32x32x32_lds(acc, Ai, SUBLD, Bj, SUBLD, -1.f);
#pragma unroll
    for (int i = 0; i < 16; ++i) Cs[sig(i, h) + n * SUBLD] = acc[i];
}

// Columns [c_lo, c_hi) of the lower block into the sub-tiles (identity
// beyond bs), 16-byte path: wave w of nwav reads columns c_lo + w, +nwav, ...
// (LB of them per batch, all loads of a batch in flight), lane l rows
// 4l..4l+3.  Needs a 16-byte aligned block with lda and bs multiples of 4;
// rows above the diagonal are read and dropped.
template <int LB>
__device__ __forceinline__ void load_cols_vec(float* T, const float* __restrict__ A, int64_t lda,
                                              int bs, int nsb, int c_lo, int c_hi, int w,
                                              int nwav) {
    const int lane = threadIdx.x & 63;
    const int r4 = 4 * lane, rb = r4 >> 5;
    for (int cbase = c_lo; cbase < c_hi; cbase += nwav * LB) {
        float4 v[LB];
#pragma unroll
        for (int u = 0; u < LB; ++u) {
            // clamped, unconditional load (selected below): a load under a
            // per-lane condition becomes a branch with its own wait
            const int c = cbase + nwav * u + w;
            const int cc = c < bs ? c : bs - 1, rc = r4 < bs ? r4 : bs - 4;
            const float4 x = *reinterpret_cast<const float4*>(A + rc + (int64_t)cc * lda);
            v[u] = (c < bs && r4 < bs) ? x : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < LB; ++u) {
            const int c = cbase + nwav * u + w, cb = c >> 5;
            if (c >= c_hi || rb >= nsb || rb < cb) continue;
            const float e[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = r4 + q;
                float x = e[q];
                if (r < c) x = 0.f;
                if (r >= bs || c >= bs) x = (r == c) ? 1.f : 0.f;
                T[sidx(rb, cb) * SUBSZ + (r & 31) + (c & 31) * SUBLD] = x;
            }
        }
    }
}

// Sub-column cb (its lower part, rows cb*32 .. bs-1) back to A by threads
// t0 .. t0+nthr-1: a thread per row, consecutive threads consecutive rows.
__device__ __forceinline__ void store_subcol(const float* T, float* __restrict__ A, int64_t lda,
                                             int bs, int cb, int t0, int nthr) {
    const int c0 = cb * 32, c1 = bs < c0 + 32 ? bs : c0 + 32;
    for (int r = c0 + (int)threadIdx.x - t0; r < bs; r += nthr) {
        const int cend = r + 1 < c1 ? r + 1 : c1;        // lower triangle only
        const float* Ts = T + sidx(r >> 5, cb) * SUBSZ + (r & 31);
        float* Ar = A + r;
        for (int c = c0; c < cend; ++c) Ar[(int64_t)c * lda] = Ts[(c & 31) * SUBLD];
    }
}

__global__ __launch_bounds__(NT) void potrf_kernel(float* __restrict__ A, int64_t lda, int bs,
                                                    int col0, float* __restrict__ dinv,
                                                    mpc_small* __restrict__ sm, int ntiles_next,
                                                    uint32_t* __restrict__ st_amax,
                                                    uint32_t* __restrict__ st_maxrs) {
    __shared__ float T[NSUB * SUBSZ];
    __shared__ float Dinv[SUBSZ];
    __shared__ float colb[2][32];                   // column k of L, broadcast
    __shared__ __align__(16) float Lr[33 * LRLD];   // row-major sub-tile; row 32: 1 / L_kk
    __shared__ int fail_col;
    __shared__ float red[256];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (sm->info != 0) return;                       // an earlier step failed
    for (int t = tid; t < ntiles_next; t += NT) {
        st_amax[t] = 0u;
        st_maxrs[t] = 0u;
    }
    if (tid == 0) fail_col = -1;
    const int nsb = (bs + 31) >> 5;

    // ---- load lower sub-tiles (padding: identity beyond bs), with many loads
    // per thread in flight.  16-byte path: sub-column 0 now; the other
    // sub-columns are read by waves 1.. while wave 0 factors sub-block 0.
    const bool vec = ((reinterpret_cast<uintptr_t>(A) & 15) == 0) && (lda % 4 == 0) && (bs % 4 == 0);
    if (vec) {
        load_cols_vec<8>(T, A, lda, bs, nsb, 0, 32, wave, NW);
    } else {
        constexpr int LB = 32;
        const int r = tid & 255, rb = r >> 5, half = tid >> 8;
        const int ncol = nsb * 32;
        for (int c0 = half * LB; c0 < ncol; c0 += (NT / 256) * LB) {
            float v[LB];
#pragma unroll
            for (int u = 0; u < LB; ++u) {
                const int c = c0 + u;
                const float y = A[(r < bs ? r : bs - 1) + (int64_t)(c < bs ? c : bs - 1) * lda];
                v[u] = (r < bs && c < bs && r >= c && rb < nsb) ? y : ((r == c) ? 1.f : 0.f);
            }
#pragma unroll
            for (int u = 0; u < LB; ++u) {
                const int c = c0 + u, cb = c >> 5;
                if (rb < nsb && rb >= cb)
                    T[sidx(rb, cb) * SUBSZ + (r & 31) + (c & 31) * SUBLD] = (r >= c) ? v[u] : 0.f;
            }
        }
    }
    mpc_lds_barrier();

    for (int kb = 0; kb < nsb; ++kb) {
        float* D = T + sidx(kb, kb) * SUBSZ;
        // ---- 1. factor + invert the diagonal sub-tile (wave 0); meanwhile
        // waves 1..3 finish the trailing update of kb-1 on the sub-columns
        // beyond kb (3b below), which the factor of sub-block kb does not read
        if (wave != 0 && kb == 0 && vec)                 // the rest of the block
            load_cols_vec<16>(T, A, lda, bs, nsb, 32, nsb * 32, wave - 1, NW - 1);
        if (wave != 0 && kb >= 1) {
            const int m = nsb - 1 - kb;                  // sub-columns kb+1 .. nsb-1
            const int npair = m * (m + 1) / 2;
            for (int pi = wave - 1; pi < npair; pi += NW - 1) {
                int ii = 0;
                while ((ii + 1) * (ii + 2) / 2 <= pi) ++ii;
                const int jj = pi - ii * (ii + 1) / 2;
                update_subtile(T, kb + 1 + ii, kb + 1 + jj, kb - 1);
            }
            // sub-column kb-1 is final: back to A while the factor runs
            store_subcol(T, A, lda, bs, kb - 1, 64, NT - 64);
        }
        if (wave == 0) {
            // Pipelined right-looking factor, row per lane (lanes 32..63 mirror
            // 0..31).  Column k+1 receives column k's update through one
            // readlane (the pivot chain); the other columns receive it one step
            // later from an LDS broadcast read right after it is written, so
            // the read latency overlaps the next pivot.  The opaque copy
            // of the lane id keeps the lane-dependent predicates inside the kb
            // loop (hoisted out of it they exhaust the scalar registers).
            int r = lane & 31;
            asm volatile("" : "+v"(r));
            // Updates are not predicated on r >= c: entries above the diagonal
            // only accumulate values no valid entry ever reads (they are not
            // broadcast, not pivots, and zeroed when the row is written back).
            float d[32];
#pragma unroll
            for (int c = 0; c < 32; ++c) {
                const float v = D[r + c * SUBLD];
                d[c] = (c <= r) ? v : 0.f;
            }
            int bad = -1;
            float lc[32];          // column k-1 of L, prefetched from the broadcast
#pragma unroll
            for (int k = 0; k < 32; ++k) {
                float piv = rdlane(d[k], k);
                if (!(piv > 0.f)) {                  // uniform across the wave
                    if (bad < 0) bad = k;
                    piv = 1.f;
                }
                // 1/sqrt(piv) from the hardware estimate and one Newton step
                // (about 1 ulp), sqrt(piv) = piv / sqrt(piv): a short
                // dependent chain per pivot, where a correctly rounded sqrt
                // and reciprocal take four times the instructions
                float rl = __builtin_amdgcn_rsqf(piv);
                const float hr = 0.5f * piv * rl;
                rl = fmaf(fmaf(-hr, rl, 0.5f), rl, rl);
                const float l = piv * rl;
                d[k] = (r == k) ? l : d[k] * rl;
                if (r == k) Lr[32 * LRLD + k] = rl;
                if (k < 31) {
                    if (lane < 32) colb[k & 1][r] = d[k];
                    const float v = rdlane(d[k], k + 1);   // l_{k+1,k}: the pivot chain
                    d[k + 1] = fmaf(-d[k], v, d[k + 1]);
                }
                if (k >= 1) {                         // deferred: column k-1 -> columns > k
                    const float lr = d[k - 1];
#pragma unroll
                    for (int c = k + 1; c < 32; ++c) d[c] = fmaf(-lr, lc[c], d[c]);
                }
                if (k < 30) {                         // prefetch column k for step k+1
#pragma unroll
                    for (int c = k + 2; c < 32; ++c) lc[c] = colb[k & 1][c];
                }
            }
#pragma unroll
            for (int c = 0; c < 32; ++c) d[c] = (c <= r) ? d[c] : 0.f;
            // the whole sub-tile, zeros above the diagonal: the trailing
            // updates of earlier kb wrote the full 32x32 block, and the
            // refinement below reads D as a lower-triangular operand; Lr keeps
            // a row-major copy for the inverse's broadcast reads
            if (lane < 32) {
#pragma unroll
                for (int c = 0; c < 32; ++c) D[r + c * SUBLD] = d[c];
#pragma unroll
                for (int c = 0; c < 32; c += 4)
                    *reinterpret_cast<float4*>(&Lr[r * LRLD + c]) =
                        make_float4(d[c], d[c + 1], d[c + 2], d[c + 3]);
            }
            if (bad >= 0 && lane == 0 && kb * 32 + bad < bs && fail_col < 0)
                fail_col = kb * 32 + bad;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            // inverse, column per lane: x[p] = (e_c[p] - sum_{q<p} L[p][q] x[q]) / L[p][p],
            // row p of L read by every lane at the same LDS address (broadcast)
            // (Lr lies beyond the 64 KB reach of the LDS instruction offset:
            // an opaque base keeps every read at base + immediate offset)
            const int c = r;
            lds_cfloat_p lrb = (lds_cfloat_p)(Lr);
            asm volatile("" : "+v"(lrb));
            float x[32];
#pragma unroll
            for (int p = 0; p < 32; ++p) {
                // row p's reads wait for x[p-2]: without this order the
                // scheduler hoists all 528 reads of L to the top, which
                // exhausts the registers and spills to scratch memory
                if (p >= 2) asm volatile("" : "+v"(lrb) : "v"(x[p - 2]));
                float a0 = (p == c) ? 1.f : 0.f, a1 = 0.f;
#pragma unroll
                for (int q4 = 0; q4 < p; q4 += 4) {
                    const f4 lp = *(lds_cf4_p)(&lrb[p * LRLD + q4]);
                    a0 = fmaf(-lp.x, x[q4], a0);
                    if (q4 + 1 < p) a1 = fmaf(-lp.y, x[q4 + 1], a1);
                    if (q4 + 2 < p) a0 = fmaf(-lp.z, x[q4 + 2], a0);
                    if (q4 + 3 < p) a1 = fmaf(-lp.w, x[q4 + 3], a1);
                }
                x[p] = (a0 + a1) * lrb[32 * LRLD + p];
            }
            if (lane < 32) {
                float* gd = dinv + (int64_t)((col0 >> 5) + kb) * 1024;
#pragma unroll
                for (int p = 0; p < 32; ++p) {
                    Dinv[p + c * SUBLD] = x[p];
                    if (kb * 32 < bs) gd[p + c * 32] = x[p];
                }
            }
        }
        mpc_lds_barrier();

        // ---- 2. sub-tiles below: X = S * Dinv^T, refined once:
        //   R = S - X D^T, X += R Dinv^T  (D the factored diagonal sub-tile,
        //   zero above its diagonal).  Worked in the transposed domain: lane
        //   (n, h) register i holds X(n, sig(i, h)), the k index of MFMA step t
        //   is sig(t, h) on both operands, so no LDS round trip is needed.
        for (int ib = kb + 1 + wave; ib < nsb; ib += NW) {
            float* S = T + sidx(ib, kb) * SUBSZ;
            const int n = lane & 31, h = lane >> 5;
            v16f s0, x = {};
#pragma unroll
            for (int i = 0; i < 16; ++i) s0[i] = S[n + sig(i, h) * SUBLD];
#pragma unroll
            for (int t = 0; t < 16; ++t)
                x = __builtin_amdgcn_mfma_f32_32x32x2f32(Dinv[n + sig(t, h) * SUBLD], s0[t], x,
                                                         0, 0, 0);
            if constexpr (kPotrfRefine) {
                v16f r = s0;
#pragma unroll
                for (int t = 0; t < 16; ++t)
                    r = __builtin_amdgcn_mfma_f32_32x32x2f32(-D[n + sig(t, h) * SUBLD], x[t], r, 0,
                                                             0, 0);
#pragma unroll
                for (int t = 0; t < 16; ++t)
                    x = __builtin_amdgcn_mfma_f32_32x32x2f32(Dinv[n + sig(t, h) * SUBLD], r[t], x,
                                                             0, 0, 0);
            }
#pragma unroll
            for (int i = 0; i < 16; ++i) S[n + sig(i, h) * SUBLD] = x[i];
        }
        mpc_lds_barrier();

        // ---- 3a. trailing update of sub-column kb+1 (what sub-block kb+1
        // needs next): S(ib,kb+1) -= S(ib,kb) S(kb+1,kb)^T, ib > kb.  The other
        // sub-columns (3b) are updated during the next factor step.
        for (int ib = kb + 1 + wave; ib < nsb; ib += NW) update_subtile(T, ib, kb + 1, kb);
        mpc_lds_barrier();
    }

    // ---- write back the last sub-column (the others went back during the
    // factor), min pivot^2, failure
    store_subcol(T, A, lda, bs, nsb - 1, 0, NT);
    float pmin = INFINITY;
    if (tid < bs) {
        const float d = T[sidx(tid >> 5, tid >> 5) * SUBSZ + (tid & 31) * (1 + SUBLD)];
        pmin = d * d;
    }
    if (tid < 256) red[tid] = pmin;
    mpc_lds_barrier();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) red[tid] = fminf(red[tid], red[tid + w]);
        mpc_lds_barrier();
    }
    if (tid == 0) {
        if (fail_col >= 0) {
            sm->info = col0 + fail_col + 1;
        } else {
            atomicMin(&sm->mu_bits, __float_as_uint(red[0]));
        }
    }
}

}  // namespace

int mpc_launch_potrf(hipStream_t st, float* A, int64_t lda, int bs, int col0, float* dinv,
                     mpc_small* sm, int ntiles_next, uint32_t* stats_amax, uint32_t* stats_maxrs) {
    if (bs <= 0 || bs > 256) {
        mpc_set_error("potrf: block size %d outside 1..256", bs);
        return MPC_ERR_ARG;
    }
    if (ntiles_next > 1024) {
        mpc_set_error("potrf: %d panel tiles exceed 1024", ntiles_next);
        return MPC_ERR_ARG;
    }
    hipLaunchKernelGGL(potrf_kernel, dim3(1), dim3(NT), 0, st, A, lda, bs, col0, dinv, sm,
                       ntiles_next, stats_amax, stats_maxrs);
    MPC_LAUNCH_CHECK();
    return 0;
}
