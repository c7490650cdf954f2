// k_potrf.hip -- FP32 Cholesky of one diagonal block (bs <= 256) in a single
// workgroup, plus the inverses of its 32x32 diagonal sub-blocks.
//
// Replaces the cusolverDnSpotrf call of switching_chol.cuh:86-95 (CPU form:
// micro_cholesky, micro_chol.hpp:14-44).  The block lives in LDS as 32x32
// sub-tiles (lower triangle only, 36 sub-tiles x 32x33 floats = 152 KB for
// bs = 256); the algorithm is right-looking over 32-column sub-panels, with
// only what the next diagonal sub-block needs on wave 0's path:
//   1. wave 0 factors the diagonal sub-tile kb (pipelined right-looking, row
//      per lane) and inverts it (column per lane, rows of L broadcast from
//      LDS);
//   2. the sub-tiles below are multiplied by that inverse (MFMA 32x32x2 f32):
//      wave 0 takes sub-tile kb+1 and at once updates diagonal sub-tile
//      kb+1 with it, the other waves the rest;
//   3. waves 1..3 apply the rank-32 update to the other trailing sub-tiles
//      and write sub-column kb back to A while wave 0 factors sub-block kb+1
//      (they are disjoint from it).  Two workgroup barriers per sub-block.
// Outputs: L in place (lower; the strict upper triangle of A is untouched),
// dinv (column-major 32x32 inverse per sub-block), the running minimum pivot
// squared and, on a non-positive pivot, the 1-based failing column.
#include "mpc_internal.hpp"

namespace {

constexpr int SUBLD = 33;                 // padded sub-tile column stride
constexpr int SUBSZ = 32 * SUBLD;         // floats per sub-tile
constexpr int MAXSB = 8;                  // 256 / 32
constexpr int NSUB = MAXSB * (MAXSB + 1) / 2;
// Four waves (one per SIMD) at about 200 registers each: the factor then
// fits on a CU beside one workgroup of the concurrent trailing update.
// Eight waves are faster alone but need a whole CU, which the update's
// stream of workgroups rarely leaves free (the look-ahead's critical path
// then waits for one).
constexpr int NW = 4;
constexpr int NT = 64 * NW;

typedef float v16f __attribute__((ext_vector_type(16)));
// LDS (address space 3) pointer: an opaque copy stays a 32-bit LDS address
typedef float f4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(3))) f4* lds_cf4_p;

__device__ __forceinline__ int sidx(int bi, int bj) { return bi * (bi + 1) / 2 + bj; }

// row of accumulator register i in lane half h (32x32 MFMA output layout)
__device__ __forceinline__ int sig(int i, int h) { return 8 * (i >> 2) + 4 * h + (i & 3); }

__device__ __forceinline__ float rdlane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// lane l of half h (lanes 32 h .. 32 h + 31) to lanes l and l + 32: the half
// swap exchanges lanes 32..63 of its first operand with lanes 0..31 of its
// second; with both operands v, the first result holds the low half twice
// and the second the high half twice
__device__ __forceinline__ float bcast_half(float v, int h) {
    const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false,
                                                    false);
    return __uint_as_float(h ? p[1] : p[0]);
}

// D[m][n] (32x32) += A[m][k] * B[k][n] over k = 0..31 with both operands in LDS:
// A(m,k) = a[m + k*lda_], B(k,n) = b[n + k*ldb_] (i.e. B stored as rows n).
__device__ __forceinline__ v16f mfma_32x32x32_lds(v16f acc, const float* a, int lda_,
                                                 const float* b, int ldb_) {
    const int lane = threadIdx.x & 63;
    const int m = lane & 31, h = lane >> 5;
    // every operand read issued before the first product: interleaved, each
    // pair of products waits a full LDS round trip
    float av[16], bv[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        const int k = 2 * s + h;
        av[s] = a[m + k * lda_];
        bv[s] = b[m + k * ldb_];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int s = 0; s < 16; ++s)
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[s], bv[s], acc, 0, 0, 0);
    return acc;
}

// S(ib, jb) -= S(ib, kb) * S(jb, kb)^T for one 32x32 sub-tile (one wave, MFMA)
__device__ __forceinline__ void update_subtile(float* T, int ib, int jb, int kb) {
    const int lane = threadIdx.x & 63;
    float* Cs = T + sidx(ib, jb) * SUBSZ;
    const float* Ai = T + sidx(ib, kb) * SUBSZ;
    const float* Bj = T + sidx(jb, kb) * SUBSZ;
    const int n = lane & 31, h = lane >> 5;
    // -C += Ai Bj^T, negated on the way in and out (a negated operand would
    // put a dependent VALU op between consecutive products)
    v16f acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = -Cs[sig(i, h) + n * SUBLD];
    // A(m=r,k=q) = Ai(r,q); B(k=q,n=c) = Bj(c,q)
    acc = mfma_32x32x32_lds(acc, Ai, SUBLD, Bj, SUBLD);
#pragma unroll
    for (int i = 0; i < 16; ++i) Cs[sig(i, h) + n * SUBLD] = -acc[i];
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

// The lower sub-tiles right of sub-column 0 ((ib, cb), 1 <= cb <= ib < nsb,
// numbered by sub-column, then row) into T (identity beyond bs, zeros above
// the diagonal), 16-byte path: wave w of nwav reads tiles w, w + nwav, ...,
// TB of them per batch with all loads of a batch in flight; in a tile, load q
// of lane l covers column (l + 64 q) / 8 and rows 4 ((l + 64 q) % 8) .. +3
// (eight lanes per 128-byte column segment).  Only the lower part is read.
template <int TB>
__device__ __forceinline__ void load_tiles_vec(float* T, const float* __restrict__ A, int64_t lda,
                                               int bs, int nsb, int w, int nwav) {
    const int lane = threadIdx.x & 63;
    const int ntile = nsb * (nsb - 1) / 2;
    for (int t0 = w; t0 < ntile; t0 += nwav * TB) {
        float4 v[TB][4];
        int tib[TB], tcb[TB];
#pragma unroll
        for (int u = 0; u < TB; ++u) {
            // clamped tile index and addresses: unconditional loads (a load
            // under a condition becomes a branch with its own wait)
            int rem = t0 + nwav * u < ntile ? t0 + nwav * u : ntile - 1, cb = 1;
            while (rem >= nsb - cb) {
                rem -= nsb - cb;
                ++cb;
            }
            tib[u] = cb + rem;
            tcb[u] = cb;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int j = lane + 64 * q;
                const int c = cb * 32 + (j >> 3), r4 = tib[u] * 32 + (j & 7) * 4;
                const int cc = c < bs ? c : bs - 1, rc = r4 < bs ? r4 : bs - 4;
                v[u][q] = *reinterpret_cast<const float4*>(A + rc + (int64_t)cc * lda);
            }
        }
#pragma unroll
        for (int u = 0; u < TB; ++u) {
            if (t0 + nwav * u >= ntile) continue;
            float* Ts = T + sidx(tib[u], tcb[u]) * SUBSZ;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int j = lane + 64 * q;
                const int c = tcb[u] * 32 + (j >> 3), r4 = tib[u] * 32 + (j & 7) * 4;
                const float e[4] = {v[u][q].x, v[u][q].y, v[u][q].z, v[u][q].w};
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int r = r4 + i;
                    float x = (r < c) ? 0.f : e[i];
                    if (r >= bs || c >= bs) x = (r == c) ? 1.f : 0.f;
                    Ts[(r & 31) + (c & 31) * SUBLD] = x;
                }
            }
        }
    }
}

// Sub-column cb (its lower part, rows cb*32 .. bs-1) back to A by threads
// t0 .. t0+nthr-1: a thread per row, consecutive threads consecutive rows.
__device__ __forceinline__ void store_subcol(const float* T, float* __restrict__ A, int64_t lda,
                                             int bs, int cb, int t0, int nthr) {
    const int c0 = cb * 32, c1 = bs < c0 + 32 ? bs : c0 + 32;
    // byte offsets from A in one 32-bit VGPR per store (a 64-bit address per
    // store would hold twice the registers); the launch checks that the
    // block's bs + bs lda elements stay below 2^31 bytes
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(A, 0, 0x7fffffff,
                                                                          0x00020000);
    const uint32_t ld32 = (uint32_t)lda;
    for (int r = c0 + (int)threadIdx.x - t0; r < bs; r += nthr) {
        // Every (r, c) is stored unconditionally: where c > r (diagonal
        // sub-tile) or c >= bs, the store goes to the lower-triangle entry
        // (max, min) of (r, min(c, c1 - 1)) with that entry's value, which
        // its own row also stores (the same value twice).  A per-column
        // predicate instead keeps 32 lane masks alive.  The row's 32 reads
        // are in flight at once, then the stores.
        const float* Tc = T + sidx(r >> 5, cb) * SUBSZ;
#pragma unroll
        for (int c8 = 0; c8 < 32; c8 += 16) {
            float v[16];
            uint32_t off[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const int cc = c0 + c8 + c < c1 ? c0 + c8 + c : c1 - 1;
                const int rs = r > cc ? r : cc, cs = r > cc ? cc : r;
                v[c] = Tc[(rs & 31) + (cs & 31) * SUBLD];
                off[c] = 4u * ((uint32_t)rs + (uint32_t)cs * ld32);
            }
#pragma unroll
            for (int c = 0; c < 16; ++c)
                __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v[c]), rsrc, off[c], 0, 0);
            __builtin_amdgcn_sched_barrier(0);      // halves kept apart (registers)
        }
    }
}

__global__ __launch_bounds__(NT) void potrf_kernel(float* __restrict__ A, int64_t lda, int bs,
                                                    int col0, float* __restrict__ dinv,
                                                    mpc_small* __restrict__ sm, int ntiles_next,
                                                    uint32_t* __restrict__ st_amax,
                                                    uint32_t* __restrict__ st_maxrs,
                                                    int mu_slot, int want_f2) {
    __shared__ float T[NSUB * SUBSZ];
    __shared__ float Dinv[SUBSZ];
    __shared__ __align__(16) float colb[2][32];     // column k of L by row parity
    __shared__ __align__(16) float Lc[32 * 32];     // L's columns, rows split by parity
    __shared__ int fail_col, fail_nan;
    __shared__ float red[256];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (sm->info != 0) return;                       // an earlier step failed
    for (int t = tid; t < ntiles_next; t += NT) {
        st_amax[t] = 0u;
        st_maxrs[t] = 0u;
    }
    if (tid == 0) { fail_col = -1; fail_nan = 0; }
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
        // waves 1..3 load the rest of the block (kb = 0) or finish the
        // trailing update of kb-1 (3, below: it does not touch sub-tile kb)
        if (wave != 0 && kb == 0 && vec)                 // the rest of the block
            load_tiles_vec<4>(T, A, lda, bs, nsb, wave - 1, NW - 1);
        if (wave == 0) {
            // Pipelined right-looking factor, row per lane, the row's columns
            // split over the two lane halves by parity: lane r + 32 h holds
            // e[u] = A[r][2u + h].  Column k+1 receives column k's update at
            // once (the pivot chain); the other columns receive it one step
            // later from an LDS broadcast read right after it is written, so
            // the read latency overlaps the next pivot.  Column k's entries
            // reach the other half through a half swap (one VALU op).
            // Updates are not predicated on r >= c: entries above the
            // diagonal only accumulate values no valid entry ever reads (they
            // are not broadcast, not pivots, and zeroed before the row is
            // written back).
            int r = lane & 31;
            asm volatile("" : "+v"(r));
            const bool hi = lane >= 32;
            float e[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int c = 2 * u + (hi ? 1 : 0);
                const float v = D[r + c * SUBLD];
                e[u] = (c <= r) ? v : 0.f;
            }
            f4 lcg[4];             // column k-1 of L at this lane's columns, by 4
            float lprev = 0.f;     // L[r][k-1]
            float rv = 1.f;        // lane r: 1 / L[r][r]
            float pv = 1.f;        // lane r: the pivot of column r as met
#pragma unroll
            for (int k = 0; k < 32; ++k) {
                const int hk = k & 1, uk = k >> 1;
                const float pr = rdlane(e[uk], k + 32 * hk);
                pv = (r == k) ? pr : pv;
                const float piv = (pr > 0.f) ? pr : 1.f;
                // 1/sqrt(piv) from the hardware estimate and one Newton step
                // (about 1 ulp), sqrt(piv) = piv / sqrt(piv): a short
                // dependent chain per pivot, where a correctly rounded sqrt
                // and reciprocal take four times the instructions
                float rl = __builtin_amdgcn_rsqf(piv);
                const float hr = 0.5f * piv * rl;
                rl = fmaf(fmaf(-hr, rl, 0.5f), rl, rl);
                const float l = piv * rl;
                rv = (r == k) ? rl : rv;
                const float ek = (r == k) ? l : e[uk] * rl;    // valid in half hk
                e[uk] = (hi == (hk == 1)) ? ek : e[uk];
                const float lk = bcast_half(ek, hk);           // L[r][k], both halves
                colb[k & 1][(r & 1) * 16 + (r >> 1)] = lk;
                if (k < 31) {
                    const int h1 = (k + 1) & 1, u1 = (k + 1) >> 1;
                    const float v = rdlane(ek, (k + 1) + 32 * hk);   // L[k+1][k]
                    const float t = fmaf(-lk, v, e[u1]);
                    e[u1] = (hi == (h1 == 1)) ? t : e[u1];
                }
                if (k >= 1) {                         // deferred: column k-1 -> columns > k
                    // whole groups live until here: a dead component's
                    // register reused early would wait for the read in flight
#pragma unroll
                    for (int g = (k >> 1) >> 2; g < 4; ++g) asm volatile("" : "+v"(lcg[g]));
                    float lc[16];
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        lc[4 * g] = lcg[g].x; lc[4 * g + 1] = lcg[g].y;
                        lc[4 * g + 2] = lcg[g].z; lc[4 * g + 3] = lcg[g].w;
                    }
#pragma unroll
                    for (int u = (k >> 1) + 1; u < 16; ++u) e[u] = fmaf(-lprev, lc[u], e[u]);
                    if ((k & 1) == 0) {               // column k+1 = 2 (k/2) + 1: half 1 only
                        const int u = k >> 1;
                        const float t = fmaf(-lprev, lc[u], e[u]);
                        e[u] = hi ? t : e[u];
                    }
                }
                if (k < 30) {                         // prefetch column k for step k+1
                    const f4* cb = reinterpret_cast<const f4*>(&colb[k & 1][hi ? 16 : 0]);
#pragma unroll
                    for (int g = ((k >> 1) + 1) >> 2; g < 4; ++g) lcg[g] = cb[g];
                    // the reads stay here, a step ahead of their use (the
                    // scheduler would sink them next to it)
                    __builtin_amdgcn_sched_barrier(0);
                }
                lprev = lk;
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) e[u] = (2 * u + (hi ? 1 : 0) <= r) ? e[u] : 0.f;
            // the whole sub-tile, zeros above the diagonal (the trailing
            // updates of earlier kb wrote the full 32x32 block); Lc keeps the
            // columns of L split by row parity for the inverse's broadcast
            // reads: Lc[q][h][v] = L[2v + h][q]
            {
                const int lo = (r & 1) * 16 + (r >> 1);
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int c = 2 * u + (hi ? 1 : 0);
                    D[r + c * SUBLD] = e[u];
                    Lc[c * 32 + lo] = e[u];
                }
            }
            {
                const uint64_t badm = __builtin_amdgcn_ballot_w64(!(pv > 0.f)) & 0xffffffffull;
                const uint64_t nanm = __builtin_amdgcn_ballot_w64(!(pv == pv)) & 0xffffffffull;
                const int bad = badm ? __builtin_ctzll(badm) : -1;
                if (bad >= 0 && lane == 0 && kb * 32 + bad < bs && fail_col < 0) {
                    fail_col = kb * 32 + bad;
                    fail_nan = (int)((nanm >> bad) & 1u);     // the failing pivot is NaN
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            // inverse, right-looking, column c = r per lane, rows split over
            // the halves by parity (acc[v] = X[2v + h][c]):
            //   x_q = acc_q / L[q][q];  acc_p -= L[p][q] x_q  (p > q)
            // x_q reaches the other half through a half swap; column q of L
            // is read from Lc by both halves (one address per half), one step
            // ahead.
            const int c = r;
            float acc[16];
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[v] = (2 * v + (hi ? 1 : 0) == c) ? 1.f : 0.f;
            lds_cf4_p lcb = (lds_cf4_p)(Lc + (hi ? 16 : 0));
            asm volatile("" : "+v"(lcb));
            // column q's reads are issued two steps ahead, in slot q % 3
            f4 lq[3][4];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                lq[0][g] = lcb[g];
                lq[1][g] = lcb[8 + g];
            }
            float xprev = 0.f;
#pragma unroll
            for (int q = 0; q < 32; ++q) {
                const int hq = q & 1, vq = q >> 1;
                // the reads of column q+2 wait for x_{q-1}: without this
                // order the scheduler hoists the reads of all 32 columns to
                // the top, which exhausts the registers
                if (q + 2 < 32) {
                    asm volatile("" : "+v"(lcb) : "v"(xprev));
#pragma unroll
                    for (int g = ((q + 2) >> 1) >> 2; g < 4; ++g)
                        lq[(q + 2) % 3][g] = lcb[(q + 2) * 8 + g];
                }
                const float rq = rdlane(rv, q);
                const float xq = acc[vq] * rq;                 // valid in half hq
                acc[vq] = (hi == (hq == 1)) ? xq : acc[vq];
                const float xb = bcast_half(xq, hq);
                // whole groups live until here: a dead component's register
                // reused early would wait for the read in flight into it
#pragma unroll
                for (int g = (q >> 1) >> 2; g < 4; ++g) asm volatile("" : "+v"(lq[q % 3][g]));
                float Lq[16];
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    Lq[4 * g] = lq[q % 3][g].x; Lq[4 * g + 1] = lq[q % 3][g].y;
                    Lq[4 * g + 2] = lq[q % 3][g].z; Lq[4 * g + 3] = lq[q % 3][g].w;
                }
#pragma unroll
                for (int v = vq + 1; v < 16; ++v) acc[v] = fmaf(-Lq[v], xb, acc[v]);
                if (hq == 0) {                                 // row q + 1 = 2 vq + 1: half 1
                    const float t = fmaf(-Lq[vq], xb, acc[vq]);
                    acc[vq] = hi ? t : acc[vq];
                }
                xprev = xb;
            }
            {
                float* gd = dinv + (int64_t)((col0 >> 5) + kb) * 1024;
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int p = 2 * v + (hi ? 1 : 0);
                    Dinv[p + c * SUBLD] = acc[v];
                    if (kb * 32 < bs) gd[p + c * 32] = acc[v];
                }
            }
        }
        mpc_lds_barrier();          // Dinv(kb) and D(kb); the trailing update of kb-1 done

        // ---- 2. sub-tiles below: X = S Dinv^T.  Worked in the transposed
        // domain: lane (n, h) register i holds X(n, sig(i, h)), the k index
        // of MFMA step t is sig(t, h) on both operands, so no LDS round trip
        // is needed.  Wave 0 takes the sub-tile that the next diagonal block
        // depends on (ib = kb+1) and applies it to that block at once
        // (S(kb+1, kb+1) -= X X^T); the other waves take ib >= kb+2.
        auto solve_below = [&](int ib) {
            float* S = T + sidx(ib, kb) * SUBSZ;
            const int n = lane & 31, h = lane >> 5;
            v16f s0, x = {};
            float dv[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                s0[i] = S[n + sig(i, h) * SUBLD];
                dv[i] = Dinv[n + sig(i, h) * SUBLD];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = 0; t < 16; ++t)
                x = __builtin_amdgcn_mfma_f32_32x32x2f32(dv[t], s0[t], x, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 16; ++i) S[n + sig(i, h) * SUBLD] = x[i];
        };
        if (wave == 0) {
            if (kb + 1 < nsb) {
                solve_below(kb + 1);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
                update_subtile(T, kb + 1, kb + 1, kb);
            }
        } else {
            for (int ib = kb + 1 + wave; ib < nsb; ib += NW - 1) solve_below(ib);
        }
        mpc_lds_barrier();          // X(ib, kb) for every ib > kb

        // ---- 3. trailing update by sub-column kb on sub-columns kb+1 ..
        // (waves 1..3, during the factor of sub-block kb+1, which wave 0
        // already updated), then sub-column kb, now final, back to A
        if (wave != 0) {
            const int m = nsb - 1 - kb;                  // sub-columns kb+1 .. nsb-1
            const int npair = m * (m + 1) / 2;
            for (int pi = wave; pi < npair; pi += NW - 1) {   // pair 0 = (kb+1, kb+1): wave 0
                int ii = 0;
                while ((ii + 1) * (ii + 2) / 2 <= pi) ++ii;
                const int jj = pi - ii * (ii + 1) / 2;
                update_subtile(T, kb + 1 + ii, kb + 1 + jj, kb);
            }
            store_subcol(T, A, lda, bs, kb, 64, NT - 64);
        }
    }

    // ---- min pivot^2 and failure (every sub-column went back to A in the loop)
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
            if (fail_nan) atomicOr(&sm->nan_flag, 1);
        } else {
            const uint32_t mine = __float_as_uint(red[0]);
            const uint32_t before = atomicMin(&sm->mu_bits, mine);
            sm->mu_step[mu_slot] = before < mine ? before : mine;
        }
        sm->schur_max[mu_slot] = 0u;     // the step's TRSM folds its ratios in
    }
    if (want_f2) {
        // ||L_jj||_F^2: thread t sums the lower-triangle entries (r, c),
        // r >= c, of sub-tiles (ib, jb), jb <= ib, at element index
        // t, t + NT, ... of each sub-tile in order; then a fixed-order tree
        mpc_lds_barrier();
        float f2 = 0.f;
        for (int ib = 0; ib < nsb; ++ib)
            for (int jb = 0; jb <= ib; ++jb)
                for (int e = tid; e < 1024; e += NT) {
                    const int r = ib * 32 + (e & 31), c = jb * 32 + (e >> 5);
                    const float v = T[sidx(ib, jb) * SUBSZ + (e & 31) + (e >> 5) * SUBLD];
                    f2 = (r < bs && c <= r) ? fmaf(v, v, f2) : f2;
                }
        if (tid < 256) red[tid] = f2;
        mpc_lds_barrier();
        for (int w = 128; w > 0; w >>= 1) {
            if (tid < w) red[tid] += red[tid + w];
            mpc_lds_barrier();
        }
        if (tid == 0) sm->diag_f2[mu_slot] = red[0];
    }
}

}  // namespace

int mpc_launch_potrf(hipStream_t st, float* A, int64_t lda, int bs, int col0, float* dinv,
                     mpc_small* sm, int ntiles_next, uint32_t* stats_amax, uint32_t* stats_maxrs,
                     int mu_slot, int want_f2) {
    if (bs <= 0 || bs > 256) {
        mpc_set_error("potrf: block size %d outside 1..256", bs);
        return MPC_ERR_ARG;
    }
    if (lda > (1 << 21)) {
        mpc_set_error("potrf: leading dimension %lld above 2^21", (long long)lda);
        return MPC_ERR_ARG;
    }
    if (ntiles_next > 1024) {
        mpc_set_error("potrf: %d panel tiles exceed 1024", ntiles_next);
        return MPC_ERR_ARG;
    }
    if (mu_slot != 0 && mu_slot != 1) {
        mpc_set_error("potrf: mu slot %d", mu_slot);
        return MPC_ERR_ARG;
    }
    hipLaunchKernelGGL(potrf_kernel, dim3(1), dim3(NT), 0, st, A, lda, bs, col0, dinv, sm,
                       ntiles_next, stats_amax, stats_maxrs, mu_slot, want_f2 ? 1 : 0);
    MPC_LAUNCH_CHECK();
    return 0;
}
