// k_tri.hip -- triangular solves of the PCG preconditioner, z = L^-T L^-1 r.
//
// Replaces the two FP64 cusolverDnDpotrs sweeps per PCG iteration
// (mixed_cholesky.cu:147,166,204; potrs on an FP64 copy of L made by
// convertFloattoDouble, :141).  Here L stays FP32 and is read once per sweep
// for all right-hand sides; vectors are FP64.
//
// A row-by-row substitution over n = 16384 rows is a dependency chain; the
// sweep is instead blocked with big diagonal blocks of Bo rows (Bo = 1024 or
// 2048) whose inverses Winv_I = L_II^{-1} are formed once per factor:
//   forward   for I:   y_I = Winv_I w_I ;  w_{>I} -= L_{>I,I} y_I
//   backward  for I:   x_I = Winv_I^T w_I; w_{<I} -= L_{I,<I}^T x_I
// so a sweep is 2 * n / Bo bandwidth-bound kernels with no serial work
// inside.  Winv_I is built by recursive doubling from the 32x32 inverses the
// POTRF kernel leaves behind:
//   [A 0; B C]^{-1} = [A^-1 0; -C^-1 B A^-1  C^-1]
// with a batched FP32 MFMA GEMM (v_mfma_f32_32x32x2f32) that skips the zero
// triangles of A^-1 and C^-1.  The temporary B A^-1 lives in the unused upper
// triangle of the Winv block.
#include "mpc_internal.hpp"
#include <algorithm>

namespace {

typedef float v16f __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------- batched GEMM
// For batch (q, I) and an output tile (T x T at m0, n0):
//   C[m][n] = alpha * sum_k A[m][k] B[k][n]    (column-major, ld per operand)
// a_lower: A[m][k] = 0 for k > m; b_lower: B[k][n] = 0 for k < n.  Both are
// applied per element while staging (the strict upper parts of inverse
// blocks hold temporaries of the previous doubling level) and per tile when
// choosing the K range.  Batch b = blockIdx.z = I * nq + q.
struct gemm_batch {
    const float* A; const float* B; float* C;
    int64_t lda, ldb, ldc;
    int64_t sA, sB, sC;          // stride between pairs q (elements)
    int64_t tA, tB, tC;          // stride between big blocks I
    int nq;
    int m, n, k;
    float alpha;
    int a_lower, b_lower;
};

template <int T>
__global__ __launch_bounds__(256) void gemm_f32_kernel(gemm_batch g) {
    constexpr int WT = T / 2;          // per-wave tile (2 x 2 waves)
    constexpr int NB = WT / 32;        // 32x32 MFMA blocks per wave side
    __shared__ float As[32][T + 4];    // As[k][m]
    __shared__ float Bs[32][T + 4];    // Bs[k][n]
    const int64_t q = blockIdx.z % g.nq, I = blockIdx.z / g.nq;
    const float* A = g.A + q * g.sA + I * g.tA;
    const float* B = g.B + q * g.sB + I * g.tB;
    float* C = g.C + q * g.sC + I * g.tC;
    const int m0 = blockIdx.x * T, n0 = blockIdx.y * T;
    if (m0 >= g.m || n0 >= g.n) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * WT, wn = (wave & 1) * WT;
    int klo = 0, khi = g.k;
    if (g.a_lower && m0 + T < khi) khi = m0 + T;
    if (g.b_lower) klo = n0;
    klo &= ~31;
    v16f acc[NB][NB];
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = v16f{};
    // Staging: every load of a K step issued together from clamped
    // (in-bounds) addresses and masked by a select when stored to LDS
    // (conditional loads would serialise); the next step's loads are issued
    // before this step's products, so their latency overlaps the MFMA work.
    constexpr int NL = T * 32 / 256;
    float xa[NL], xb[NL];
    auto gload = [&](int k0) {
#pragma unroll
        for (int u = 0; u < NL; ++u) {
            const int idx = tid + 256 * u;
            const int gm = m0 + idx % T, gk = k0 + idx / T;
            xa[u] = A[(gm < g.m ? gm : g.m - 1) + (int64_t)(gk < g.k ? gk : g.k - 1) * g.lda];
            const int gk2 = k0 + (idx & 31), gn = n0 + (idx >> 5);
            xb[u] = B[(gk2 < g.k ? gk2 : g.k - 1) + (int64_t)(gn < g.n ? gn : g.n - 1) * g.ldb];
        }
    };
    if (klo < khi) gload(klo);
    for (int k0 = klo; k0 < khi; k0 += 32) {
#pragma unroll
        for (int u = 0; u < NL; ++u) {
            const int idx = tid + 256 * u;
            const int mm = idx % T, kk = idx / T, gm = m0 + mm, gk = k0 + kk;
            const bool aok = gm < g.m && gk < g.k && (!g.a_lower || gk <= gm);
            As[kk][mm] = aok ? xa[u] : 0.f;
            const int k2 = idx & 31, nn = idx >> 5, gk2 = k0 + k2, gn = n0 + nn;
            const bool bok = gk2 < g.k && gn < g.n && (!g.b_lower || gk2 >= gn);
            Bs[k2][nn] = bok ? xb[u] : 0.f;
        }
        __syncthreads();
        if (k0 + 32 < khi) gload(k0 + 32);
        const int r = lane & 31, h = lane >> 5;
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            float av[NB], bv[NB];
#pragma unroll
            for (int a = 0; a < NB; ++a) av[a] = As[2 * s + h][wm + 32 * a + r];
#pragma unroll
            for (int b = 0; b < NB; ++b) bv[b] = Bs[2 * s + h][wn + 32 * b + r];
#pragma unroll
            for (int a = 0; a < NB; ++a)
#pragma unroll
                for (int b = 0; b < NB; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }
    const int n = lane & 31, h = lane >> 5;
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                int mm = m0 + wm + 32 * a + 8 * (i >> 2) + 4 * h + (i & 3);
                int nn = n0 + wn + 32 * b + n;
                if (mm < g.m && nn < g.n) C[mm + (int64_t)nn * g.ldc] = g.alpha * acc[a][b][i];
            }
}

// copy the 32x32 inverses onto the diagonal of the Winv blocks
__global__ __launch_bounds__(256) void winv_init_kernel(const float* __restrict__ dinv, int64_t n,
                                                       int Bo, float* __restrict__ W, int64_t b0) {
    const int64_t b = b0 + blockIdx.x;
    const int64_t r0 = b * 32;
    const int64_t I = r0 / Bo;
    const int64_t p = r0 - I * Bo;
    float* Wi = W + I * (int64_t)Bo * Bo;
    for (int idx = threadIdx.x; idx < 1024; idx += 256) {
        int r = idx & 31, c = idx >> 5;
        float v = (r0 + r < n && r0 + c < n) ? dinv[b * 1024 + r + c * 32] : (r == c ? 1.f : 0.f);
        Wi[(p + r) + (p + c) * (int64_t)Bo] = v;
    }
}

// WT = Winv^T for each big block (lower part only is meaningful in Winv)
__global__ __launch_bounds__(256) void transpose_blocks_kernel(const float* __restrict__ W,
                                                              float* __restrict__ WT, int Bo,
                                                              int64_t I0) {
    __shared__ float t[32][33];
    const int64_t I = I0 + blockIdx.z;
    const float* Wi = W + I * (int64_t)Bo * Bo;
    float* Ti = WT + I * (int64_t)Bo * Bo;
    const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
    if (r0 < c0) return;                              // upper tiles of Winv are unused
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int j = ty; j < 32; j += 8) {
        float v = Wi[(r0 + tx) + (int64_t)(c0 + j) * Bo];
        if (r0 == c0 && r0 + tx < c0 + j) v = 0.f;     // strict upper part of a diagonal tile
        t[j][tx] = v;
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) Ti[(c0 + tx) + (int64_t)(r0 + j) * Bo] = t[tx][j];
}

// ---------------------------------------------------------------- TRSV kernels
// Both sweeps are left-looking: the update a big block receives from the
// blocks solved before it is one GEMV over all of them, split into many
// workgroups (partial sums per split, in a fixed order), and the partial sums
// are folded into the right-hand side inside the triangular GEMV that
// applies the block's inverse.  Per block: two or three launches, each wide
// enough to fill the chip, instead of a narrow per-block update of every
// later block.
constexpr int NR = 4;       // right-hand sides per pass
constexpr int TG = 64;      // tile edge of the triangular GEMV
// forward: columns per split (4 waves x FW_CPW).  At n = 16384 (Bo = 2048)
// 128 beat 64, 256 and 512: more workgroups with fewer loads in flight each.
constexpr int FW_CW = 128;
constexpr int FW_CPW = FW_CW / 4;
constexpr int BW_CPW = 8;   // backward: columns per wave (32 per workgroup)
constexpr int BW_RW = 512;  // backward: rows per split (BW_RPL per lane)
constexpr int BW_RPL = BW_RW / 64;

// Forward, big block I = rows [p0, p0 + bI):
//   part[(s nrhs + j) bI + i] = sum_{c in split s} L[p0 + i][c] Y[c][j],  c < p0
// 64 rows per workgroup (lane = row: each column load is a 256-byte
// segment), split s = FW_CW columns, wave w its FW_CPW of them, all loads of
// a lane in flight before the first product; the split's Y values are staged
// in LDS and read back by all lanes at once.
__global__ __launch_bounds__(256) void fwd_gemv_partial_kernel(
    const float* __restrict__ L, int64_t ldl, int64_t p0, int bI, const double* __restrict__ Y,
    int64_t ldy, int nrhs, double* __restrict__ part) {
    __shared__ double red[4][NR][64];
    __shared__ double ys[NR][FW_CW];        // the split's Y values, read back broadcast
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i = blockIdx.x * 64 + lane;
    const int ic = i < bI ? i : bI - 1;
    const int sp = blockIdx.y;
    const int j0 = blockIdx.z * NR;
    const int nj = nrhs - j0 < NR ? nrhs - j0 : NR;
    const int64_t cs = (int64_t)sp * FW_CW;
    const float* Lr = L + (p0 + ic);
    float l[FW_CPW];
#pragma unroll
    for (int u = 0; u < FW_CPW; ++u) {
        const int64_t c = cs + wave * FW_CPW + u;
        const float x = Lr[(c < p0 ? c : p0 - 1) * ldl];     // clamped, then masked
        l[u] = c < p0 ? x : 0.f;
    }
#pragma unroll
    for (int q = 0; q < NR * FW_CW / 256; ++q) {
        const int idx = tid + 256 * q, j = idx / FW_CW, cc = idx % FW_CW;
        const int64_t c = cs + cc;
        const int jj = j < nj ? j : nj - 1;
        const double v = Y[(int64_t)(j0 + jj) * ldy + (c < p0 ? c : p0 - 1)];
        ys[j][cc] = (j < nj && c < p0) ? v : 0.0;
    }
    __syncthreads();
    double acc[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        double a0 = 0.0, a1 = 0.0;
        if (j < nj) {
#pragma unroll
            for (int u = 0; u < FW_CPW; u += 2) {
                a0 = fma((double)l[u], ys[j][wave * FW_CPW + u], a0);
                a1 = fma((double)l[u + 1], ys[j][wave * FW_CPW + u + 1], a1);
            }
        }
        acc[j] = a0 + a1;
    }
#pragma unroll
    for (int j = 0; j < NR; ++j) red[wave][j][lane] = acc[j];
    __syncthreads();
    if (wave == 0 && i < bI) {
#pragma unroll
        for (int j = 0; j < NR; ++j)
            if (j < nj)
                part[((int64_t)sp * nrhs + j0 + j) * bI + i] =
                    ((red[0][j][lane] + red[1][j][lane]) + red[2][j][lane]) + red[3][j][lane];
    }
}

// Backward, big block I = columns [p0, p0 + bI), rows below it [p1, n):
//   part[(s nrhs + j) bI + c] = sum_{r in split s} L[r][p0 + c] X[r][j]
// 32 columns per workgroup (8 per wave), split s = BW_RW rows; lane k of a
// wave holds rows lane + 64k, k < BW_RPL, of each of its columns (coalesced
// 256-byte loads, 8 BW_RPL in flight), then the 64 lane sums of a column are added through
// LDS: lane t sums 8 of column t/8's, three shuffle steps finish it.
__global__ __launch_bounds__(256) void bwd_gemv_partial_kernel(
    const float* __restrict__ L, int64_t ldl, int64_t p0, int bI, int64_t n,
    const double* __restrict__ X, int64_t ldx, int nrhs, double* __restrict__ part) {
    __shared__ double buf[4][BW_CPW][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t p1 = p0 + bI;
    const int sp = blockIdx.y;
    const int64_t rs = p1 + (int64_t)sp * BW_RW;
    const int c0 = blockIdx.x * (4 * BW_CPW) + wave * BW_CPW;
    const int j0 = blockIdx.z * NR;
    const int nj = nrhs - j0 < NR ? nrhs - j0 : NR;
    float l[BW_CPW][BW_RPL];
#pragma unroll
    for (int cc = 0; cc < BW_CPW; ++cc) {
        const int c = c0 + cc < bI ? c0 + cc : bI - 1;
        const float* col = L + (p0 + c) * ldl;
#pragma unroll
        for (int k = 0; k < BW_RPL; ++k) {
            const int64_t r = rs + lane + 64 * k;
            l[cc][k] = col[r < n ? r : n - 1];
        }
    }
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        if (j < nj) {
            const double* x = X + (int64_t)(j0 + j) * ldx;
            double xv[BW_RPL];
#pragma unroll
            for (int k = 0; k < BW_RPL; ++k) {
                const int64_t r = rs + lane + 64 * k;
                const double v = x[r < n ? r : n - 1];
                xv[k] = r < n ? v : 0.0;                     // rows beyond n weigh nothing
            }
#pragma unroll
            for (int cc = 0; cc < BW_CPW; ++cc) {
                double a = 0.0;
#pragma unroll
                for (int k = 0; k < BW_RPL; ++k) a = fma((double)l[cc][k], xv[k], a);
                buf[wave][cc][lane] = a;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int col = lane >> 3, q = lane & 7;
            double v = 0.0;
#pragma unroll
            for (int t = 0; t < 8; ++t) v += buf[wave][col][8 * q + t];
            v += __shfl_xor(v, 1);
            v += __shfl_xor(v, 2);
            v += __shfl_xor(v, 4);
            if (q == 0 && c0 + col < bI)
                part[((int64_t)sp * nrhs + j0 + j) * bI + c0 + col] = v;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();                 // buf is rewritten next
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
}

// Triangular GEMV on a diagonal big block, out[i] = sum_j M[i + j*ld] v[j]
// over j <= i (lower: M = Winv, forward) or j >= i (upper: M = Winv^T,
// backward), i, j < bI, where v[j] = V[j] - sum_{s < S} pin[(s nrhs + r) bI + j]
// (the off-diagonal GEMV's partial sums).  One workgroup of TT_NW waves per
// 64 x 64 tile (row block ib, column chunk jb) of the triangle; lane = row.
// Wave w adds the partial sums s = w, w + TT_NW, .. of the tile's 64 v
// entries (batches of 8 independent loads); the waves' sums are combined in
// LDS in wave order.  Wave w then multiplies its TT_CW columns of the tile
// (their loads in flight at once, v read back from LDS) and the waves'
// products are added in wave order.  Partial sums go to
// part[(jb * nrhs + j) * bI + i]; tri_gemv_reduce adds the chunks of a row
// in order (deterministic).
#ifndef MPC_TT_NW
#define MPC_TT_NW 4
#endif
constexpr int TT_NW = MPC_TT_NW;          // waves per tile
constexpr int TT_CW = TG / TT_NW;          // tile columns per wave
template <bool UPPER>
__global__ __launch_bounds__(64 * TT_NW) void tri_gemv_tile_kernel(const float* __restrict__ M,
                                                                  int ld, int bI,
                                                                  const double* __restrict__ V,
                                                                  int64_t ldv, int nrhs,
                                                                  const double* __restrict__ pin,
                                                                  int S,
                                                                  double* __restrict__ part) {
    __shared__ double vsum[TT_NW][64];
    __shared__ double vs[64];
    __shared__ double acc_w[TT_NW][64];
    // tile t -> (ib, jb): lower ib >= jb, t = ib (ib + 1) / 2 + jb; upper mirrors it
    const int t = blockIdx.x;
    int a = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (a * (a + 1) / 2 > t) --a;
    while ((a + 1) * (a + 2) / 2 <= t) ++a;
    const int b = t - a * (a + 1) / 2;
    const int ib = UPPER ? b : a, jb = UPPER ? a : b;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = ib * TG + lane;
    const int j0 = jb * TG;
    const int c0 = j0 + TT_CW * wave;      // this wave's first column
    const int r0 = blockIdx.y * NR;
    const int nj = nrhs - r0 < NR ? nrhs - r0 : NR;
    // This wave's 16 columns of the tile, all loads issued before any use.  A
    // load guarded by a condition (or whose use is) would be sunk into a
    // branch with its own wait, serialising the loads; hence three shapes:
    float m[TT_CW];
    const bool full = j0 + TG <= bI && ib * TG + TG <= bI;
    if (full && ib != jb) {                  // interior: no mask
        const float* Mi = M + i + (int64_t)c0 * ld;
#pragma unroll
        for (int u = 0; u < TT_CW; ++u) m[u] = Mi[(int64_t)u * ld];
    } else if (full) {                       // diagonal tile: per-lane triangle select
        const float* Mi = M + i + (int64_t)c0 * ld;
#pragma unroll
        for (int u = 0; u < TT_CW; ++u) m[u] = Mi[(int64_t)u * ld];
#pragma unroll
        for (int u = 0; u < TT_CW; ++u) {
            const int cl = TT_CW * wave + u;              // column within the tile
            const bool keep = UPPER ? cl >= lane : cl <= lane;
            m[u] = keep ? m[u] : 0.f;
        }
    } else {                                 // ragged last block: clamped addresses
        const int ic = i < bI ? i : bI - 1;
#pragma unroll
        for (int u = 0; u < TT_CW; ++u) {
            const int j = c0 + u;
            m[u] = M[ic + (int64_t)(j < bI ? j : bI - 1) * ld];
        }
#pragma unroll
        for (int u = 0; u < TT_CW; ++u) {
            const int j = c0 + u;
            const bool keep = j < bI && (UPPER ? j >= i : j <= i);
            m[u] = keep ? m[u] : 0.f;
        }
    }
    // lane u of every wave builds a part of v[j0 + u]
    const int jl = j0 + lane < bI ? j0 + lane : bI - 1;
    const int64_t sstride = (int64_t)nrhs * bI;
    for (int r = 0; r < nj; ++r) {
        const double* pr = pin + (int64_t)(r0 + r) * bI + jl;
        double sum = 0.0;
        int sp = wave;
        for (; sp + 7 * TT_NW < S; sp += 8 * TT_NW) {      // batches of 8 independent loads
            double pv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) pv[q] = pr[(int64_t)(sp + q * TT_NW) * sstride];
#pragma unroll
            for (int q = 0; q < 8; ++q) sum += pv[q];
        }
        for (; sp < S; sp += TT_NW) sum += pr[(int64_t)sp * sstride];
        vsum[wave][lane] = sum;
        __syncthreads();
        if (wave == 0) {
            double tot = vsum[0][lane];
#pragma unroll
            for (int w = 1; w < TT_NW; ++w) tot += vsum[w][lane];
            vs[lane] = V[(int64_t)(r0 + r) * ldv + jl] - tot;
        }
        __syncthreads();
        double a0 = 0.0, a1 = 0.0;
#pragma unroll
        for (int u = 0; u < TT_CW; u += 2) {
            a0 = fma((double)m[u], vs[TT_CW * wave + u], a0);
            a1 = fma((double)m[u + 1], vs[TT_CW * wave + u + 1], a1);
        }
        acc_w[wave][lane] = a0 + a1;
        __syncthreads();
        if (wave == 0 && i < bI) {
            double tot = acc_w[0][lane];
#pragma unroll
            for (int w = 1; w < TT_NW; ++w) tot += acc_w[w][lane];
            part[((int64_t)jb * nrhs + r0 + r) * bI + i] = tot;
        }
        __syncthreads();                    // vsum, vs and acc_w are reused
    }
}

// 64 rows per workgroup (lane = row, one TG chunk row).  Wave w sums the
// chunks jlo + w, jlo + w + 4, ... in batches of 8 independent loads
// (clamped chunk index, zero weight beyond jhi): one batch when the block
// has at most 32 chunks.  The four wave sums are then added in order.
__global__ __launch_bounds__(256) void tri_gemv_reduce_kernel(const double* __restrict__ part,
                                                             int bI, int nrhs, int upper,
                                                             double* __restrict__ out,
                                                             int64_t ldo) {
    __shared__ double red[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i0 = blockIdx.x * 64 + lane;
    const int i = i0 < bI ? i0 : bI - 1;
    const int r = blockIdx.y;
    const int nb = (bI + TG - 1) / TG, ib = i / TG;
    const int jlo = upper ? ib : 0, jhi = upper ? nb - 1 : ib;
    double s = 0.0;
    for (int j0 = jlo + w; j0 <= jhi; j0 += 32) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int jb = j0 + 4 * u <= jhi ? j0 + 4 * u : jhi;
            v[u] = part[((int64_t)jb * nrhs + r) * bI + i];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) s += j0 + 4 * u <= jhi ? v[u] : 0.0;
    }
    red[w][lane] = s;
    __syncthreads();
    if (w == 0 && i0 < bI)
        out[i0 + (int64_t)r * ldo] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// 128x128 tiles when they give at least two workgroups per CU of a 256-CU
// device, else 64x64 tiles (the deep levels of the doubling are a few large
// products: four times the workgroups for the same work)
int launch_gemm(hipStream_t st, const gemm_batch& g, int nbatch) {
    const int64_t t128 = (int64_t)cdiv(g.m, 128) * cdiv(g.n, 128) * nbatch;
    if (g.m >= 128 && g.n >= 128 && t128 >= 512) {
        dim3 grid(cdiv(g.m, 128), cdiv(g.n, 128), (unsigned)nbatch);
        hipLaunchKernelGGL(gemm_f32_kernel<128>, grid, dim3(256), 0, st, g);
    } else {
        dim3 grid(cdiv(g.m, 64), cdiv(g.n, 64), (unsigned)nbatch);
        hipLaunchKernelGGL(gemm_f32_kernel<64>, grid, dim3(256), 0, st, g);
    }
    MPC_LAUNCH_CHECK();
    return 0;
}

}  // namespace

// Build Winv_I = L_II^{-1} for every big block (Bo x Bo, column-major, ld Bo,
// block I at W + I*Bo*Bo) from L and its 32x32 diagonal inverses, then
// WT_I = Winv_I^T.  Full big blocks are batched per level; a ragged last big
// block is handled by its own launches with clipped sizes.
static int double_level(hipStream_t st, const float* L, int64_t ldl, int64_t n, int Bo, float* W,
                        int h, int64_t I0, int nblk, int64_t rows, int q0, int nq, int hc) {
    const int64_t base = I0 * Bo;
    const int64_t p = 2 * (int64_t)h * q0;
    float* Wi = W + I0 * (int64_t)Bo * Bo;
    const int64_t bs = (int64_t)Bo * Bo;                 // stride between big blocks in W
    gemm_batch g1;                                       // T = B * Ainv -> upper part (p, p+h)
    g1.A = L + (base + p + h) + (base + p) * ldl;  g1.lda = ldl;
    g1.B = Wi + p + p * (int64_t)Bo;              g1.ldb = Bo;
    g1.C = Wi + p + (p + h) * (int64_t)Bo;        g1.ldc = Bo;
    g1.sA = 2 * (int64_t)h * (1 + ldl);
    g1.sB = g1.sC = 2 * (int64_t)h * (1 + Bo);
    g1.tA = (int64_t)Bo * (1 + ldl);
    g1.tB = g1.tC = bs;
    g1.nq = nq;
    g1.m = hc; g1.n = h; g1.k = h; g1.alpha = 1.f; g1.a_lower = 0; g1.b_lower = 1;
    int rc = launch_gemm(st, g1, nq * nblk);
    if (rc) return rc;
    gemm_batch g2;                                       // X21 = -Cinv * T -> (p+h, p)
    g2.A = Wi + (p + h) + (p + h) * (int64_t)Bo;  g2.lda = Bo;
    g2.B = Wi + p + (p + h) * (int64_t)Bo;        g2.ldb = Bo;
    g2.C = Wi + (p + h) + p * (int64_t)Bo;        g2.ldc = Bo;
    g2.sA = g2.sB = g2.sC = 2 * (int64_t)h * (1 + Bo);
    g2.tA = g2.tB = g2.tC = bs;
    g2.nq = nq;
    g2.m = hc; g2.n = h; g2.k = hc; g2.alpha = -1.f; g2.a_lower = 1; g2.b_lower = 0;
    (void)rows;
    return launch_gemm(st, g2, nq * nblk);
}

// Big blocks [I0, I1) only (each depends on L_II and the 32x32 inverses of
// its diagonal alone, so blocks can be built as soon as they are factored).
int mpc_build_block_inverses_range(hipStream_t st, const float* L, int64_t ldl, int64_t n, int Bo,
                                   const float* dinv, float* W, float* WT, int64_t I0,
                                   int64_t I1) {
    if (Bo % 64 || Bo < 64) {
        mpc_set_error("block inverse: Bo must be a multiple of 64");
        return MPC_ERR_ARG;
    }
    const int64_t nI = (n + Bo - 1) / Bo;
    const int64_t nfull = n / Bo;                 // big blocks with Bo rows
    if (I1 > nI) I1 = nI;
    if (I0 >= I1) return 0;
    const int64_t b0 = I0 * (Bo / 32), b1 = std::min<int64_t>((n + 31) / 32, I1 * (Bo / 32));
    hipLaunchKernelGGL(winv_init_kernel, dim3((unsigned)(b1 - b0)), dim3(256), 0, st, dinv, n, Bo, W,
                       b0);
    MPC_LAUNCH_CHECK();
    const int64_t f1 = std::min(I1, nfull);      // full blocks in the range: [I0, f1)
    int rc;
    for (int h = 32; h < Bo; h *= 2) {
        const int pairs = Bo / (2 * h);
        if (f1 > I0 &&
            (rc = double_level(st, L, ldl, n, Bo, W, h, I0, (int)(f1 - I0), Bo, 0, pairs, h)))
            return rc;
        if (nI > nfull && nfull >= I0 && nfull < I1) {   // ragged last big block
            const int64_t I = nfull, rows = n - I * Bo;
            int npair = 0;
            for (int q = 0; q < pairs; ++q)
                if (2 * h * q + h < rows) npair = q + 1;
            if (npair == 0) continue;
            const int full = (2 * h * (npair - 1) + 2 * h <= rows) ? npair : npair - 1;
            if (full > 0 && (rc = double_level(st, L, ldl, n, Bo, W, h, I, 1, rows, 0, full, h)))
                return rc;
            if (npair > full) {
                const int hc = (int)(rows - 2 * (int64_t)h * full - h);
                if ((rc = double_level(st, L, ldl, n, Bo, W, h, I, 1, rows, full, 1, hc))) return rc;
            }
        }
    }
    dim3 tg((unsigned)(Bo / 32), (unsigned)(Bo / 32), (unsigned)(I1 - I0));
    hipLaunchKernelGGL(transpose_blocks_kernel, tg, dim3(256), 0, st, W, WT, Bo, I0);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_build_block_inverses(hipStream_t st, const float* L, int64_t ldl, int64_t n, int Bo,
                             const float* dinv, float* W, float* WT) {
    return mpc_build_block_inverses_range(st, L, ldl, n, Bo, dinv, W, WT, 0, (n + Bo - 1) / Bo);
}

static int tri_gemv(hipStream_t st, const float* M, int Bo, int bI, int upper, const double* V,
                    int64_t ldv, int nrhs, const double* pin, int S, double* part, double* out,
                    int64_t ldo) {
    const int nb = (bI + TG - 1) / TG;
    const dim3 grid((unsigned)(nb * (nb + 1) / 2), cdiv(nrhs, NR));
    if (upper)
        hipLaunchKernelGGL(tri_gemv_tile_kernel<true>, grid, dim3(64 * TT_NW), 0, st, M, Bo, bI, V,
                           ldv, nrhs, pin, S, part);
    else
        hipLaunchKernelGGL(tri_gemv_tile_kernel<false>, grid, dim3(64 * TT_NW), 0, st, M, Bo, bI, V,
                           ldv, nrhs, pin, S, part);
    hipLaunchKernelGGL(tri_gemv_reduce_kernel, dim3(cdiv(bI, 64), (unsigned)nrhs), dim3(256), 0, st,
                       part, bI, nrhs, upper, out, ldo);
    MPC_LAUNCH_CHECK();
    return 0;
}

// splits of the off-diagonal GEMVs: at most n / FW_CW (forward, over p0 < n
// columns) and n / BW_RW (backward, over fewer than n rows)
static int64_t gemv_splits_max(int64_t n, int Bo) {
    (void)Bo;
    return std::max<int64_t>((n + FW_CW - 1) / FW_CW, (n + BW_RW - 1) / BW_RW);
}

int64_t mpc_block_prec_partial_doubles(int64_t n, int Bo, int nrhs) {
    return ((int64_t)(Bo + TG - 1) / TG + gemv_splits_max(n, Bo)) * Bo * nrhs;
}

// z = L^{-T} L^{-1} r for nrhs columns: R (in), Y (forward result), Z (out).
// late_ev (optional): the stream waits for it before the first use of big
// block I_late's inverse (which may still be in the making).
int mpc_apply_block_prec(hipStream_t st, const float* L, int64_t ldl, int64_t n, int Bo,
                         const float* Winv, const float* WinvT, const double* R, double* Y,
                         double* Z, int64_t ld, int nrhs, double* part, hipEvent_t late_ev,
                         int64_t I_late) {
    const int64_t nI = (n + Bo - 1) / Bo;
    const unsigned gz = cdiv(nrhs, NR);
    double* gpart = part + (int64_t)((Bo + TG - 1) / TG) * Bo * nrhs;   // GEMV partial sums
    int rc;
    for (int64_t I = 0; I < nI; ++I) {
        const int64_t p0 = I * Bo;
        const int bI = (int)(n - p0 < Bo ? n - p0 : Bo);
        int S = 0;
        if (p0 > 0) {
            S = (int)cdiv(p0, FW_CW);
            hipLaunchKernelGGL(fwd_gemv_partial_kernel, dim3(cdiv(bI, 64), (unsigned)S, gz),
                               dim3(256), 0, st, L, ldl, p0, bI, Y, ld, nrhs, gpart);
        }
        const float* Wi = Winv + I * (int64_t)Bo * Bo;
        if (late_ev && I == I_late && hipStreamWaitEvent(st, late_ev, 0) != hipSuccess) {
            mpc_set_error("apply_block_prec: stream wait failed");
            return MPC_ERR_HIP;
        }
        if ((rc = tri_gemv(st, Wi, Bo, bI, 0, R + p0, ld, nrhs, gpart, S, part, Y + p0, ld)))
            return rc;
    }
    for (int64_t I = nI - 1; I >= 0; --I) {
        const int64_t p0 = I * Bo;
        const int bI = (int)(n - p0 < Bo ? n - p0 : Bo);
        const int64_t p1 = p0 + bI;
        int S = 0;
        if (p1 < n) {
            S = (int)cdiv(n - p1, BW_RW);
            hipLaunchKernelGGL(bwd_gemv_partial_kernel, dim3(cdiv(bI, 4 * BW_CPW), (unsigned)S, gz),
                               dim3(256), 0, st, L, ldl, p0, bI, n, Z, ld, nrhs, gpart);
        }
        const float* WTi = WinvT + I * (int64_t)Bo * Bo;
        if ((rc = tri_gemv(st, WTi, Bo, bI, 1, Y + p0, ld, nrhs, gpart, S, part, Z + p0, ld)))
            return rc;
    }
    MPC_LAUNCH_CHECK();
    return 0;
}
