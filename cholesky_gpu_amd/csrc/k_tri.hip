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

namespace {

typedef float v16f __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------- batched GEMM
// For batch b and output tile (64 x 64 at m0, n0):
//   C[m][n] = alpha * sum_k A[m][k] B[k][n]    (column-major, ld per operand)
// over k in [klo, khi) where the triangular structure allows:
//   a_lower: A[m][k] = 0 for k > m  -> khi <= m0 + 64
//   b_lower: B[k][n] = 0 for k < n  -> klo >= n0
struct gemm_batch {
    const float* A; const float* B; float* C;
    int64_t lda, ldb, ldc;
    int64_t sA, sB, sC;          // batch strides (elements)
    int m, n, k;
    float alpha;
    int a_lower, b_lower;
};

__global__ __launch_bounds__(256) void gemm_f32_kernel(gemm_batch g) {
    __shared__ float As[32][64 + 4];   // As[k][m]
    __shared__ float Bs[32][64 + 4];   // Bs[k][n]
    const int64_t bidx = blockIdx.z;
    const float* A = g.A + bidx * g.sA;
    const float* B = g.B + bidx * g.sB;
    float* C = g.C + bidx * g.sC;
    const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
    if (m0 >= g.m || n0 >= g.n) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    int klo = 0, khi = g.k;
    if (g.a_lower && m0 + 64 < khi) khi = m0 + 64;
    if (g.b_lower) klo = n0;
    klo &= ~31;
    v16f acc = {};
    for (int k0 = klo; k0 < khi; k0 += 32) {
        // stage A (64 x 32) and B (32 x 64)
        for (int idx = tid; idx < 64 * 32; idx += 256) {
            int mm = idx & 63, kk = idx >> 6;
            int gm = m0 + mm, gk = k0 + kk;
            // exact triangles: the strict upper parts of inverse blocks hold
            // temporaries of the previous doubling level, so mask per element
            const bool aok = gm < g.m && gk < g.k && (!g.a_lower || gk <= gm);
            As[kk][mm] = aok ? A[gm + (int64_t)gk * g.lda] : 0.f;
            int k2 = idx & 31, nn = idx >> 5;
            int gk2 = k0 + k2, gn = n0 + nn;
            const bool bok = gk2 < g.k && gn < g.n && (!g.b_lower || gk2 >= gn);
            Bs[k2][nn] = bok ? B[gk2 + (int64_t)gn * g.ldb] : 0.f;
        }
        __syncthreads();
        const int r = lane & 31, h = lane >> 5;
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            float a = As[2 * s + h][wm + r];
            float b = Bs[2 * s + h][wn + r];
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int n = lane & 31, h = lane >> 5;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        int mm = m0 + wm + 8 * (i >> 2) + 4 * h + (i & 3);
        int nn = n0 + wn + n;
        if (mm < g.m && nn < g.n) C[mm + (int64_t)nn * g.ldc] = g.alpha * acc[i];
    }
}

// copy the 32x32 inverses onto the diagonal of the Winv blocks, zero the rest
// of each 32-column strip's lower part handled later by the doubling
__global__ __launch_bounds__(256) void winv_init_kernel(const float* __restrict__ dinv, int64_t n,
                                                       int Bo, float* __restrict__ W) {
    // one workgroup per 32-block b (global), writes the 32x32 diagonal tile
    const int64_t b = blockIdx.x;
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

// ---------------------------------------------------------------- TRSV kernels
constexpr int NR = 8;   // right-hand sides per pass

// W[r][j] -= sum_{c < cw} L[r][c0 + c] * Y[c0 + c][j]  for r in [r0, n)
// 64 rows per workgroup (lane = row), 4 waves split the columns.
__global__ __launch_bounds__(256) void fwd_update_kernel(const float* __restrict__ L, int64_t ldl,
                                                        int64_t n, int64_t r0, int64_t c0, int cw,
                                                        const double* __restrict__ Y, int64_t ldy,
                                                        double* __restrict__ W, int64_t ldw,
                                                        int nrhs) {
    __shared__ double red[4][NR][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t r = r0 + (int64_t)blockIdx.x * 64 + lane;
    const int j0 = blockIdx.y * NR;
    const int nj = nrhs - j0 < NR ? nrhs - j0 : NR;
    double acc[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) acc[j] = 0.0;
    if (r < n) {
        const float* Lr = L + r + c0 * ldl;
        const double* Yc = Y + c0 + (int64_t)j0 * ldy;
        int c = wave;
        for (; c + 12 < cw; c += 16) {
            float l0 = Lr[(int64_t)c * ldl], l1 = Lr[(int64_t)(c + 4) * ldl];
            float l2 = Lr[(int64_t)(c + 8) * ldl], l3 = Lr[(int64_t)(c + 12) * ldl];
#pragma unroll
            for (int j = 0; j < NR; ++j)
                if (j < nj) {
                    const double* y = Yc + (int64_t)j * ldy;
                    acc[j] = fma((double)l0, y[c], acc[j]);
                    acc[j] = fma((double)l1, y[c + 4], acc[j]);
                    acc[j] = fma((double)l2, y[c + 8], acc[j]);
                    acc[j] = fma((double)l3, y[c + 12], acc[j]);
                }
        }
        for (; c < cw; c += 4) {
            float l0 = Lr[(int64_t)c * ldl];
#pragma unroll
            for (int j = 0; j < NR; ++j)
                if (j < nj) acc[j] = fma((double)l0, Yc[(int64_t)j * ldy + c], acc[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < NR; ++j) red[wave][j][lane] = acc[j];
    __syncthreads();
    if (wave == 0 && r < n) {
#pragma unroll
        for (int j = 0; j < NR; ++j)
            if (j < nj) {
                double s = (red[0][j][lane] + red[1][j][lane]) + (red[2][j][lane] + red[3][j][lane]);
                W[r + (int64_t)(j0 + j) * ldw] -= s;
            }
    }
}

// Y[p0 + r][j] = sum_{c <= r} Winv[r][c] W[p0 + c][j]  (rows of one big block)
__global__ __launch_bounds__(256) void fwd_diag_kernel(const float* __restrict__ Wi, int Bo, int bI,
                                                      int64_t p0, const double* __restrict__ W,
                                                      int64_t ldw, double* __restrict__ Y,
                                                      int64_t ldy, int nrhs) {
    __shared__ double red[4][NR][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = blockIdx.x * 64 + lane;
    const int j0 = blockIdx.y * NR;
    const int nj = nrhs - j0 < NR ? nrhs - j0 : NR;
    double acc[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) acc[j] = 0.0;
    if (r < bI) {
        const double* Wc = W + p0 + (int64_t)j0 * ldw;
        for (int c = wave; c <= r; c += 4) {
            float w = Wi[r + (int64_t)c * Bo];
#pragma unroll
            for (int j = 0; j < NR; ++j)
                if (j < nj) acc[j] = fma((double)w, Wc[(int64_t)j * ldw + c], acc[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < NR; ++j) red[wave][j][lane] = acc[j];
    __syncthreads();
    if (wave == 0 && r < bI) {
#pragma unroll
        for (int j = 0; j < NR; ++j)
            if (j < nj)
                Y[p0 + r + (int64_t)(j0 + j) * ldy] =
                    (red[0][j][lane] + red[1][j][lane]) + (red[2][j][lane] + red[3][j][lane]);
    }
}

// W[c][j] -= sum_{r < bI} L[p0 + r][c] * X[p0 + r][j]   for c < p0
// one wave per column, 4 columns per wave, 16 per workgroup
__global__ __launch_bounds__(256) void bwd_update_kernel(const float* __restrict__ L, int64_t ldl,
                                                        int64_t p0, int bI,
                                                        const double* __restrict__ X, int64_t ldx,
                                                        double* __restrict__ W, int64_t ldw,
                                                        int nrhs) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j0 = blockIdx.y * NR;
    const int nj = nrhs - j0 < NR ? nrhs - j0 : NR;
    const int64_t cbase = (int64_t)blockIdx.x * 16 + wave * 4;
    for (int q = 0; q < 4; ++q) {
        const int64_t c = cbase + q;
        if (c >= p0) break;
        const float* col = L + p0 + c * ldl;
        double acc[NR];
#pragma unroll
        for (int j = 0; j < NR; ++j) acc[j] = 0.0;
        for (int r = lane; r < bI; r += 64) {
            float l = col[r];
#pragma unroll
            for (int j = 0; j < NR; ++j)
                if (j < nj) acc[j] = fma((double)l, X[p0 + r + (int64_t)(j0 + j) * ldx], acc[j]);
        }
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            double s = acc[j];
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
            acc[j] = s;
        }
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < NR; ++j)
                if (j < nj) W[c + (int64_t)(j0 + j) * ldw] -= acc[j];
        }
    }
}

// X[p0 + c][j] = sum_{r >= c} Winv[r][c] W[p0 + r][j]
__global__ __launch_bounds__(256) void bwd_diag_kernel(const float* __restrict__ Wi, int Bo, int bI,
                                                      int64_t p0, const double* __restrict__ W,
                                                      int64_t ldw, double* __restrict__ X,
                                                      int64_t ldx, int nrhs) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j0 = blockIdx.y * NR;
    const int nj = nrhs - j0 < NR ? nrhs - j0 : NR;
    const int cbase = blockIdx.x * 16 + wave * 4;
    for (int q = 0; q < 4; ++q) {
        const int c = cbase + q;
        if (c >= bI) break;
        const float* col = Wi + (int64_t)c * Bo;
        double acc[NR];
#pragma unroll
        for (int j = 0; j < NR; ++j) acc[j] = 0.0;
        for (int r = c + lane; r < bI; r += 64) {
            float w = col[r];
#pragma unroll
            for (int j = 0; j < NR; ++j)
                if (j < nj) acc[j] = fma((double)w, W[p0 + r + (int64_t)(j0 + j) * ldw], acc[j]);
        }
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            double s = acc[j];
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
            acc[j] = s;
        }
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < NR; ++j)
                if (j < nj) X[p0 + c + (int64_t)(j0 + j) * ldx] = acc[j];
        }
    }
}

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

int launch_gemm(hipStream_t st, const gemm_batch& g, int nbatch) {
    dim3 grid(cdiv(g.m, 64), cdiv(g.n, 64), (unsigned)nbatch);
    hipLaunchKernelGGL(gemm_f32_kernel, grid, dim3(256), 0, st, g);
    MPC_LAUNCH_CHECK();
    return 0;
}

}  // namespace

// Build Winv_I = L_II^{-1} for every big block (Bo x Bo, column-major, ld Bo,
// block I at W + I*Bo*Bo) from L and its 32x32 diagonal inverses.
int mpc_build_block_inverses(hipStream_t st, const float* L, int64_t ldl, int64_t n, int Bo,
                             const float* dinv, float* W) {
    if (Bo % 64 || Bo < 64) {
        mpc_set_error("block inverse: Bo must be a multiple of 64");
        return MPC_ERR_ARG;
    }
    const int64_t nI = (n + Bo - 1) / Bo;
    const int64_t nb32 = (n + 31) / 32;
    hipLaunchKernelGGL(winv_init_kernel, dim3((unsigned)nb32), dim3(256), 0, st, dinv, n, Bo, W);
    MPC_LAUNCH_CHECK();
    // doubling: h = 32, 64, ..., Bo/2.  Pairs of h-blocks p = 2h q inside each
    // big block.  Rows beyond n (ragged last big block) are treated as part of
    // an identity extension: their L entries are never read (m, n, k clipped).
    for (int h = 32; h < Bo; h *= 2) {
        const int pairs = Bo / (2 * h);
        for (int64_t I = 0; I < nI; ++I) {
            const int64_t base = I * Bo;
            const int64_t rows = n - base < Bo ? n - base : Bo;
            int npair = 0;
            for (int q = 0; q < pairs; ++q)
                if (2 * h * q + h < rows) npair = q + 1;
            if (npair == 0) continue;
            // all pairs of this big block in one batched launch; the last pair
            // may be ragged, so clip its sizes separately
            const int full = (2 * h * (npair - 1) + 2 * h <= rows) ? npair : npair - 1;
            float* Wi = W + I * (int64_t)Bo * Bo;
            for (int part = 0; part < 2; ++part) {
                const int q0 = part == 0 ? 0 : full;
                const int qn = part == 0 ? full : npair - full;
                if (qn <= 0) continue;
                const int64_t p = 2 * (int64_t)h * q0;
                const int hc = part == 0 ? h : (int)(rows - p - h);   // C-block size
                // T = B * Ainv  (T stored in the upper triangle: position (p, p + h))
                gemm_batch g1;
                g1.A = L + (base + p + h) + (base + p) * ldl;  g1.lda = ldl;
                g1.B = Wi + p + p * (int64_t)Bo;              g1.ldb = Bo;
                g1.C = Wi + p + (p + h) * (int64_t)Bo;        g1.ldc = Bo;
                g1.sA = 2 * (int64_t)h * (1 + ldl);
                g1.sB = 2 * (int64_t)h * (1 + Bo);
                g1.sC = 2 * (int64_t)h * (1 + Bo);
                g1.m = hc; g1.n = h; g1.k = h; g1.alpha = 1.f; g1.a_lower = 0; g1.b_lower = 1;
                int rc = launch_gemm(st, g1, qn);
                if (rc) return rc;
                // X21 = -Cinv * T  -> position (p + h, p)
                gemm_batch g2;
                g2.A = Wi + (p + h) + (p + h) * (int64_t)Bo;  g2.lda = Bo;
                g2.B = Wi + p + (p + h) * (int64_t)Bo;        g2.ldb = Bo;
                g2.C = Wi + (p + h) + p * (int64_t)Bo;        g2.ldc = Bo;
                g2.sA = g2.sB = g2.sC = 2 * (int64_t)h * (1 + Bo);
                g2.m = hc; g2.n = h; g2.k = hc; g2.alpha = -1.f; g2.a_lower = 1; g2.b_lower = 0;
                rc = launch_gemm(st, g2, qn);
                if (rc) return rc;
            }
        }
    }
    return 0;
}

// z = L^{-T} L^{-1} w for nrhs columns.  W (in/out, destroyed), Y (work), Z (out)
int mpc_apply_block_prec(hipStream_t st, const float* L, int64_t ldl, int64_t n, int Bo,
                         const float* Winv, double* W, double* Y, double* Z, int64_t ld, int nrhs) {
    const int64_t nI = (n + Bo - 1) / Bo;
    const unsigned gz = cdiv(nrhs, NR);
    for (int64_t I = 0; I < nI; ++I) {
        const int64_t p0 = I * Bo;
        const int bI = (int)(n - p0 < Bo ? n - p0 : Bo);
        const float* Wi = Winv + I * (int64_t)Bo * Bo;
        hipLaunchKernelGGL(fwd_diag_kernel, dim3(cdiv(bI, 64), gz), dim3(256), 0, st, Wi, Bo, bI,
                           p0, W, ld, Y, ld, nrhs);
        const int64_t r0 = p0 + bI;
        if (r0 < n)
            hipLaunchKernelGGL(fwd_update_kernel, dim3(cdiv(n - r0, 64), gz), dim3(256), 0, st, L,
                               ldl, n, r0, p0, bI, Y, ld, W, ld, nrhs);
    }
    // backward: Y is the right-hand side now (destroyed), Z the solution
    for (int64_t I = nI - 1; I >= 0; --I) {
        const int64_t p0 = I * Bo;
        const int bI = (int)(n - p0 < Bo ? n - p0 : Bo);
        const float* Wi = Winv + I * (int64_t)Bo * Bo;
        hipLaunchKernelGGL(bwd_diag_kernel, dim3(cdiv(bI, 16), gz), dim3(256), 0, st, Wi, Bo, bI,
                           p0, Y, ld, Z, ld, nrhs);
        if (p0 > 0)
            hipLaunchKernelGGL(bwd_update_kernel, dim3(cdiv(p0, 16), gz), dim3(256), 0, st, L, ldl,
                               p0, bI, Z, ld, Y, ld, nrhs);
    }
    MPC_LAUNCH_CHECK();
    return 0;
}
