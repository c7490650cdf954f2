// k_pcg.hip -- bandwidth-bound kernels of the PCG loop.
//
// Replaces, from precond_CG (mixed_cholesky.cu:145-217): cublasDgemv (A p),
// cusolverDnDpotrs (z = L^-T L^-1 r, on an FP64 copy of L made by
// convertFloattoDouble), and the cublasDdot/Daxpy/Dscal/Dcopy calls that
// return scalars to the host.  Here A and L stay FP32 in HBM and are read once
// per product; vectors, accumulation, inner products and the alpha/beta
// scalars are FP64 and never leave the device.  Every reduction has a fixed
// order (two-stage, partials then an ordered sum), so results are bitwise
// reproducible run to run.
#include "mpc_internal.hpp"

namespace {

constexpr int RB = 256;          // rows per workgroup
constexpr int NRC = 8;           // right-hand sides per pass

// ---------------------------------------------------------------- GEMV
// part[s][j][r] = sum_{c in split s} A[r,c] X[c,j]
__global__ __launch_bounds__(RB) void gemv_partial_kernel(const float* __restrict__ A, int64_t lda,
                                                         int64_t n, int nrhs,
                                                         const double* __restrict__ X, int64_t ldx,
                                                         double* __restrict__ part, int64_t cols_per) {
    const int64_t r = (int64_t)blockIdx.x * RB + threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.y * cols_per;
    int64_t c1 = c0 + cols_per;
    if (c1 > n) c1 = n;
    const int j0 = blockIdx.z * NRC;
    const int nj = nrhs - j0 < NRC ? nrhs - j0 : NRC;
    double acc[NRC];
#pragma unroll
    for (int j = 0; j < NRC; ++j) acc[j] = 0.0;
    if (r < n) {
        const float* a = A + r;
        int64_t c = c0;
        for (; c + 4 <= c1; c += 4) {
            float a0 = a[c * lda], a1 = a[(c + 1) * lda], a2 = a[(c + 2) * lda], a3 = a[(c + 3) * lda];
#pragma unroll
            for (int j = 0; j < NRC; ++j) {
                if (j < nj) {
                    const double* x = X + (int64_t)(j0 + j) * ldx + c;
                    acc[j] = fma((double)a0, x[0], acc[j]);
                    acc[j] = fma((double)a1, x[1], acc[j]);
                    acc[j] = fma((double)a2, x[2], acc[j]);
                    acc[j] = fma((double)a3, x[3], acc[j]);
                }
            }
        }
        for (; c < c1; ++c) {
            float a0 = a[c * lda];
#pragma unroll
            for (int j = 0; j < NRC; ++j)
                if (j < nj) acc[j] = fma((double)a0, X[(int64_t)(j0 + j) * ldx + c], acc[j]);
        }
        const int64_t ns = gridDim.y;
#pragma unroll
        for (int j = 0; j < NRC; ++j)
            if (j < nj) part[((int64_t)(j0 + j) * ns + blockIdx.y) * n + r] = acc[j];
    }
}

__global__ __launch_bounds__(RB) void gemv_reduce_kernel(int64_t n, int nrhs, int nsplit,
                                                        const double* __restrict__ part,
                                                        double* __restrict__ Y, int64_t ldy) {
    const int64_t r = (int64_t)blockIdx.x * RB + threadIdx.x;
    const int j = blockIdx.y;
    if (r >= n) return;
    double s = 0.0;
    for (int q = 0; q < nsplit; ++q) s += part[((int64_t)j * nsplit + q) * n + r];
    Y[(int64_t)j * ldy + r] = s;
}

// ---------------------------------------------------------------- SYMV
// y = A x for symmetric A reading only the upper triangle (i <= j) of the
// column-major FP32 matrix: half the HBM traffic of a full GEMV.  One
// workgroup per 128 x 128 tile (I <= J), staged in LDS once; its row sums
// (A_IJ x_J) and column sums (A_IJ^T x_I) go to fixed partial slots
//   P[K][L][r]:  L >= K row-sum partial of tile (K, L); L < K column-sum
//               partial of tile (L, K),
// and symv_reduce sums P[K][0..nt) in order (deterministic).
constexpr int ST = 128;

__global__ __launch_bounds__(256) void symv_tile_kernel(const float* __restrict__ A, int64_t lda,
                                                       int64_t n, int nt, int nrhs,
                                                       const double* __restrict__ X, int64_t ldx,
                                                       double* __restrict__ P) {
    __shared__ float As[ST * (ST + 1)];
    __shared__ double xi[ST], xj[ST];
    // tile index -> (I, J), I <= J, row-major over the upper triangle
    const int64_t t = blockIdx.x;
    int64_t I = (int64_t)((2.0 * nt + 1.0 - sqrt((2.0 * nt + 1.0) * (2.0 * nt + 1.0) - 8.0 * t)) * 0.5);
    auto start = [&](int64_t i) { return i * nt - i * (i - 1) / 2; };
    while (I > 0 && start(I) > t) --I;
    while (start(I + 1) <= t) ++I;
    const int64_t J = I + (t - start(I));
    const int tid = threadIdx.x;
    const int64_t r0 = I * ST, c0 = J * ST;
    const bool diag = I == J;
    for (int idx = tid; idx < ST * ST; idx += 256) {
        int r = idx & (ST - 1), c = idx >> 7;
        int64_t gr = r0 + r, gc = c0 + c;
        float v = 0.f;
        if (gr < n && gc < n && (!diag || r <= c)) v = A[gr + gc * lda];
        As[c * (ST + 1) + r] = v;
    }
    for (int j = 0; j < nrhs; ++j) {
        if (tid < ST) {
            xi[tid] = r0 + tid < n ? X[j * ldx + r0 + tid] : 0.0;
            xj[tid] = c0 + tid < n ? X[j * ldx + c0 + tid] : 0.0;
        }
        __syncthreads();
        // threads 0..127: row sums of row tid; 128..255: column sums of column tid-128
        double s = 0.0;
        if (tid < ST) {
            const int r = tid;
            for (int c = 0; c < ST; ++c) s = fma((double)As[c * (ST + 1) + r], xj[c], s);
            if (diag) {   // mirrored strict lower part: sum_{c < r} A[c][r] x[c]
                for (int c = 0; c < r; ++c) s = fma((double)As[r * (ST + 1) + c], xi[c], s);
            }
            P[((I * nt + J) * nrhs + j) * ST + r] = s;
        } else if (!diag) {
            const int c = tid - ST;
            for (int r = 0; r < ST; ++r) s = fma((double)As[c * (ST + 1) + r], xi[r], s);
            P[((J * nt + I) * nrhs + j) * ST + c] = s;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void symv_reduce_kernel(int64_t n, int nt, int nrhs,
                                                         const double* __restrict__ P,
                                                         double* __restrict__ Y, int64_t ldy) {
    const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (row >= n) return;
    const int64_t K = row / ST, r = row % ST;
    double s = 0.0;
    for (int64_t L = 0; L < nt; ++L) s += P[((K * nt + L) * nrhs + j) * ST + r];
    Y[j * ldy + row] = s;
}

// ---------------------------------------------------------------- TRSV steps
// Solve the diagonal block in LDS: y = L_jj^{-1} w (forward) or
// L_jj^{-T} w (backward) via the 32x32 inverses, for one right-hand side.
// ys[bs] holds w on entry and the solution on exit.  All 256 threads.
__device__ void diag_solve(const float* __restrict__ Ljj, int64_t ldl, int bs,
                           const float* __restrict__ dinv_j, double* ys, double* tmp, bool trans) {
    const int t = threadIdx.x;
    const int nsb = (bs + 31) >> 5;
    const int lane = t & 31, grp = t >> 5;   // 8 groups of 32 threads
    for (int it = 0; it < nsb; ++it) {
        const int b = trans ? nsb - 1 - it : it;
        const int rb0 = 32 * b;
        const int rows = bs - rb0 < 32 ? bs - rb0 : 32;
        // tmp[lane] = w_b[lane] - sum over solved blocks (split across 8 groups)
        double part = 0.0;
        if (lane < rows) {
            if (!trans) {
                for (int q = grp; q < rb0; q += 8)
                    part += (double)Ljj[(rb0 + lane) + (int64_t)q * ldl] * ys[q];
            } else {
                for (int q = rb0 + 32 + grp; q < bs; q += 8)
                    part += (double)Ljj[q + (int64_t)(rb0 + lane) * ldl] * ys[q];
            }
        }
        tmp[32 + grp * 32 + lane] = part;
        __syncthreads();
        if (t < 32) {
            double s = t < rows ? ys[rb0 + t] : 0.0;
            for (int g = 0; g < 8; ++g) s -= tmp[32 + g * 32 + t];
            tmp[t] = t < rows ? s : 0.0;
        }
        __syncthreads();
        if (t < 32 && t < rows) {
            const float* D = dinv_j + b * 1024;     // column-major 32x32
            double s = 0.0;
            if (!trans) {
                for (int q = 0; q <= t; ++q) s += (double)D[t + q * 32] * tmp[q];
            } else {
                for (int q = t; q < rows; ++q) s += (double)D[q + t * 32] * tmp[q];
            }
            ys[rb0 + t] = s;
        }
        __syncthreads();
    }
}

// forward step j: y_j = L_jj^{-1} w_j -> Y;  w_i -= L_ij y_j for rows below.
// W rows of block j are only read (every workgroup re-solves them), so the
// solution goes to the separate Y.
__global__ __launch_bounds__(RB) void trsv_fwd_kernel(const float* __restrict__ L, int64_t ldl,
                                                     int64_t n, int nb, int64_t c0, int bs,
                                                     const float* __restrict__ dinv_j,
                                                     double* __restrict__ W, int64_t ldw, int nrhs,
                                                     double* __restrict__ Y, int64_t ldy) {
    __shared__ double ys[256];
    __shared__ double tmp[32 + 256];
    const int t = threadIdx.x;
    const int64_t e = c0 + bs;
    const float* Ljj = L + c0 + c0 * ldl;
    const int64_t r = e + (int64_t)blockIdx.x * RB + t;
    for (int j = 0; j < nrhs; ++j) {
        double* w = W + (int64_t)j * ldw;
        if (t < bs) ys[t] = w[c0 + t];
        __syncthreads();
        diag_solve(Ljj, ldl, bs, dinv_j, ys, tmp, false);
        if (blockIdx.x == 0 && t < bs) Y[(int64_t)j * ldy + c0 + t] = ys[t];
        if (r < n) {
            const float* Lr = L + r + c0 * ldl;
            double acc = 0.0;
            for (int c = 0; c < bs; ++c) acc = fma((double)Lr[(int64_t)c * ldl], ys[c], acc);
            w[r] -= acc;
        }
        __syncthreads();
    }
}

// backward step j: x_j = L_jj^{-T} w_j -> Xo;  w_c -= sum_{r in block j} L[r,c] x_r
// for c < c0 (rows of W that this step does not read)
__global__ __launch_bounds__(RB) void trsv_bwd_kernel(const float* __restrict__ L, int64_t ldl,
                                                     int64_t n, int nb, int64_t c0, int bs,
                                                     const float* __restrict__ dinv_j,
                                                     double* __restrict__ W, int64_t ldw, int nrhs,
                                                     int cols_per_block, double* __restrict__ Xo,
                                                     int64_t ldxo) {
    __shared__ double ys[256];
    __shared__ double tmp[32 + 256];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const float* Ljj = L + c0 + c0 * ldl;
    for (int j = 0; j < nrhs; ++j) {
        double* w = W + (int64_t)j * ldw;
        if (t < bs) ys[t] = w[c0 + t];
        __syncthreads();
        diag_solve(Ljj, ldl, bs, dinv_j, ys, tmp, true);
        if (blockIdx.x == 0 && t < bs) Xo[(int64_t)j * ldxo + c0 + t] = ys[t];
        // columns c in [blockIdx.x * cols_per_block, ...) below c0
        const int64_t cbeg = (int64_t)blockIdx.x * cols_per_block;
        for (int64_t cc = cbeg + wave; cc < cbeg + cols_per_block && cc < c0; cc += 4) {
            const float* col = L + c0 + cc * ldl;
            double s = 0.0;
            for (int q = lane; q < bs; q += 64) s = fma((double)col[q], ys[q], s);
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
            if (lane == 0) w[cc] -= s;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- dots and updates
__global__ __launch_bounds__(RB) void dot_partial_kernel(int64_t n, const double* __restrict__ a,
                                                        int64_t lda_, const double* __restrict__ b,
                                                        int64_t ldb, double* __restrict__ part) {
    __shared__ double red[RB];
    const int j = blockIdx.y;
    const int64_t r0 = (int64_t)blockIdx.x * RB * 8;
    double s = 0.0;
    for (int q = 0; q < 8; ++q) {
        int64_t r = r0 + q * RB + threadIdx.x;
        if (r < n) s = fma(a[j * lda_ + r], b[j * ldb + r], s);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = RB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)j * gridDim.x + blockIdx.x] = red[0];
}

__global__ __launch_bounds__(RB) void dot_final_kernel(int nparts, const double* __restrict__ part,
                                                      double* __restrict__ out) {
    __shared__ double red[RB];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int q = threadIdx.x; q < nparts; q += RB) s += part[(int64_t)j * nparts + q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = RB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = red[0];
}

__global__ __launch_bounds__(RB) void update_xr_kernel(int64_t n, double* __restrict__ X,
                                                      int64_t ldx, double* __restrict__ R,
                                                      int64_t ldr, const double* __restrict__ P,
                                                      int64_t ldp, const double* __restrict__ Q,
                                                      int64_t ldq, const double* __restrict__ rz,
                                                      const double* __restrict__ pq,
                                                      const int* __restrict__ active) {
    const int j = blockIdx.y;
    if (active && !active[j]) return;
    const int64_t r = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (r >= n) return;
    const double alpha = rz[j] / pq[j];
    X[j * ldx + r] = fma(alpha, P[j * ldp + r], X[j * ldx + r]);
    R[j * ldr + r] = fma(-alpha, Q[j * ldq + r], R[j * ldr + r]);
}

__global__ __launch_bounds__(RB) void update_p_kernel(int64_t n, double* __restrict__ P, int64_t ldp,
                                                     const double* __restrict__ Z, int64_t ldz,
                                                     const double* __restrict__ rz_new,
                                                     const double* __restrict__ rz_old,
                                                     const int* __restrict__ active) {
    const int j = blockIdx.y;
    if (active && !active[j]) return;
    const int64_t r = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (r >= n) return;
    const double beta = rz_new[j] / rz_old[j];
    P[j * ldp + r] = fma(beta, P[j * ldp + r], Z[j * ldz + r]);
}

__global__ __launch_bounds__(RB) void axpby_kernel(int64_t n, double a, const double* __restrict__ X,
                                                  int64_t ldx, double b, double* __restrict__ Y,
                                                  int64_t ldy) {
    const int j = blockIdx.y;
    const int64_t r = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (r >= n) return;
    Y[j * ldy + r] = (a == 0.0 && b == 0.0) ? 0.0 : a * X[j * ldx + r] + b * Y[j * ldy + r];
}

__global__ __launch_bounds__(RB) void copy2d_kernel(int64_t n, const double* __restrict__ X,
                                                   int64_t ldx, double* __restrict__ Y, int64_t ldy) {
    const int j = blockIdx.y;
    const int64_t r = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (r < n) Y[j * ldy + r] = X[j * ldx + r];
}

__global__ __launch_bounds__(RB) void copy_f32_kernel(int64_t rows, int64_t cols,
                                                     const float* __restrict__ A, int64_t lda,
                                                     float* __restrict__ B, int64_t ldb) {
    const int64_t r = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (r >= rows) return;
    for (int64_t c = blockIdx.y; c < cols; c += gridDim.y) B[r + c * ldb] = A[r + c * lda];
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + RB - 1) / RB); }

}  // namespace

int mpc_gemv_nsplit(int64_t n) {
    int64_t rb = (n + RB - 1) / RB;
    int64_t ns = (2048 + rb - 1) / rb;
    if (ns < 1) ns = 1;
    if (ns > 64) ns = 64;
    if (ns > n) ns = n;
    return (int)ns;
}

int mpc_launch_gemv(hipStream_t st, const float* A, int64_t lda, int64_t n, int nrhs,
                    const double* X, int64_t ldx, double* Y, int64_t ldy, double* part) {
    const int ns = mpc_gemv_nsplit(n);
    const int64_t cols_per = (n + ns - 1) / ns;
    dim3 g1(blocks_for(n), (unsigned)ns, (unsigned)((nrhs + NRC - 1) / NRC));
    hipLaunchKernelGGL(gemv_partial_kernel, g1, dim3(RB), 0, st, A, lda, n, nrhs, X, ldx, part,
                       cols_per);
    MPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(gemv_reduce_kernel, dim3(blocks_for(n), (unsigned)nrhs), dim3(RB), 0, st, n,
                       nrhs, ns, part, Y, ldy);
    MPC_LAUNCH_CHECK();
    return 0;
}

int64_t mpc_symv_partial_doubles(int64_t n, int nrhs) {
    const int64_t nt = (n + ST - 1) / ST;
    return nt * nt * nrhs * ST;
}

int mpc_launch_symv(hipStream_t st, const float* A, int64_t lda, int64_t n, int nrhs,
                    const double* X, int64_t ldx, double* Y, int64_t ldy, double* part) {
    const int64_t nt = (n + ST - 1) / ST;
    const int64_t ntile = nt * (nt + 1) / 2;
    hipLaunchKernelGGL(symv_tile_kernel, dim3((unsigned)ntile), dim3(256), 0, st, A, lda, n, (int)nt,
                       nrhs, X, ldx, part);
    MPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(symv_reduce_kernel, dim3(blocks_for(n), (unsigned)nrhs), dim3(256), 0, st, n,
                       (int)nt, nrhs, part, Y, ldy);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_trsv_step_fwd(hipStream_t st, const float* L, int64_t ldl, int64_t n, int nb,
                             int j, const float* dinv, double* W, int64_t ldw, int nrhs,
                             double* Y, int64_t ldy) {
    const int64_t c0 = (int64_t)j * nb;
    const int bs = (int)(n - c0 < nb ? n - c0 : nb);
    const int64_t below = n - c0 - bs;
    unsigned g = below > 0 ? blocks_for(below) : 1;
    hipLaunchKernelGGL(trsv_fwd_kernel, dim3(g), dim3(RB), 0, st, L, ldl, n, nb, c0, bs,
                       dinv + (c0 / 32) * 1024, W, ldw, nrhs, Y, ldy);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_trsv_step_bwd(hipStream_t st, const float* L, int64_t ldl, int64_t n, int nb,
                             int j, const float* dinv, double* W, int64_t ldw, int nrhs,
                             double* Xo, int64_t ldxo) {
    const int64_t c0 = (int64_t)j * nb;
    const int bs = (int)(n - c0 < nb ? n - c0 : nb);
    const int cpb = 64;
    unsigned g = c0 > 0 ? (unsigned)((c0 + cpb - 1) / cpb) : 1;
    hipLaunchKernelGGL(trsv_bwd_kernel, dim3(g), dim3(RB), 0, st, L, ldl, n, nb, c0, bs,
                       dinv + (c0 / 32) * 1024, W, ldw, nrhs, cpb, Xo, ldxo);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_dot(hipStream_t st, int64_t n, int nrhs, const double* a, int64_t lda_,
                   const double* b, int64_t ldb, double* out, double* part) {
    const unsigned nparts = (unsigned)((n + RB * 8 - 1) / (RB * 8));
    hipLaunchKernelGGL(dot_partial_kernel, dim3(nparts, (unsigned)nrhs), dim3(RB), 0, st, n, a,
                       lda_, b, ldb, part);
    MPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(dot_final_kernel, dim3((unsigned)nrhs), dim3(RB), 0, st, (int)nparts, part,
                       out);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_pcg_update_xr(hipStream_t st, int64_t n, int nrhs, double* X, int64_t ldx,
                             double* R, int64_t ldr, const double* P, int64_t ldp,
                             const double* Q, int64_t ldq, const double* rz, const double* pq,
                             const int* active) {
    hipLaunchKernelGGL(update_xr_kernel, dim3(blocks_for(n), (unsigned)nrhs), dim3(RB), 0, st, n, X,
                       ldx, R, ldr, P, ldp, Q, ldq, rz, pq, active);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_pcg_update_p(hipStream_t st, int64_t n, int nrhs, double* P, int64_t ldp,
                            const double* Z, int64_t ldz, const double* rz_new,
                            const double* rz_old, const int* active) {
    hipLaunchKernelGGL(update_p_kernel, dim3(blocks_for(n), (unsigned)nrhs), dim3(RB), 0, st, n, P,
                       ldp, Z, ldz, rz_new, rz_old, active);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_axpby(hipStream_t st, int64_t n, int nrhs, double a, const double* X, int64_t ldx,
                     double b, double* Y, int64_t ldy) {
    hipLaunchKernelGGL(axpby_kernel, dim3(blocks_for(n), (unsigned)nrhs), dim3(RB), 0, st, n, a, X,
                       ldx, b, Y, ldy);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_copy2d(hipStream_t st, int64_t n, int nrhs, const double* X, int64_t ldx,
                      double* Y, int64_t ldy) {
    hipLaunchKernelGGL(copy2d_kernel, dim3(blocks_for(n), (unsigned)nrhs), dim3(RB), 0, st, n, X,
                       ldx, Y, ldy);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_copy_f32(hipStream_t st, int64_t rows, int64_t cols, const float* A, int64_t lda,
                        float* B, int64_t ldb) {
    unsigned gy = (unsigned)(cols < 4096 ? cols : 4096);
    hipLaunchKernelGGL(copy_f32_kernel, dim3(blocks_for(rows), gy), dim3(RB), 0, st, rows, cols, A,
                       lda, B, ldb);
    MPC_LAUNCH_CHECK();
    return 0;
}
