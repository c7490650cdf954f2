// k_pcg.hip -- bandwidth-bound kernels of the PCG loop.
//
// Replaces, from precond_CG (mixed_cholesky.cu:145-217): cublasDgemv (A p),
// cusolverDnDpotrs (z = L^-T L^-1 r, on an FP64 copy of L made by
// convertFloattoDouble), and the cublasDdot/Daxpy/Dscal/Dcopy calls that
// return scalars to the host.  Here A and L stay FP32 in HBM (A may also be an
// FP64 operator, for solves that must match FP64 exactly) and are read once
// per product; vectors, accumulation, inner products and the alpha/beta
// scalars are FP64 and never leave the device.  Every reduction has a fixed
// order (two-stage, partials then an ordered sum), so results are bitwise
// reproducible run to run.
#include "mpc_internal.hpp"

namespace {

constexpr int RB = 256;          // rows per workgroup
constexpr int NRC = 8;           // right-hand sides per pass

// ---------------------------------------------------------------- SYMV
// y = A x for symmetric A reading only the upper triangle (i <= j) of the
// column-major FP32 (or FP64) matrix: half the HBM traffic of a full GEMV.
// One workgroup per ST x ST tile (I <= J), staged in LDS once; its row sums
// (A_IJ x_J) and column sums (A_IJ^T x_I) go to fixed partial slots
//   P[K][L][r]:  L >= K row-sum partial of tile (K, L); L < K column-sum
//               partial of tile (L, K),
// and symv_reduce sums P[K][0..nt) in order (deterministic).  Tile edge 64:
// 17 KB (FP32) or 33 KB (FP64) of LDS, so several tiles share a CU and one
// tile's loads overlap another's sums.  (A 128 edge for FP32, 65 KB and two
// workgroups per CU, was slower.)
constexpr int SYMV_ST = 64;
template <typename T> struct vec16 { using type = float4; };
template <> struct vec16<double> { using type = double2; };

template <typename T>
__global__ __launch_bounds__(256) void symv_tile_kernel(const T* __restrict__ A, int64_t lda,
                                                       int64_t n, int nt, int nrhs,
                                                       const double* __restrict__ X, int64_t ldx,
                                                       double* __restrict__ P) {
    constexpr int ST = SYMV_ST;
    __shared__ T As[ST * (ST + 1)];
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
    // Stage the tile.  Full tiles of an aligned matrix: 16-byte loads, all of
    // a thread's loads issued before any LDS store (one HBM latency per tile).
    using V16 = typename vec16<T>::type;
    constexpr int VE = 16 / (int)sizeof(T);        // elements per 16-byte load
    constexpr int RV = ST / VE;                     // loads per tile column
    constexpr int CPP = 256 / RV;                   // columns per pass
    constexpr int NP = ST / CPP;                    // passes
    const bool fast = r0 + ST <= n && c0 + ST <= n && lda % VE == 0 &&
                      (reinterpret_cast<uintptr_t>(A) & 15) == 0;
    if (fast) {
        const int q = tid % RV, cc = tid / RV;
        V16 v[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p)
            v[p] = *reinterpret_cast<const V16*>(A + r0 + VE * q + (c0 + p * CPP + cc) * lda);
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const int c = p * CPP + cc;
            const T* e = reinterpret_cast<const T*>(&v[p]);
#pragma unroll
            for (int u = 0; u < VE; ++u) {
                const int r = VE * q + u;
                As[c * (ST + 1) + r] = (!diag || r <= c) ? e[u] : T(0);
            }
        }
    } else {
        for (int idx = tid; idx < ST * ST; idx += 256) {
            const int r = idx % ST, c = idx / ST;
            const int64_t gr = r0 + r, gc = c0 + c;
            T v = 0;
            if (gr < n && gc < n && (!diag || r <= c)) v = A[gr + gc * lda];
            As[c * (ST + 1) + r] = v;
        }
    }
    for (int j = 0; j < nrhs; ++j) {
        if (tid < ST) {
            xi[tid] = r0 + tid < n ? X[j * ldx + r0 + tid] : 0.0;
            xj[tid] = c0 + tid < n ? X[j * ldx + c0 + tid] : 0.0;
        }
        __syncthreads();
        // threads [0, ST): row sums of row tid; [ST, 2 ST): column sums of column tid - ST
        double s = 0.0;
        if (tid < ST) {
            const int r = tid;
            for (int c = 0; c < ST; ++c) s = fma((double)As[c * (ST + 1) + r], xj[c], s);
            if (diag) {   // mirrored strict lower part: sum_{c < r} A[c][r] x[c]
                for (int c = 0; c < r; ++c) s = fma((double)As[r * (ST + 1) + c], xi[c], s);
            }
            P[((I * nt + J) * nrhs + j) * ST + r] = s;
        } else if (tid < 2 * ST && !diag) {
            const int c = tid - ST;
            for (int r = 0; r < ST; ++r) s = fma((double)As[c * (ST + 1) + r], xi[r], s);
            P[((J * nt + I) * nrhs + j) * ST + c] = s;
        }
        __syncthreads();
    }
}

// 64 rows per workgroup (lane = row).  Wave w sums the partials of tiles
// [w q, (w+1) q), q = ceil(nt / 4), in batches of 16 independent loads
// (clamped tile index, zero weight past the end) in tile order; the four
// quarter sums are then added in order.
template <int ST>
__global__ __launch_bounds__(256) void symv_reduce_kernel(int64_t n, int nt, int nrhs,
                                                         const double* __restrict__ P,
                                                         double* __restrict__ Y, int64_t ldy) {
    __shared__ double red[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t row0 = (int64_t)blockIdx.x * 64 + lane;
    const int64_t row = row0 < n ? row0 : n - 1;
    const int j = blockIdx.y;
    const int64_t K = row / ST, r = row % ST;
    const int q = (nt + 3) / 4, l0 = w * q, l1 = l0 + q < nt ? l0 + q : nt;
    double s = 0.0;
    for (int L0 = l0; L0 < l1; L0 += 16) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int64_t L = L0 + u < l1 ? L0 + u : l1 - 1;
            v[u] = P[((K * nt + L) * nrhs + j) * ST + r];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) s += L0 + u < l1 ? v[u] : 0.0;
    }
    red[w][lane] = s;
    __syncthreads();
    if (w == 0 && row0 < n)
        Y[j * ldy + row0] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

// ---------------------------------------------------------------- dots and updates
// blockIdx.z selects the second operand: 0 -> b, 1 -> b2 (two dots with the
// same first operand in one launch); partials of z at part + z nrhs nparts
__global__ __launch_bounds__(RB) void dot_partial_kernel(int64_t n, const double* __restrict__ a,
                                                        int64_t lda_, const double* __restrict__ b0,
                                                        int64_t ldb0, const double* __restrict__ b2,
                                                        int64_t ldb2, double* __restrict__ part) {
    __shared__ double red[RB];
    const int j = blockIdx.y;
    const double* __restrict__ b = blockIdx.z ? b2 : b0;
    const int64_t ldb = blockIdx.z ? ldb2 : ldb0;
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
    if (threadIdx.x == 0)
        part[((int64_t)blockIdx.z * gridDim.y + j) * gridDim.x + blockIdx.x] = red[0];
}

// blockIdx.y selects the dot: 0 -> out, 1 -> out2
__global__ __launch_bounds__(RB) void dot_final_kernel(int nparts, const double* __restrict__ part,
                                                      double* __restrict__ out0,
                                                      double* __restrict__ out2) {
    __shared__ double red[RB];
    const int j = blockIdx.x;
    double* __restrict__ out = blockIdx.y ? out2 : out0;
    part += (int64_t)blockIdx.y * gridDim.x * nparts;
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

// End of one PCG iteration, per right-hand side j that is still active:
// record column 0's CG coefficients (Lanczos estimate), rz <- rz_new, count
// the iteration, and test convergence (stop_rule 0: sqrt(rr)/||b||, the
// recursive residual, confirmed later on the true residual; 1: the reference
// rule sqrt(|r.z|)/||b||).  A column that meets tol becomes a candidate and is
// frozen (active = 0) so later iterations leave it untouched.
__global__ __launch_bounds__(256) void pcg_status_kernel(int nrhs, double* __restrict__ rz,
                                                        const double* __restrict__ rz_new,
                                                        const double* __restrict__ pq,
                                                        const double* __restrict__ rr,
                                                        const double* __restrict__ bb, double tol,
                                                        int stop_rule, int* __restrict__ active,
                                                        int* __restrict__ cand,
                                                        int* __restrict__ iters,
                                                        double* __restrict__ relres,
                                                        double* __restrict__ ahist,
                                                        double* __restrict__ bhist, int hmax) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= nrhs || !active[j]) return;
    const int k = iters[j];
    if (j == 0 && k < hmax) {
        ahist[k] = rz[0] / pq[0];
        bhist[k] = rz_new[0] / rz[0];
    }
    rz[j] = rz_new[j];
    iters[j] = k + 1;
    double nb = sqrt(bb[j]);
    if (nb == 0.0) nb = 1.0;
    const double res = stop_rule == 1 ? sqrt(fabs(rz_new[j])) / nb : sqrt(rr[j]) / nb;
    relres[j] = res;
    if (res <= tol) {
        cand[j] = 1;
        active[j] = 0;
    }
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

int64_t mpc_symv_partial_doubles(int64_t n, int nrhs) {
    const int64_t nt = (n + SYMV_ST - 1) / SYMV_ST;
    return nt * nt * SYMV_ST * nrhs;
}

template <typename T>
static int launch_symv_t(hipStream_t st, const T* A, int64_t lda, int64_t n, int nrhs,
                         const double* X, int64_t ldx, double* Y, int64_t ldy, double* part) {
    constexpr int ST = SYMV_ST;
    const int64_t nt = (n + ST - 1) / ST;
    const int64_t ntile = nt * (nt + 1) / 2;
    hipLaunchKernelGGL(symv_tile_kernel<T>, dim3((unsigned)ntile), dim3(256), 0, st, A, lda, n,
                       (int)nt, nrhs, X, ldx, part);
    hipLaunchKernelGGL(symv_reduce_kernel<ST>, dim3((unsigned)((n + 63) / 64), (unsigned)nrhs),
                       dim3(256), 0, st,
                       n, (int)nt, nrhs, part, Y, ldy);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_symv(hipStream_t st, const float* A, int64_t lda, int64_t n, int nrhs,
                    const double* X, int64_t ldx, double* Y, int64_t ldy, double* part) {
    return launch_symv_t<float>(st, A, lda, n, nrhs, X, ldx, Y, ldy, part);
}

int mpc_launch_symv64(hipStream_t st, const double* A, int64_t lda, int64_t n, int nrhs,
                      const double* X, int64_t ldx, double* Y, int64_t ldy, double* part) {
    return launch_symv_t<double>(st, A, lda, n, nrhs, X, ldx, Y, ldy, part);
}

int mpc_launch_dot(hipStream_t st, int64_t n, int nrhs, const double* a, int64_t lda_,
                   const double* b, int64_t ldb, double* out, double* part) {
    const unsigned nparts = (unsigned)((n + RB * 8 - 1) / (RB * 8));
    hipLaunchKernelGGL(dot_partial_kernel, dim3(nparts, (unsigned)nrhs, 1), dim3(RB), 0, st, n, a,
                       lda_, b, ldb, b, ldb, part);
    MPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(dot_final_kernel, dim3((unsigned)nrhs, 1), dim3(RB), 0, st, (int)nparts, part,
                       out, out);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_dot2(hipStream_t st, int64_t n, int nrhs, const double* a, int64_t lda_,
                    const double* b, int64_t ldb, double* out_b, const double* c, int64_t ldc,
                    double* out_c, double* part) {
    const unsigned nparts = (unsigned)((n + RB * 8 - 1) / (RB * 8));
    hipLaunchKernelGGL(dot_partial_kernel, dim3(nparts, (unsigned)nrhs, 2), dim3(RB), 0, st, n, a,
                       lda_, b, ldb, c, ldc, part);
    MPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(dot_final_kernel, dim3((unsigned)nrhs, 2), dim3(RB), 0, st, (int)nparts, part,
                       out_b, out_c);
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

int mpc_launch_pcg_status(hipStream_t st, int nrhs, double* rz, const double* rz_new,
                          const double* pq, const double* rr, const double* bb, double tol,
                          int stop_rule, int* active, int* cand, int* iters, double* relres,
                          double* ahist, double* bhist, int hmax) {
    hipLaunchKernelGGL(pcg_status_kernel, dim3((unsigned)((nrhs + 255) / 256)), dim3(256), 0, st,
                       nrhs, rz, rz_new, pq, rr, bb, tol, stop_rule, active, cand, iters, relres,
                       ahist, bhist, hmax);
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
