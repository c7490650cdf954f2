// k_matgen.hip -- device-side problem generation.
//
// generatePSD (matgen.cuh:188-298) builds A = Q^T Lambda Q from a QR of a
// random n x n matrix and two dense FP64 GEMMs (~5n^3 flops plus an FP64 copy
// of A).  Here A = H Lambda H^T with H = H_1..H_r (r <= 4 Householder
// reflectors, compact WY form H = I - W T W^T) is written in ONE pass over the
// n x n FP32 output: entry (i, j) needs only four 4-vectors per row,
//   A_ij = lam_i [i==j] - lam_j <p_i, w_j> - lam_i <p_j, w_i> + <p_i, g_j>,
// P = W T, g = (W^T Lambda W) p, all FP64 with explicit fma in a fixed order,
// evaluated at (min(i,j), max(i,j)) so A is exactly symmetric.  The HBM write
// of A is the only cost (4 n^2 bytes).
//
// Points for the RBF workloads come from a counter-based Philox4x32-10 stream
// (bit-identical to the host generator); the RBF kernel matrix is assembled
// in 64x64 tiles from LDS-staged, lengthscale-scaled coordinates.
#include "mpc_internal.hpp"

namespace {

__device__ __forceinline__ double dot4(const double* a, const double* b) {
    return fma(a[0], b[0], fma(a[1], b[1], fma(a[2], b[2], a[3] * b[3])));
}

// each thread: one row i, columns [blockIdx.y * cols, ...)
__global__ __launch_bounds__(256) void hh_matrix_kernel(int64_t n, const double* __restrict__ lam,
                                                       const double* __restrict__ W,
                                                       const double* __restrict__ P,
                                                       const double* __restrict__ g,
                                                       float* __restrict__ A, int64_t lda,
                                                       int64_t cols) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double wi[4], pi[4], gi[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        wi[a] = W[4 * i + a];
        pi[a] = P[4 * i + a];
        gi[a] = g[4 * i + a];
    }
    const double li = lam[i];
    int64_t j0 = (int64_t)blockIdx.y * cols, j1 = j0 + cols;
    if (j1 > n) j1 = n;
    for (int64_t j = j0; j < j1; ++j) {
        double wj[4], pj[4], gj[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            wj[a] = W[4 * j + a];
            pj[a] = P[4 * j + a];
            gj[a] = g[4 * j + a];
        }
        const double lj = lam[j];
        // (lo, hi) = (min, max)
        const bool ilo = i <= j;
        const double* plo = ilo ? pi : pj;
        const double* phi = ilo ? pj : pi;
        const double* wlo = ilo ? wi : wj;
        const double* whi = ilo ? wj : wi;
        const double* ghi = ilo ? gj : gi;
        const double llo = ilo ? li : lj, lhi = ilo ? lj : li;
        double t = dot4(plo, ghi);
        t = fma(-lhi, dot4(plo, whi), t);
        t = fma(-llo, dot4(phi, wlo), t);
        if (i == j) t += llo;
        A[i + j * lda] = (float)t;
    }
}

// ---------------------------------------------------------------- Philox
__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
        uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    }
}
__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}
__device__ __forceinline__ double dev_uniform(uint64_t seed, uint64_t stream, uint64_t idx) {
    uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)stream, 0x554e4946u};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    return u53(c[0], c[1]);
}
__device__ __forceinline__ double dev_normal(uint64_t seed, uint64_t stream, uint64_t idx) {
    uint64_t q = idx >> 1;
    uint32_t c[4] = {(uint32_t)q, (uint32_t)(q >> 32), (uint32_t)stream, 0x4e4f524du};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    double u1 = 1.0 - u53(c[0], c[1]);
    double u2 = u53(c[2], c[3]);
    double r = sqrt(-2.0 * log(u1));
    double th = 6.283185307179586 * u2;
    return (idx & 1) ? r * sin(th) : r * cos(th);
}

__global__ __launch_bounds__(256) void points_kernel(float* __restrict__ X, int64_t n, int d,
                                                    int64_t ldx, int kind, double p_zero,
                                                    uint64_t seed) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n * d) return;
    const int64_t i = t / d;
    const int c = (int)(t - i * d);
    float v;
    if (kind == 0) {
        v = (float)dev_normal(seed, 0, (uint64_t)t);
    } else {
        double u = dev_uniform(seed, 1, (uint64_t)t);
        v = u < p_zero ? 0.0f : (float)(1.0 - dev_uniform(seed, 2, (uint64_t)t));
    }
    X[i * ldx + c] = v;
}

// ---------------------------------------------------------------- RBF
constexpr int RT = 64;      // output tile
constexpr int RD = 16;      // dims per LDS stage

// T = float: FP32 coordinates and sums (the factor's FP32 matrix).  T =
// double: the same kernel evaluated in FP64 (the exact operator for solves
// that must match FP64), with an optional FP32 copy rounded from it.
template <typename T>
__global__ __launch_bounds__(256) void rbf_kernel(const float* __restrict__ X, int64_t n, int d,
                                                 int64_t ldx, const T* __restrict__ inv_ls,
                                                 int n_ls, T variance, T nugget,
                                                 float* __restrict__ K32, int64_t ldk32,
                                                 double* __restrict__ K64, int64_t ldk64) {
    __shared__ T xi[RD][RT + 1];
    __shared__ T xj[RD][RT + 1];
    const int64_t i0 = (int64_t)blockIdx.x * RT, j0 = (int64_t)blockIdx.y * RT;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;   // 16 x 16 threads, 4x4 each
    T acc[4][4] = {};
    for (int c0 = 0; c0 < d; c0 += RD) {
        for (int idx = threadIdx.x; idx < RD * RT; idx += 256) {
            int p = idx / RD, c = idx % RD;
            int cc = c0 + c;
            const int ccl = cc < d ? cc : d - 1;
            const T s = inv_ls[n_ls == 1 ? 0 : ccl];
            const int64_t ri = i0 + p, rj = j0 + p;
            const T a = (T)X[(ri < n ? ri : n - 1) * ldx + ccl] * s;
            const T b = (T)X[(rj < n ? rj : n - 1) * ldx + ccl] * s;
            xi[c][p] = (cc < d && ri < n) ? a : T(0);
            xj[c][p] = (cc < d && rj < n) ? b : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < RD; ++c) {
            T a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) { a[u] = xi[c][tx + 16 * u]; b[u] = xj[c][ty + 16 * u]; }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    T df = a[u] - b[v];
                    acc[u][v] = fma(df, df, acc[u][v]);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            int64_t i = i0 + tx + 16 * u, j = j0 + ty + 16 * v;
            if (i < n && j < n) {
                T val = variance * exp(T(-0.5) * acc[u][v]);
                if (i == j) val += nugget;
                if (K64) K64[i + j * ldk64] = (double)val;
                if (K32) K32[i + j * ldk32] = (float)val;
            }
        }
}

// Rademacher probes z[p][i] = -1 if u < 1/2 else +1, u the Philox uniform of
// stream 3 at index p * n + i; or unit vectors z[p] = e_p (exact trace) when
// unit != 0.  Z is n x nprobe column-major (column p at Z + p * ldz).
__global__ __launch_bounds__(256) void probe_kernel(double* __restrict__ Z, int64_t ldz, int64_t n,
                                                   int nprobe, uint64_t seed, int unit) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y;
    if (i >= n || p >= nprobe) return;
    double v;
    if (unit) {
        v = (i == p) ? 1.0 : 0.0;
    } else {
        v = dev_uniform(seed, 3, (uint64_t)p * (uint64_t)n + (uint64_t)i) < 0.5 ? -1.0 : 1.0;
    }
    Z[(int64_t)p * ldz + i] = v;
}

__global__ __launch_bounds__(256) void convert_f8_kernel(const float* __restrict__ x,
                                                        uint8_t* __restrict__ y, int64_t count,
                                                        int e4m3) {
    const int64_t i = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 4;
    if (i >= count) return;
    float v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = i + q < count ? x[i + q] : 0.f;
    int w = 0;
    if (e4m3) {
        w = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], w, false);
        w = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], w, true);
    } else {
        w = __builtin_amdgcn_cvt_pk_bf8_f32(v[0], v[1], w, false);
        w = __builtin_amdgcn_cvt_pk_bf8_f32(v[2], v[3], w, true);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (i + q < count) y[i + q] = (uint8_t)((uint32_t)w >> (8 * q));
}

// inverses of the 32x32 diagonal sub-blocks of a lower factor (one wave each)
__global__ __launch_bounds__(64) void dinv_from_L_kernel(const float* __restrict__ L, int64_t ldl,
                                                        int64_t n, float* __restrict__ dinv) {
    const int64_t b = blockIdx.x;
    const int64_t r0 = b * 32;
    const int lane = threadIdx.x, c = lane & 31;
    __shared__ float Ls[32][33];
    for (int idx = lane; idx < 1024; idx += 64) {
        int r = idx & 31, cc = idx >> 5;
        int64_t gr = r0 + r, gc = r0 + cc;
        float v = (gr < n && gc < n) ? (r >= cc ? L[gr + gc * ldl] : 0.f) : (r == cc ? 1.f : 0.f);
        Ls[r][cc] = v;
    }
    __syncthreads();
    float x[32];
#pragma unroll
    for (int p = 0; p < 32; ++p) {
        float acc = (p == c) ? 1.f : 0.f;
#pragma unroll
        for (int q = 0; q < p; ++q) acc = fmaf(-Ls[p][q], x[q], acc);
        x[p] = acc / Ls[p][p];
    }
    if (lane < 32) {
#pragma unroll
        for (int p = 0; p < 32; ++p) dinv[b * 1024 + p + c * 32] = x[p];
    }
}

}  // namespace

int mpc_launch_hh_matrix(hipStream_t st, int64_t n, const double* lam, const double* W,
                         const double* P, const double* g, float* A, int64_t lda) {
    const int64_t cols = 64;
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)((n + cols - 1) / cols));
    hipLaunchKernelGGL(hh_matrix_kernel, grid, dim3(256), 0, st, n, lam, W, P, g, A, lda, cols);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_points(hipStream_t st, float* X, int64_t n, int d, int64_t ldx, int kind,
                      double p_zero, uint64_t seed) {
    int64_t total = n * d;
    hipLaunchKernelGGL(points_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, X, n,
                       d, ldx, kind, p_zero, seed);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_rbf(hipStream_t st, const float* X, int64_t n, int d, int64_t ldx,
                   const float* inv_ls, int n_ls, double variance, double nugget, float* K,
                   int64_t ldk) {
    dim3 grid((unsigned)((n + RT - 1) / RT), (unsigned)((n + RT - 1) / RT));
    hipLaunchKernelGGL(rbf_kernel<float>, grid, dim3(256), 0, st, X, n, d, ldx, inv_ls, n_ls,
                       (float)variance, (float)nugget, K, ldk, (double*)nullptr, (int64_t)0);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_rbf64(hipStream_t st, const float* X, int64_t n, int d, int64_t ldx,
                     const double* inv_ls, int n_ls, double variance, double nugget, double* K64,
                     int64_t ldk64, float* K32, int64_t ldk32) {
    dim3 grid((unsigned)((n + RT - 1) / RT), (unsigned)((n + RT - 1) / RT));
    hipLaunchKernelGGL(rbf_kernel<double>, grid, dim3(256), 0, st, X, n, d, ldx, inv_ls, n_ls,
                       variance, nugget, K32, ldk32, K64, ldk64);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_probes(hipStream_t st, double* Z, int64_t ldz, int64_t n, int nprobe, uint64_t seed,
                      int unit) {
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)nprobe);
    hipLaunchKernelGGL(probe_kernel, grid, dim3(256), 0, st, Z, ldz, n, nprobe, seed, unit);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_convert_f8(hipStream_t st, const float* x, uint8_t* y, int64_t count, int prec) {
    int64_t nthr = (count + 3) / 4;
    hipLaunchKernelGGL(convert_f8_kernel, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, st, x,
                       y, count, prec == MPC_PREC_E4M3 ? 1 : 0);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_dinv_from_L(hipStream_t st, const float* L, int64_t ldl, int64_t n, float* dinv) {
    hipLaunchKernelGGL(dinv_from_L_kernel, dim3((unsigned)((n + 31) / 32)), dim3(64), 0, st, L, ldl,
                       n, dinv);
    MPC_LAUNCH_CHECK();
    return 0;
}
