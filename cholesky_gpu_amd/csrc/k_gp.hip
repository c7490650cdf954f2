// k_gp.hip -- fused gradient contraction of the GP loss.
//
// No counterpart in the reference (its RBF.cu is an empty stub); the GP
// workload is defined by the differentiable-solve configuration.  For
//   loss = 1/2 b^T A^-1 b + 1/2 logdet A,
//   A_ij = variance * exp(-1/2 sum_c ((x_ic - x_jc) / ls_c)^2) + noise [i == j],
// every parameter theta has dloss/dtheta = sum_ij M_ij dA_ij/dtheta with
//   M = -1/2 alpha alpha^T + (w / 2) sum_p A^-1 z_p z_p^T = U V^T,
//   U = [-alpha/2, w/2 A^-1 z_1, ...],  V = [alpha, z_1, ...],
// (alpha = A^-1 b; w = 1/P for Rademacher probes, 1 for unit probes, which
// make the trace exact).  dA/dtheta is symmetric, so only the lower triangle
// is visited with the symmetrised weight m_ij = (U V^T)_ij + (U V^T)_ji.
// The kernel entries are recomputed in FP64 from the FP32 points instead of
// being stored; per entry:
//   dA_ij/dls_c     = K_ij (x_ic - x_jc)^2 / ls_c^3   (isotropic: summed over c)
//   dA_ij/dvariance = K_ij / variance
//   dA_ij/dnoise    = [i == j]
// Each workgroup writes its partial sums to a fixed slot; a second kernel adds
// the slots in order, so the result is bitwise reproducible.
#include "mpc_internal.hpp"

namespace {

constexpr int GT = 64;        // tile edge
constexpr int GRC = 32;       // solve columns staged per pass
constexpr int GLD = GRC + 1;

__global__ __launch_bounds__(256) void gp_grad_tile_kernel(
    const float* __restrict__ X, int64_t n, int d, int64_t ldx, const double* __restrict__ inv_ls,
    int n_ls, double variance, const double* __restrict__ U, const double* __restrict__ V,
    int64_t ldu, int R, int nt, double* __restrict__ part) {
    __shared__ double Ui[GT * GLD], Vi[GT * GLD], Uj[GT * GLD], Vj[GT * GLD];
    __shared__ double xi[MPC_GP_MAXD][GT], xj[MPC_GP_MAXD][GT];
    __shared__ double red[4][MPC_GP_MAXD + 2];
    // lower tile (I >= J): t = I (I + 1) / 2 + J
    const int64_t t = blockIdx.x;
    int64_t I = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (I * (I + 1) / 2 > t) --I;
    while ((I + 1) * (I + 2) / 2 <= t) ++I;
    const int64_t J = t - I * (I + 1) / 2;
    const int64_t i0 = I * GT, j0 = J * GT;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    (void)nt;

    for (int idx = tid; idx < d * GT; idx += 256) {
        const int c = idx / GT, p = idx % GT;
        const double s = inv_ls[n_ls == 1 ? 0 : c];
        const int64_t ri = i0 + p < n ? i0 + p : n - 1, rj = j0 + p < n ? j0 + p : n - 1;
        const double a = (double)X[ri * ldx + c] * s, b = (double)X[rj * ldx + c] * s;
        xi[c][p] = i0 + p < n ? a : 0.0;
        xj[c][p] = j0 + p < n ? b : 0.0;
    }
    // m[u][v] for i = i0 + tx + 16u, j = j0 + ty + 16v
    double m[4][4] = {};
    for (int r0 = 0; r0 < R; r0 += GRC) {
        __syncthreads();
        for (int idx = tid; idx < GT * GRC; idx += 256) {
            const int p = idx % GT, r = idx / GT;
            const bool okr = r0 + r < R;
            const int64_t gi = i0 + p, gj = j0 + p;
            const int64_t gic = gi < n ? gi : n - 1, gjc = gj < n ? gj : n - 1;
            const int64_t o = (int64_t)(okr ? r0 + r : R - 1) * ldu;
            const double ui = U[o + gic], vi = V[o + gic], uj = U[o + gjc], vj = V[o + gjc];
            Ui[p * GLD + r] = okr && gi < n ? ui : 0.0;
            Vi[p * GLD + r] = okr && gi < n ? vi : 0.0;
            Uj[p * GLD + r] = okr && gj < n ? uj : 0.0;
            Vj[p * GLD + r] = okr && gj < n ? vj : 0.0;
        }
        __syncthreads();
        for (int r = 0; r < GRC; ++r) {
            double ui[4], vi[4], uj[4], vj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                ui[u] = Ui[(tx + 16 * u) * GLD + r];
                vi[u] = Vi[(tx + 16 * u) * GLD + r];
                uj[u] = Uj[(ty + 16 * u) * GLD + r];
                vj[u] = Vj[(ty + 16 * u) * GLD + r];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) m[u][v] = fma(ui[u], vj[v], fma(uj[v], vi[u], m[u][v]));
        }
    }
    // m = (U V^T)_ij + (U V^T)_ji, which counts the diagonal twice: halved there
    double gl[MPC_GP_MAXD], gvar = 0.0, gnoise = 0.0;
#pragma unroll
    for (int c = 0; c < MPC_GP_MAXD; ++c) gl[c] = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int pi = tx + 16 * u, pj = ty + 16 * v;
            const int64_t i = i0 + pi, j = j0 + pj;
            if (i >= n || j >= n || j > i) continue;
            double w = m[u][v];
            if (i == j) {
                w *= 0.5;
                gnoise += w;
            }
            double s2 = 0.0;
            for (int c = 0; c < d; ++c) {
                const double df = xi[c][pi] - xj[c][pj];
                s2 = fma(df, df, s2);
            }
            const double kij = variance * exp(-0.5 * s2);
            const double wk = w * kij;
            gvar += wk;
            if (n_ls == 1) {
                gl[0] += wk * s2;          // times 1/ls on the host
            } else {
                for (int c = 0; c < d; ++c) {
                    const double df = xi[c][pi] - xj[c][pj];
                    gl[c] = fma(wk, df * df, gl[c]);
                }
            }
        }
    // workgroup reduction (wave shuffles, then 4 waves in order)
    const int nout = (n_ls == 1 ? 1 : d) + 2;
    const int lane = tid & 63, wave = tid >> 6;
    for (int q = 0; q < nout; ++q) {
        double v = q < nout - 2 ? gl[q] : (q == nout - 2 ? gvar : gnoise);
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) red[wave][q] = v;
    }
    __syncthreads();
    if (tid < nout) part[t * nout + tid] = ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
}

// out[q] = sum over tiles of part[t][q], tiles added in a fixed order
__global__ __launch_bounds__(256) void gp_grad_reduce_kernel(const double* __restrict__ part,
                                                            int64_t ntile, int nout,
                                                            double* __restrict__ out) {
    __shared__ double s[256];
    const int q = blockIdx.x, tid = threadIdx.x;
    double a = 0.0;
    for (int64_t t = tid; t < ntile; t += 256) a += part[t * nout + q];
    s[tid] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) s[tid] += s[tid + w];
        __syncthreads();
    }
    if (tid == 0) out[q] = s[0];
}

// out = sum_i log L_ii (fixed order: strided per thread, then a tree)
__global__ __launch_bounds__(256) void logdiag_kernel(const float* __restrict__ L, int64_t ldl,
                                                     int64_t n, double* __restrict__ out) {
    __shared__ double s[256];
    const int tid = threadIdx.x;
    double a = 0.0;
    for (int64_t i = tid; i < n; i += 256) a += log((double)L[i + i * ldl]);
    s[tid] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) s[tid] += s[tid + w];
        __syncthreads();
    }
    if (tid == 0) *out = s[0];
}

}  // namespace

int mpc_launch_logdiag(hipStream_t st, const float* L, int64_t ldl, int64_t n, double* out) {
    hipLaunchKernelGGL(logdiag_kernel, dim3(1), dim3(256), 0, st, L, ldl, n, out);
    MPC_LAUNCH_CHECK();
    return 0;
}

int64_t mpc_gp_partial_doubles(int64_t n, int nout) {
    const int64_t nt = (n + GT - 1) / GT;
    return nt * (nt + 1) / 2 * nout;
}

int mpc_launch_gp_grad(hipStream_t st, const float* X, int64_t n, int d, int64_t ldx,
                       const double* inv_ls, int n_ls, double variance, const double* U,
                       const double* V, int64_t ldu, int R, double* part, double* out) {
    if (d > MPC_GP_MAXD || d < 1 || (n_ls != 1 && n_ls != d)) {
        mpc_set_error("gp gradient: d must be 1..%d and n_ls 1 or d", MPC_GP_MAXD);
        return MPC_ERR_ARG;
    }
    const int64_t nt = (n + GT - 1) / GT;
    const int64_t ntile = nt * (nt + 1) / 2;
    const int nout = (n_ls == 1 ? 1 : d) + 2;
    hipLaunchKernelGGL(gp_grad_tile_kernel, dim3((unsigned)ntile), dim3(256), 0, st, X, n, d, ldx,
                       inv_ls, n_ls, variance, U, V, ldu, R, (int)nt, part);
    hipLaunchKernelGGL(gp_grad_reduce_kernel, dim3((unsigned)nout), dim3(256), 0, st, part, ntile,
                       nout, out);
    MPC_LAUNCH_CHECK();
    return 0;
}
