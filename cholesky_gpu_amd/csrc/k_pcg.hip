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

// ---------------------------------------------------------------- SYMV
// y = A x for symmetric A reading only the upper triangle (i <= j) of the
// column-major FP32 (or FP64) matrix: half the HBM traffic of a full GEMV.
// One workgroup per SY_E x SY_E super-tile (I <= J) of SY_S x SY_S tiles of
// 64, walked tile by tile through LDS (the next tile's loads are in flight
// while the current one is summed).  Waves 0-1 form row sums (A_IJ x_J: lane
// = row, wave = half of the tile's columns), waves 2-3 column sums
// (A_IJ^T x_I: lane = column, wave = half of the rows), accumulated in
// registers over the super-tile, so each super-tile writes one partial per
// row of I and one per row of J:
//   P[K][L][j][r]:  L >= K row-sum partial of super-tile (K, L); L < K
//                   column-sum partial of super-tile (L, K)
// (a diagonal super-tile writes its row and column sums as one partial), and
// symv_reduce adds P[K][0..nst) in order (deterministic).  The partials take
// nst^2 SY_E SJ doubles, SJ <= 8 right-hand sides per pass: n^2 SJ / 32
// bytes, 1/128 of the FP32 matrix per right-hand side; more right-hand sides
// than 8 run in passes of 8 over the same buffer.
constexpr int SY_T = 64;
constexpr int SY_S = 4;
constexpr int SY_E = SY_T * SY_S;
constexpr int SY_JMAX = 8;
template <typename T> struct vec16 { using type = float4; };
template <> struct vec16<double> { using type = double2; };

template <typename T, int NJ>
__global__ __launch_bounds__(256) void symv_super_kernel(const T* __restrict__ A, int64_t lda,
                                                        int64_t n, int nst, int j0, int nj,
                                                        const double* __restrict__ X, int64_t ldx,
                                                        double* __restrict__ P) {
    constexpr int ST = SY_T, S = SY_S, E = SY_E;
    using V16 = typename vec16<T>::type;
    constexpr int VE = 16 / (int)sizeof(T);        // elements per 16-byte load
    constexpr int RV = ST / VE;                     // loads per tile column
    constexpr int CPP = 256 / RV;                   // columns per pass
    constexpr int NP = ST / CPP;                    // passes
    static_assert(S * NJ * 64 <= ST * (ST + 1) * (int)sizeof(T) / 8, "column buffer fits As");
    __shared__ T As[ST * (ST + 1)];
    __shared__ double xr[NJ][E], xc[NJ][E];
    __shared__ double hb[NJ][64];
    // super-tile index -> (I, J), I <= J, row-major over the upper triangle
    const int64_t t = blockIdx.x;
    int64_t I = (int64_t)((2.0 * nst + 1.0 - sqrt((2.0 * nst + 1.0) * (2.0 * nst + 1.0) - 8.0 * t)) * 0.5);
    auto start = [&](int64_t i) { return i * nst - i * (i - 1) / 2; };
    while (I > 0 && start(I) > t) --I;
    while (start(I + 1) <= t) ++I;
    const int64_t J = I + (t - start(I));
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t R0 = I * E, C0 = J * E;
    const bool diag = I == J;
    const int na = (int)((n - R0 + ST - 1) / ST < S ? (n - R0 + ST - 1) / ST : S);
    const int nbb = (int)((n - C0 + ST - 1) / ST < S ? (n - C0 + ST - 1) / ST : S);
    // x of the rows (xr) and of the columns (xc); a diagonal super-tile reads
    // both from xr and keeps its row partials in xc
    for (int idx = tid; idx < NJ * E; idx += 256) {
        const int j = idx / E, k = idx % E;
        const int jj = j < nj ? j : nj - 1;
        const double vr = X[(int64_t)(j0 + jj) * ldx + (R0 + k < n ? R0 + k : n - 1)];
        xr[j][k] = (j < nj && R0 + k < n) ? vr : 0.0;
        const double vc = X[(int64_t)(j0 + jj) * ldx + (C0 + k < n ? C0 + k : n - 1)];
        xc[j][k] = (!diag && j < nj && C0 + k < n) ? vc : 0.0;
    }
    const double* xcol = diag ? &xr[0][0] : &xc[0][0];
    const bool fast_ok = lda % VE == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0;
    const int q = tid % RV, cq = tid / RV;
    V16 v[NP];
    auto load_tile = [&](int a, int b) {
        const int64_t r0 = R0 + a * ST, c0 = C0 + b * ST;
        if (fast_ok && r0 + ST <= n && c0 + ST <= n) {
#pragma unroll
            for (int p = 0; p < NP; ++p)
                v[p] = *reinterpret_cast<const V16*>(A + r0 + VE * q + (c0 + p * CPP + cq) * lda);
        } else {
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                T* e = reinterpret_cast<T*>(&v[p]);
                const int64_t gc = c0 + p * CPP + cq;
#pragma unroll
                for (int u = 0; u < VE; ++u) {
                    const int64_t gr = r0 + VE * q + u;
                    const T x = A[(gr < n ? gr : n - 1) + (gc < n ? gc : n - 1) * lda];
                    e[u] = (gr < n && gc < n) ? x : T(0);
                }
            }
        }
    };
    auto store_tile = [&](bool dtile) {
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const int c = p * CPP + cq;
            const T* e = reinterpret_cast<const T*>(&v[p]);
#pragma unroll
            for (int u = 0; u < VE; ++u) {
                const int r = VE * q + u;
                As[c * (ST + 1) + r] = (!dtile || r <= c) ? e[u] : T(0);
            }
        }
    };
    const bool rowrole = wave < 2;
    const int half = wave & 1;
    double cacc[S][NJ];
#pragma unroll
    for (int b = 0; b < S; ++b)
#pragma unroll
        for (int j = 0; j < NJ; ++j) cacc[b][j] = 0.0;
    if (na > 0) load_tile(0, 0);
    for (int a = 0; a < na; ++a) {
        const int b0 = diag ? a : 0;
        double racc[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j) racc[j] = 0.0;
#pragma unroll
        for (int b = 0; b < S; ++b) {
            if (b < b0 || b >= nbb) continue;
            const bool dtile = diag && a == b;
            __syncthreads();                          // the previous tile is summed
            store_tile(dtile);
            {                                         // the next tile's loads
                int an = a, bn = b + 1;
                if (bn >= nbb) { an = a + 1; bn = diag ? an : 0; }
                if (an < na) load_tile(an, bn);
            }
            __syncthreads();
            if (rowrole) {
                const int r = lane;
                double s[NJ];
#pragma unroll
                for (int j = 0; j < NJ; ++j) s[j] = 0.0;
#pragma unroll 8
                for (int cc = 0; cc < 32; ++cc) {
                    const int c = half * 32 + cc;
                    const double av = (double)As[c * (ST + 1) + r];
#pragma unroll
                    for (int j = 0; j < NJ; ++j) s[j] = fma(av, xcol[j * E + b * ST + c], s[j]);
                }
#pragma unroll
                for (int j = 0; j < NJ; ++j) racc[j] += s[j];
            } else {
                const int c = lane;
                double s[NJ];
#pragma unroll
                for (int j = 0; j < NJ; ++j) s[j] = 0.0;
                if (dtile) {                          // strict lower mirror only
#pragma unroll 8
                    for (int rr = 0; rr < 32; ++rr) {
                        const int r = half * 32 + rr;
                        const double av = r < c ? (double)As[c * (ST + 1) + r] : 0.0;
#pragma unroll
                        for (int j = 0; j < NJ; ++j) s[j] = fma(av, xr[j][a * ST + r], s[j]);
                    }
                } else {
#pragma unroll 8
                    for (int rr = 0; rr < 32; ++rr) {
                        const int r = half * 32 + rr;
                        const double av = (double)As[c * (ST + 1) + r];
#pragma unroll
                        for (int j = 0; j < NJ; ++j) s[j] = fma(av, xr[j][a * ST + r], s[j]);
                    }
                }
#pragma unroll
                for (int j = 0; j < NJ; ++j) cacc[b][j] += s[j];
            }
        }
        // row partials of tile-row a: the two column halves added in order
        if (rowrole && half == 1) {
#pragma unroll
            for (int j = 0; j < NJ; ++j) hb[j][lane] = racc[j];
        }
        __syncthreads();
        if (rowrole && half == 0) {
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const double tot = racc[j] + hb[j][lane];
                if (diag) xc[j][a * ST + lane] = tot;
                else if (j < nj) P[((I * nst + J) * NJ + j) * E + a * ST + lane] = tot;
            }
        }
    }
    // column partials: the two row halves added in order (through As), plus
    // the row partials of a diagonal super-tile
    __syncthreads();
    double* cb = reinterpret_cast<double*>(As);
    if (!rowrole && half == 1) {
#pragma unroll
        for (int b = 0; b < S; ++b)
#pragma unroll
            for (int j = 0; j < NJ; ++j) cb[(b * NJ + j) * 64 + lane] = cacc[b][j];
    }
    __syncthreads();
    if (!rowrole && half == 0) {
#pragma unroll
        for (int b = 0; b < S; ++b) {
            if (b >= nbb) continue;
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                if (j >= nj) continue;
                double tot = cacc[b][j] + cb[(b * NJ + j) * 64 + lane];
                if (diag) tot = xc[j][b * ST + lane] + tot;
                P[((J * nst + I) * NJ + j) * E + b * ST + lane] = tot;
            }
        }
    }
}

// 64 rows per workgroup (lane = row).  Wave w sums the partials of
// super-tiles [w q, (w+1) q), q = ceil(nst / 4), in batches of 16
// independent loads (clamped index, zero weight past the end) in order; the
// four quarter sums are then added in order.
__global__ __launch_bounds__(256) void symv_reduce_kernel(int64_t n, int nst, int NJ, int j0,
                                                         const double* __restrict__ P,
                                                         double* __restrict__ Y, int64_t ldy) {
    __shared__ double red[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t row0 = (int64_t)blockIdx.x * 64 + lane;
    const int64_t row = row0 < n ? row0 : n - 1;
    const int j = blockIdx.y;
    const int64_t K = row / SY_E, r = row % SY_E;
    const int q = (nst + 3) / 4, l0 = w * q, l1 = l0 + q < nst ? l0 + q : nst;
    double s = 0.0;
    for (int L0 = l0; L0 < l1; L0 += 16) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int64_t L = L0 + u < l1 ? L0 + u : l1 - 1;
            v[u] = P[((K * nst + L) * NJ + j) * SY_E + r];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) s += L0 + u < l1 ? v[u] : 0.0;
    }
    red[w][lane] = s;
    __syncthreads();
    if (w == 0 && row0 < n)
        Y[(int64_t)(j0 + j) * ldy + row0] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
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

// ---------------------------------------------------------------- block PCG
// s x s products and updates of the block recurrence (s <= MPC_BLOCK_PCG_MAX
// right-hand sides, n x s blocks column-major).
constexpr int BS = MPC_BLOCK_PCG_MAX;   // largest block
constexpr int GCH = 2048;         // rows per gram workgroup
constexpr int GSUB = 128;         // rows staged in LDS at a time

// G = X^T Y: workgroup c forms the s^2 partial dots of rows
// [c GCH, (c+1) GCH) (thread p = a + s b, rows ascending) and the last one
// to finish adds the partials of every chunk in order.
__global__ __launch_bounds__(256) void gram_kernel(int64_t n, int s, const double* __restrict__ X,
                                                  int64_t ldx, const double* __restrict__ Y,
                                                  int64_t ldy, double* __restrict__ part,
                                                  double* __restrict__ G, int* __restrict__ tick) {
    __shared__ double xs[BS][GSUB + 1], ys[BS][GSUB + 1];
    const int tid = threadIdx.x;
    const int s2 = s * s;
    const int a = tid % s, b = tid / s;
    const int64_t r0 = (int64_t)blockIdx.x * GCH;
    double acc = 0.0;
    for (int64_t q0 = r0; q0 < r0 + GCH && q0 < n; q0 += GSUB) {
        for (int idx = tid; idx < s * GSUB; idx += 256) {
            const int c = idx / GSUB, r = idx % GSUB;
            const int64_t gr = q0 + r;
            const int64_t rc = gr < n ? gr : n - 1;
            const double vx = X[c * ldx + rc], vy = Y[c * ldy + rc];
            xs[c][r] = gr < n ? vx : 0.0;
            ys[c][r] = gr < n ? vy : 0.0;
        }
        __syncthreads();
        if (tid < s2) {
#pragma unroll 8
            for (int r = 0; r < GSUB; ++r) acc = fma(xs[a][r], ys[b][r], acc);
        }
        __syncthreads();
    }
    if (tid < s2) part[(int64_t)blockIdx.x * s2 + tid] = acc;
    if (!mpc_last_arrival(tick, (int)gridDim.x)) return;
    if (tid < s2) {
        double t = 0.0;
        for (int c = 0; c < (int)gridDim.x; ++c) t += part[(int64_t)c * s2 + tid];
        G[tid] = t;
    }
}

// Xo = G^-1 H for s x s G (its lower triangle, SPD) and H, through the
// Cholesky factor of G in LDS (one thread factors, thread c solves column
// c).  A pivot not above kSpdRelPivot times its diagonal entry (G singular
// to working precision, e.g. dependent columns), or not finite, sets *flag
// and gives Xo = 0, so the updates that use Xo leave their blocks unchanged.
constexpr double kSpdRelPivot = 1e-13;
__global__ __launch_bounds__(64) void spd_solve_kernel(int s, const double* __restrict__ Gm,
                                                      const double* __restrict__ H,
                                                      double* __restrict__ Xo,
                                                      int* __restrict__ flag) {
    __shared__ double C[BS][BS + 1];
    __shared__ int bad;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < s * s; idx += 64) C[idx % s][idx / s] = Gm[idx];
    __syncthreads();
    if (tid == 0) {
        int b = 0;
        for (int j = 0; j < s && !b; ++j) {
            const double g = C[j][j];
            double d = g;
            for (int k = 0; k < j; ++k) d -= C[j][k] * C[j][k];
            if (!(d > kSpdRelPivot * g) || !(d < INFINITY)) { b = 1; break; }
            d = sqrt(d);
            C[j][j] = d;
            for (int i = j + 1; i < s; ++i) {
                double v = C[i][j];
                for (int k = 0; k < j; ++k) v -= C[i][k] * C[j][k];
                C[i][j] = v / d;
            }
        }
        bad = b;
        if (b) flag[0] = 1;
    }
    __syncthreads();
    if (tid < s) {
        double x[BS];
#pragma unroll
        for (int i = 0; i < BS; ++i) x[i] = 0.0;
        if (!bad) {
#pragma unroll
            for (int i = 0; i < BS; ++i) {                 // C y = h
                if (i < s) {
                    double v = H[i + tid * s];
#pragma unroll
                    for (int k = 0; k < i; ++k) v -= C[i][k] * x[k];
                    x[i] = v / C[i][i];
                }
            }
#pragma unroll
            for (int i = BS - 1; i >= 0; --i) {            // C^T x = y
                if (i < s) {
                    double v = x[i];
#pragma unroll
                    for (int k = i + 1; k < BS; ++k)
                        if (k < s) v -= C[k][i] * x[k];
                    x[i] = v / C[i][i];
                }
            }
        }
#pragma unroll
        for (int i = 0; i < BS; ++i)
            if (i < s) Xo[i + tid * s] = x[i];
    }
}

// Y[r][j] = (Z ? Z[r][j] : Y[r][j]) + a sum_k M[r][k] Cm[k][j], one thread
// per row; the row of M is read before any write, so Y may be M.
__global__ __launch_bounds__(256) void block_axpy_kernel(int64_t n, int s, const double* M,
                                                        int64_t ldm,
                                                        const double* __restrict__ Cm, double a,
                                                        const double* Z, int64_t ldz, double* Y,
                                                        int64_t ldy) {
    __shared__ double cs[BS * BS];
    for (int idx = threadIdx.x; idx < s * s; idx += 256) cs[idx] = Cm[idx];
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    double m[BS];
#pragma unroll
    for (int k = 0; k < BS; ++k) m[k] = k < s ? M[k * ldm + r] : 0.0;
    for (int j = 0; j < s; ++j) {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < BS; ++k)
            if (k < s) v = fma(m[k], cs[k + j * s], v);
        const double base = Z ? Z[j * ldz + r] : Y[j * ldy + r];
        Y[j * ldy + r] = fma(a, v, base);
    }
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + RB - 1) / RB); }

}  // namespace

// per pass of at most SY_JMAX right-hand sides
static int symv_pass_rhs(int nrhs) {
    return nrhs <= 1 ? 1 : nrhs <= 2 ? 2 : nrhs <= 4 ? 4 : SY_JMAX;
}

int64_t mpc_symv_partial_doubles(int64_t n, int nrhs) {
    const int64_t nst = (n + SY_E - 1) / SY_E;
    return nst * nst * SY_E * symv_pass_rhs(nrhs);
}

template <typename T>
static int launch_symv_t(hipStream_t st, const T* A, int64_t lda, int64_t n, int nrhs,
                         const double* X, int64_t ldx, double* Y, int64_t ldy, double* part) {
    const int64_t nst = (n + SY_E - 1) / SY_E;
    const int64_t ntile = nst * (nst + 1) / 2;
    const int NJ = symv_pass_rhs(nrhs);
    for (int j0 = 0; j0 < nrhs; j0 += NJ) {
        const int nj = nrhs - j0 < NJ ? nrhs - j0 : NJ;
        const dim3 g((unsigned)ntile), b(256);
        switch (NJ) {
        case 1: hipLaunchKernelGGL((symv_super_kernel<T, 1>), g, b, 0, st, A, lda, n, (int)nst, j0, nj, X, ldx, part); break;
        case 2: hipLaunchKernelGGL((symv_super_kernel<T, 2>), g, b, 0, st, A, lda, n, (int)nst, j0, nj, X, ldx, part); break;
        case 4: hipLaunchKernelGGL((symv_super_kernel<T, 4>), g, b, 0, st, A, lda, n, (int)nst, j0, nj, X, ldx, part); break;
        default: hipLaunchKernelGGL((symv_super_kernel<T, SY_JMAX>), g, b, 0, st, A, lda, n, (int)nst, j0, nj, X, ldx, part); break;
        }
        hipLaunchKernelGGL(symv_reduce_kernel, dim3((unsigned)((n + 63) / 64), (unsigned)nj), dim3(256),
                           0, st, n, (int)nst, NJ, j0, part, Y, ldy);
        MPC_LAUNCH_CHECK();
    }
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

int64_t mpc_gram_partial_doubles(int64_t n, int s) {
    return ((n + GCH - 1) / GCH) * (int64_t)s * s;
}

int mpc_launch_gram(hipStream_t st, int64_t n, int s, const double* X, int64_t ldx,
                    const double* Y, int64_t ldy, double* part, double* G, int* tick) {
    if (s < 1 || s > BS) {
        mpc_set_error("gram: block of %d columns (1..%d)", s, BS);
        return MPC_ERR_ARG;
    }
    hipLaunchKernelGGL(gram_kernel, dim3((unsigned)((n + GCH - 1) / GCH)), dim3(256), 0, st, n, s,
                       X, ldx, Y, ldy, part, G, tick);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_spd_solve(hipStream_t st, int s, const double* G, const double* H, double* X,
                         int* flag) {
    hipLaunchKernelGGL(spd_solve_kernel, dim3(1), dim3(64), 0, st, s, G, H, X, flag);
    MPC_LAUNCH_CHECK();
    return 0;
}

int mpc_launch_block_axpy(hipStream_t st, int64_t n, int s, const double* M, int64_t ldm,
                          const double* Cm, double a, const double* Z, int64_t ldz, double* Y,
                          int64_t ldy) {
    hipLaunchKernelGGL(block_axpy_kernel, dim3(blocks_for(n)), dim3(256), 0, st, n, s, M, ldm, Cm,
                       a, Z, ldz, Y, ldy);
    MPC_LAUNCH_CHECK();
    return 0;
}
