// api.cpp -- C ABI of libmpchol: handle, generators, factor and PCG drivers.
//
// The drivers are single-stream schedules of the kernels in k_*.hip:
//   factor (switching_chol, switching_chol.cuh:76-224): per block column j
//     POTRF(A_jj) -> TRSM(panel) + row stats -> decide + quantise -> SYRK(A22)
//     with no host synchronisation inside the loop (the reference reads a
//     device flag on the host every step, switching_chol.cuh:124).
//   PCG (precond_CG, mixed_cholesky.cu:145-217): multi-RHS, FP64 vectors,
//     device-resident alpha/beta, one small host read per iteration for the
//     stopping test on the FP64 residual (confirmed on the true residual).
#include "mpc_internal.hpp"
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

// ---------------------------------------------------------------- errors
static thread_local char g_err[1024] = "";

void mpc_set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

extern "C" const char* mpc_last_error_string(void) { return g_err; }
extern "C" const char* mpc_version_string(void) { return "mpchol 0.1 (gfx950)"; }

int mpc_ws_reserve(mpc_workspace& w, size_t bytes) {
    if (bytes <= w.bytes && w.ptr) return 0;
    if (w.ptr) {
        (void)hipFree(w.ptr);
        w.ptr = nullptr;
        w.bytes = 0;
    }
    size_t b = bytes < 256 ? 256 : bytes;
    hipError_t e = hipMalloc(&w.ptr, b);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        w.ptr = nullptr;
        mpc_set_error("device allocation of %zu bytes failed: %s", b, hipGetErrorString(e));
        return MPC_ERR_NOMEM;
    }
    w.bytes = b;
    return 0;
}

// Arrival counters for kernels whose last workgroup finishes a split
// reduction.  Every such launch leaves its counters at zero, so they are
// cleared only when the buffer is (re)allocated, which happens outside any
// stream capture (before the first use at a given size).
int* mpc_tickets(mpc_handle* h, size_t count) {
    const size_t bytes = sizeof(int) * count;
    if (!h->tick.ptr || bytes > h->tick.bytes) {
        if (mpc_ws_reserve(h->tick, bytes)) return nullptr;
        if (hipMemsetAsync(h->tick.ptr, 0, h->tick.bytes, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess) {
            mpc_set_error("ticket buffer: clear failed");
            return nullptr;
        }
    }
    return (int*)h->tick.ptr;
}

// ---------------------------------------------------------------- small init kernels
namespace {
// N events owned by one call: destroyed on every return path
template <int N>
struct owned_events {
    hipEvent_t e[N] = {};
    owned_events() = default;
    owned_events(const owned_events&) = delete;
    owned_events& operator=(const owned_events&) = delete;
    ~owned_events() {
        for (auto& x : e)
            if (x) (void)hipEventDestroy(x);
    }
    int create() {
        for (auto& x : e) MPC_HIP(hipEventCreate(&x));
        return 0;
    }
};

__global__ void init_small_kernel(mpc_small* sm) {
    if (threadIdx.x == 0) {
        sm->info = 0;
        sm->nan_flag = 0;
        sm->mu_bits = 0x7f800000u;
        sm->mu_step[0] = sm->mu_step[1] = 0x7f800000u;
    }
    for (int t = threadIdx.x; t < 1024; t += blockDim.x) {
        for (int b = 0; b < 2; ++b) {
            sm->tile_amax[b][t] = 0u;
            sm->tile_maxrs[b][t] = 0u;
        }
        for (int b = 0; b < MPC_PANEL_SLOTS; ++b) sm->tile_prec[b][t] = 0;
    }
}
__global__ void init_factor_kernel(const float* A, int64_t lda, int64_t n, float* tracker,
                                   float* diag0, uint8_t* map, int64_t nt) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) tracker[i] = diag0[i] = A[i + i * lda];
    if (map && i < nt * nt) {
        int64_t r = i % nt, c = i / nt;
        map[i] = r == c ? (uint8_t)MPC_PREC_FP32 : (uint8_t)255;
    }
}
__global__ void fill_int_kernel(int* p, int n, int v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
}  // namespace

// ---------------------------------------------------------------- handle
extern "C" int mpc_handle_create(mpc_handle** out, int device) {
    if (!out) { mpc_set_error("handle pointer is NULL"); return MPC_ERR_ARG; }
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        mpc_set_error("no HIP device");
        return MPC_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= ndev) { mpc_set_error("device %d out of range", device); return MPC_ERR_ARG; }
    MPC_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    MPC_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        mpc_set_error("device %d is %s; libmpchol is built for gfx950 only", device, prop.gcnArchName);
        return MPC_ERR_UNSUPPORTED;
    }
    mpc_handle* h = new mpc_handle();
    h->device = device;
    h->num_cu = prop.multiProcessorCount;
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete h;
        mpc_set_error("hipStreamCreate failed: %s", hipGetErrorString(e));
        return MPC_ERR_HIP;
    }
    h->own_stream = true;
    e = hipStreamCreateWithFlags(&h->pstream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->wstream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->fstream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_col, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_panel, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_wready, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_wdone, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_wpre, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_bulk, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_far, hipEventDisableTiming);
    int rc = 0;
    if (e != hipSuccess) {
        mpc_set_error("stream/event creation failed: %s", hipGetErrorString(e));
        rc = MPC_ERR_HIP;
    }
    if (!rc) rc = mpc_ws_reserve(h->small, sizeof(mpc_small));
    if (!rc) rc = mpc_noalloc_init_device(device);
    if (rc) { mpc_handle_destroy(h); return rc; }
    *out = h;
    return 0;
}

// stream s waits for a big-block inverse build still running on wstream
static int wait_winv(mpc_handle* h, hipStream_t s) {
    if (h->winv_pending) {
        MPC_HIP(hipStreamWaitEvent(s, h->ev_wdone, 0));
        h->winv_pending = false;
    }
    return 0;
}

static void ws_free(mpc_workspace& w) {
    if (w.ptr) (void)hipFree(w.ptr);
    w.ptr = nullptr;
    w.bytes = 0;
}

extern "C" int mpc_handle_destroy(mpc_handle* h) {
    if (!h) return 0;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->pstream) (void)hipStreamSynchronize(h->pstream);
    if (h->wstream) (void)hipStreamSynchronize(h->wstream);
    if (h->fstream) (void)hipStreamSynchronize(h->fstream);
    for (mpc_workspace* w : {&h->panel, &h->scales, &h->dinv, &h->small, &h->rowstat, &h->amat,
                             &h->pcg, &h->scratch, &h->winv, &h->gp64, &h->gp32, &h->gpvec,
                             &h->tlist, &h->tick})
        ws_free(*w);
    if (h->pcg_exec) (void)hipGraphExecDestroy(h->pcg_exec);
    if (h->ev_col) (void)hipEventDestroy(h->ev_col);
    if (h->ev_panel) (void)hipEventDestroy(h->ev_panel);
    if (h->ev_wready) (void)hipEventDestroy(h->ev_wready);
    if (h->ev_wdone) (void)hipEventDestroy(h->ev_wdone);
    if (h->ev_wpre) (void)hipEventDestroy(h->ev_wpre);
    if (h->ev_bulk) (void)hipEventDestroy(h->ev_bulk);
    if (h->ev_far) (void)hipEventDestroy(h->ev_far);
    if (h->pstream) (void)hipStreamDestroy(h->pstream);
    if (h->wstream) (void)hipStreamDestroy(h->wstream);
    if (h->fstream) (void)hipStreamDestroy(h->fstream);
    if (h->own_stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

extern "C" int mpc_set_stream(mpc_handle* h, void* s) {
    MPC_CHECK(h, "handle is NULL");
    MPC_HIP(hipStreamSynchronize(h->stream));
    if (h->own_stream) (void)hipStreamDestroy(h->stream);
    if (s) {
        h->stream = (hipStream_t)s;
        h->own_stream = false;
    } else {
        MPC_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        h->own_stream = true;
    }
    return 0;
}

extern "C" void* mpc_get_stream(mpc_handle* h) { return h ? (void*)h->stream : nullptr; }

extern "C" int mpc_synchronize(mpc_handle* h) {
    MPC_CHECK(h, "handle is NULL");
    MPC_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

// ---------------------------------------------------------------- host RNG (matches device)
void mpc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1, n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
static inline double host_u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}
double mpc_uniform_host(uint64_t seed, uint64_t stream, uint64_t idx) {
    uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)stream, 0x554e4946u};
    uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)}, o[4];
    mpc_philox4x32(c, k, o);
    return host_u53(o[0], o[1]);
}
double mpc_normal_host(uint64_t seed, uint64_t stream, uint64_t idx) {
    uint64_t q = idx >> 1;
    uint32_t c[4] = {(uint32_t)q, (uint32_t)(q >> 32), (uint32_t)stream, 0x4e4f524du};
    uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)}, o[4];
    mpc_philox4x32(c, k, o);
    double u1 = 1.0 - host_u53(o[0], o[1]), u2 = host_u53(o[2], o[3]);
    double r = std::sqrt(-2.0 * std::log(u1)), th = 6.283185307179586 * u2;
    return (idx & 1) ? r * std::sin(th) : r * std::cos(th);
}

// ---------------------------------------------------------------- matgen
extern "C" void mpc_matgen_opts_default(mpc_matgen_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->kind = MPC_MAT_HOUSEHOLDER_SPD;
    o->dist = MPC_DIST_LOGSPACED;
    o->cond = 1e6;
    o->lam_max = 1e6;
    o->n_reflectors = 4;
    o->seed = 20240531ull;
}

// spectrum: DistType semantics of construct_diag_* (kernels_macros.cuh:635-675)
static void host_spectrum(const mpc_matgen_opts* o, int64_t n, std::vector<double>& lam) {
    lam.resize(n);
    int64_t nc = std::min<int64_t>(o->n_cluster, n);
    for (int64_t i = 0; i < nc; ++i) {
        double t = nc > 1 ? (double)i / (double)(nc - 1) : 0.0;
        lam[i] = std::exp(std::log(o->cluster_lo) + t * (std::log(o->cluster_hi) - std::log(o->cluster_lo)));
    }
    int64_t m = n - nc;
    for (int64_t i = 0; i < m; ++i) {
        double t = m > 1 ? (double)i / (double)(m - 1) : 0.0, v;
        switch (o->dist) {
        case MPC_DIST_RANDOM_LOG_UNIFORM: {
            double u = 1.0 - mpc_uniform_host(o->seed, 7, (uint64_t)i);
            v = std::pow(10.0, std::log10(1.0 / o->cond) * (1.0 - u));
            break; }
        case MPC_DIST_CLUSTERED: v = i < m - 1 ? 1.0 : 1.0 / o->cond; break;
        case MPC_DIST_ARITHMETIC: v = 1.0 - (1.0 - t) * (1.0 - 1.0 / o->cond); break;
        case MPC_DIST_GEOMETRIC: v = std::pow(o->cond, -t); break;
        default: v = std::exp(std::log(1.0 / o->cond) * (1.0 - t)); break;
        }
        lam[nc + i] = v * o->lam_max;
    }
}

extern "C" int mpc_matgen(mpc_handle* h, float* dA, int64_t n, int64_t lda,
                          const mpc_matgen_opts* o, double* d_lambda) {
    MPC_CHECK(h && dA && o, "matgen: NULL argument");
    MPC_CHECK(n >= 1 && lda >= n, "matgen: n=%lld lda=%lld", (long long)n, (long long)lda);
    MPC_CHECK(o->kind == MPC_MAT_HOUSEHOLDER_SPD, "matgen: unknown kind %d", o->kind);
    MPC_CHECK(o->n_reflectors >= 1 && o->n_reflectors <= 4, "matgen: n_reflectors must be 1..4");
    MPC_CHECK(o->dist >= 0 && o->dist <= MPC_DIST_LOGSPACED, "matgen: unknown dist %d", o->dist);
    MPC_CHECK(o->cond >= 1.0 && o->lam_max > 0.0, "matgen: cond >= 1 and lam_max > 0 required");
    MPC_CHECK(o->n_cluster == 0 || (o->cluster_lo > 0 && o->cluster_hi >= o->cluster_lo),
              "matgen: cluster bounds");
    MPC_HIP(hipSetDevice(h->device));
    std::vector<double> lam;
    host_spectrum(o, n, lam);
    const int r = o->n_reflectors;
    std::vector<double> W(4 * n, 0.0), P(4 * n, 0.0), g(4 * n, 0.0);
    for (int k = 0; k < r; ++k) {
        double s = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            double v = mpc_normal_host(o->seed, (uint64_t)k, (uint64_t)i);
            W[4 * i + k] = v;
            s += v * v;
        }
        double inv = 1.0 / std::sqrt(s);
        for (int64_t i = 0; i < n; ++i) W[4 * i + k] *= inv;
    }
    double T[4][4] = {{0}};
    for (int k = 0; k < r; ++k) {
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
    for (int64_t i = 0; i < n; ++i)
        for (int b = 0; b < 4; ++b) {
            double s = 0.0;
            for (int a = 0; a <= b; ++a) s += W[4 * i + a] * T[a][b];
            P[4 * i + b] = s;
        }
    for (int64_t i = 0; i < n; ++i)
        for (int a = 0; a < 4; ++a) {
            double s = 0.0;
            for (int b = 0; b < 4; ++b) s += G[a][b] * P[4 * i + b];
            g[4 * i + a] = s;
        }
    // device copies: lam | W | P | g
    size_t bytes = sizeof(double) * (size_t)n * 13;
    int rc = mpc_ws_reserve(h->scratch, bytes);
    if (rc) return rc;
    double* d = (double*)h->scratch.ptr;
    double *dl = d, *dW = d + n, *dP = d + 5 * n, *dg = d + 9 * n;
    MPC_HIP(hipMemcpyAsync(dl, lam.data(), sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
    MPC_HIP(hipMemcpyAsync(dW, W.data(), sizeof(double) * 4 * n, hipMemcpyHostToDevice, h->stream));
    MPC_HIP(hipMemcpyAsync(dP, P.data(), sizeof(double) * 4 * n, hipMemcpyHostToDevice, h->stream));
    MPC_HIP(hipMemcpyAsync(dg, g.data(), sizeof(double) * 4 * n, hipMemcpyHostToDevice, h->stream));
    rc = mpc_launch_hh_matrix(h->stream, n, dl, dW, dP, dg, dA, lda);
    if (rc) return rc;
    if (d_lambda)
        MPC_HIP(hipMemcpyAsync(d_lambda, dl, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream));
    MPC_HIP(hipStreamSynchronize(h->stream));   // host vectors go out of scope
    return 0;
}

// ---------------------------------------------------------------- RBF
extern "C" int mpc_rbf_points(mpc_handle* h, float* dX, int64_t n, int d, int64_t ldx, int kind,
                              double p_zero, uint64_t seed) {
    MPC_CHECK(h && dX, "rbf_points: NULL argument");
    MPC_CHECK(n >= 1 && d >= 1 && ldx >= d, "rbf_points: bad shape");
    MPC_CHECK(kind == 0 || kind == 1, "rbf_points: kind must be 0 or 1");
    MPC_HIP(hipSetDevice(h->device));
    return mpc_launch_points(h->stream, dX, n, d, ldx, kind, p_zero, seed);
}

extern "C" int mpc_rbf_assemble(mpc_handle* h, const float* dX, int64_t n, int d, int64_t ldx,
                                const double* ls, int n_ls, double variance, double nugget,
                                float* dK, int64_t ldk) {
    MPC_CHECK(h && dX && ls && dK, "rbf_assemble: NULL argument");
    MPC_CHECK(n >= 1 && d >= 1 && ldx >= d && ldk >= n, "rbf_assemble: bad shape");
    MPC_CHECK(n_ls == 1 || n_ls == d, "rbf_assemble: n_ls must be 1 or d");
    for (int i = 0; i < n_ls; ++i) MPC_CHECK(ls[i] > 0, "rbf_assemble: lengthscales must be > 0");
    MPC_HIP(hipSetDevice(h->device));
    std::vector<float> inv(n_ls);
    for (int i = 0; i < n_ls; ++i) inv[i] = (float)(1.0 / ls[i]);
    int rc = mpc_ws_reserve(h->scratch, sizeof(float) * n_ls + 256);
    if (rc) return rc;
    MPC_HIP(hipMemcpyAsync(h->scratch.ptr, inv.data(), sizeof(float) * n_ls, hipMemcpyHostToDevice,
                           h->stream));
    rc = mpc_launch_rbf(h->stream, dX, n, d, ldx, (const float*)h->scratch.ptr, n_ls, variance,
                        nugget, dK, ldk);
    if (rc) return rc;
    MPC_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int mpc_median_pairwise_distance(mpc_handle* h, const float* dX, int64_t n, int d,
                                            int64_t ldx, int64_t nsub, uint64_t seed, double* out) {
    MPC_CHECK(h && dX && out, "median_pairwise_distance: NULL argument");
    MPC_CHECK(nsub >= 2 && nsub <= n, "median_pairwise_distance: nsub must be in [2, n]");
    MPC_HIP(hipSetDevice(h->device));
    // seeded partial Fisher-Yates on the host picks the subsample
    std::vector<int64_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    for (int64_t i = 0; i < nsub; ++i) {
        int64_t j = i + (int64_t)(mpc_uniform_host(seed, 11, (uint64_t)i) * (double)(n - i));
        if (j >= n) j = n - 1;
        std::swap(idx[i], idx[j]);
    }
    std::vector<float> Xs((size_t)nsub * d), row(d);
    MPC_HIP(hipStreamSynchronize(h->stream));
    for (int64_t i = 0; i < nsub; ++i)
        MPC_HIP(hipMemcpy(Xs.data() + i * d, dX + idx[i] * ldx, sizeof(float) * d,
                          hipMemcpyDeviceToHost));
    std::vector<double> dist;
    dist.reserve((size_t)nsub * (nsub - 1) / 2);
    for (int64_t i = 0; i < nsub; ++i)
        for (int64_t j = i + 1; j < nsub; ++j) {
            double s = 0.0;
            for (int c = 0; c < d; ++c) {
                double t = (double)Xs[i * d + c] - (double)Xs[j * d + c];
                s += t * t;
            }
            dist.push_back(std::sqrt(s));
        }
    size_t mid = dist.size() / 2;
    std::nth_element(dist.begin(), dist.begin() + mid, dist.end());
    double med = dist[mid];
    if (dist.size() % 2 == 0) {
        double lo = *std::max_element(dist.begin(), dist.begin() + mid);
        med = 0.5 * (med + lo);
    }
    *out = med;
    return 0;
}

// ---------------------------------------------------------------- factor
extern "C" void mpc_factor_opts_default(mpc_factor_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->nb = 256;
    o->update_prec = MPC_PREC_SWITCH;
    o->ladder[0] = MPC_PREC_E4M3;
    o->ladder[1] = MPC_PREC_FP32;
    o->n_ladder = 2;
    o->eps_prime = 0.25;
    o->fallback = 1;
    o->profile = 0;
    o->agg_panels = 2;
    o->rule = MPC_RULE_MINPIVOT;
}

static int validate_factor_opts(const mpc_factor_opts* o) {
    MPC_CHECK(o->nb == 64 || o->nb == 128 || o->nb == 256, "factor: nb must be 64, 128 or 256");
    MPC_CHECK(o->rule >= MPC_RULE_MINPIVOT && o->rule <= MPC_RULE_SCHUR,
              "factor: rule %d is not an MPC_RULE_*", o->rule);
    MPC_CHECK(o->agg_panels >= 1 && o->agg_panels <= MPC_AGG_MAX, "factor: agg_panels must be 1..%d",
              MPC_AGG_MAX);
    if (o->update_prec == MPC_PREC_SWITCH) {
        MPC_CHECK(o->n_ladder >= 1 && o->n_ladder <= 4, "factor: n_ladder must be 1..4");
        bool f16 = false, other16 = false;
        for (int i = 0; i < o->n_ladder; ++i) {
            MPC_CHECK(mpc_valid_prec(o->ladder[i]), "factor: ladder[%d] = %d invalid", i, o->ladder[i]);
            if (o->ladder[i] == MPC_PREC_FP16) f16 = true;
            if (o->ladder[i] == MPC_PREC_BF16 || mpc_is_fp8(o->ladder[i])) other16 = true;
        }
        MPC_CHECK(!(f16 && other16), "factor: FP16 can only be combined with FP32 in a ladder");
        MPC_CHECK(o->eps_prime > 0, "factor: eps_prime must be > 0");
    } else {
        MPC_CHECK(mpc_valid_prec(o->update_prec), "factor: update_prec %d invalid", o->update_prec);
    }
    return 0;
}

static inline int tile_rows_for(int nb) { return nb < 128 ? 128 : nb; }

void mpc_syrk_masks(int update_prec, const int* ladder, int n_ladder, unsigned* class_mask,
                    unsigned* fp8_mask) {
    int precs[4];
    int np = 0;
    if (update_prec >= 0) precs[np++] = update_prec;
    else
        for (int i = 0; i < n_ladder && i < 4; ++i) precs[np++] = ladder[i];
    auto cls = [](int p, int q) {
        if (p == MPC_PREC_FP32 || q == MPC_PREC_FP32) return MPC_PREC_FP32;
        if (p == MPC_PREC_BF16 || q == MPC_PREC_BF16) return MPC_PREC_BF16;
        if (p == MPC_PREC_FP16 || q == MPC_PREC_FP16) return MPC_PREC_FP16;
        return MPC_PREC_E4M3;
    };
    unsigned cm = 0, fm = 0;
    for (int a = 0; a < np; ++a) {
        for (int b = 0; b < np; ++b) cm |= 1u << cls(precs[a], precs[b]);
        if (precs[a] == MPC_PREC_E4M3 || precs[a] == MPC_PREC_MXE4M3) fm |= 1u;
        if (precs[a] == MPC_PREC_E5M2 || precs[a] == MPC_PREC_MXE5M2) fm |= 2u;
    }
    *class_mask = cm;
    *fp8_mask = fm;
}

// one factorisation attempt; A in place
// W (optional): 2 nI Bo^2 floats receiving the big-block inverses of L (and
// their transposes) for the PCG sweeps, built on wstream as each big diagonal
// block of L becomes final, overlapping the rest of the factor.
// panels whose trailing updates beyond their group are aggregated: the
// option, except in profile mode and unless precision tiles are exactly one
// block (tile_rows == nb)
static int agg_panels_of(const mpc_factor_opts* o) {
    return (o->profile != 0 || tile_rows_for(o->nb) != o->nb) ? 1 : o->agg_panels;
}

// first_src (optional, aggregated schedule only): dA holds only the first
// agg_panels block columns (the caller copied them); the rest of its lower
// triangle is read from first_src (same leading dimension) by the first
// aggregated update, which writes it to dA.
static int factor_once(mpc_handle* h, float* dA, int64_t n, int64_t lda, const mpc_factor_opts* o,
                       double eps, uint8_t* map, float* tracker, mpc_factor_info* fi, float* W,
                       const float* first_src = nullptr) {
    const int nb = o->nb;
    const int64_t nt = (n + nb - 1) / nb;
    const int tr = tile_rows_for(nb);
    hipStream_t st = h->stream;
    int rc;
    if ((rc = wait_winv(h, st))) return rc;       // it reads dinv and the factor
    // panels live in a ring: 2 buffers, or 2 * agg_panels when the updates
    // beyond a group of agg_panels block columns are aggregated (the group
    // being aggregated and the group being factored)
    const bool prof = o->profile != 0;
    const int AG = agg_panels_of(o);
    if (first_src && (AG < 2 || (int64_t)AG * nb >= n)) {
        mpc_set_error("factor: a deferred copy needs the aggregated schedule");
        return MPC_ERR_ARG;
    }
    const int nslot = 2 * AG;
    if ((rc = mpc_ws_reserve(h->panel, (size_t)nslot * (n + 128) * nb * 4))) return rc;
    if ((rc = mpc_ws_reserve(h->scales, (size_t)nslot * (n + 128) * (nb / 32)))) return rc;
    if ((rc = mpc_ws_reserve(h->dinv, (size_t)(nt * nb / 32 + 8) * 1024 * sizeof(float)))) return rc;
    // row sums of squares, then the original diagonal (switching rule SCHUR)
    if ((rc = mpc_ws_reserve(h->rowstat, 2 * (size_t)(n + 128) * sizeof(float)))) return rc;
    // SYRK work lists: one region per stream that issues trailing updates
    const int64_t list_ints = mpc_syrk_list_ints(n, MPC_SYRK_TILE);
    if ((rc = mpc_ws_reserve(h->tlist, sizeof(int) * 3 * (size_t)list_ints))) return rc;
    auto lists_of = [&](hipStream_t s) {
        const int r = s == h->pstream ? 0 : (s == h->fstream ? 2 : 1);
        return (int*)h->tlist.ptr + r * list_ints;
    };
    mpc_small* sm = (mpc_small*)h->small.ptr;
    uint8_t* payload = (uint8_t*)h->panel.ptr;
    uint8_t* scales = (uint8_t*)h->scales.ptr;
    float* dinv = (float*)h->dinv.ptr;
    float* rowss = (float*)h->rowstat.ptr;
    float* diag0 = rowss + (n + 128);
    const int rule = o->rule;
    const int ladder[4] = {o->ladder[0], o->ladder[1], o->ladder[2], o->ladder[3]};
    unsigned cmask = 0, fmask = 0;
    mpc_syrk_masks(o->update_prec, ladder, o->n_ladder, &cmask, &fmask);

    hipLaunchKernelGGL(init_small_kernel, dim3(1), dim3(256), 0, st, sm);
    int64_t ninit = std::max<int64_t>(n, nt * nt);
    // (the diagonal of A: from first_src when dA holds only the first columns)
    hipLaunchKernelGGL(init_factor_kernel, dim3((unsigned)((ninit + 255) / 256)), dim3(256), 0, st,
                       first_src ? first_src : dA, lda, n, tracker, diag0, map, nt);
    MPC_LAUNCH_CHECK();

    const int64_t pstride = (int64_t)(n + 128) * nb * 4;     // payload bytes per buffer
    const int64_t sstride = (int64_t)(n + 128) * (nb / 32);  // scale bytes per buffer
    // POTRF, TRSM and quantisation of panel j on stream `ps`: tile statistics
    // in the step-parity buffers, the quantised panel in ring slot j % nslot
    auto panel = [&](int64_t j, hipStream_t ps) -> int {
        const int64_t c0 = j * nb, bs = std::min<int64_t>(nb, n - c0), e = c0 + bs, s = n - e;
        const int ntl = (int)((s + tr - 1) / tr), b = (int)(j & 1), q = (int)(j % nslot);
        float* Ajj = dA + c0 + c0 * lda;
        int r;
        if ((r = mpc_launch_potrf(ps, Ajj, lda, (int)bs, (int)c0, dinv, sm, ntl,
                                  sm->tile_amax[b], sm->tile_maxrs[b], b,
                                  rule == MPC_RULE_FROBENIUS)))
            return r;
        if (s == 0) return 0;
        float* A21 = dA + e + c0 * lda;
        if ((r = mpc_launch_trsm_stats(ps, A21, lda, s, nb, dinv + (c0 / 32) * 1024, Ajj,
                                       rowss + e, tracker + e, tr, sm, sm->tile_amax[b],
                                       sm->tile_maxrs[b], 0, -1,
                                       rule == MPC_RULE_SCHUR ? diag0 + e : nullptr, b)))
            return r;
        return mpc_launch_quantize_ex(ps, A21, s, nb, lda, tr, nullptr, sm->tile_amax[b],
                                      sm->tile_maxrs[b], sm, o->update_prec, ladder, o->n_ladder,
                                      eps, sm->tile_prec[q], map ? map + (j + 1) + j * nt : nullptr,
                                      payload + q * pstride, (int64_t)nb * 4, scales + q * sstride,
                                      &sm->nan_flag, 0, -1, b, tr / nb, nt - (j + 1), rule,
                                      rowss + e);
    };
    // trailing update of panel j on stream us restricted to SYRK tile columns
    // [t0, t1) of A22 (tiles of MPC_SYRK_TILE, or of 64 * mixed when mixed)
    auto update = [&](int64_t j, int64_t t0, int64_t t1, hipStream_t us, int mixed) -> int {
        const int64_t e = std::min<int64_t>((j + 1) * nb, n), s = n - e;
        const int q = (int)(j % nslot);
        return mpc_launch_syrk_classes(us, dA + e + e * lda, s, lda, nb, payload + q * pstride,
                                       (int64_t)nb * 4, scales + q * sstride, tr,
                                       sm->tile_prec[q], sm, cmask, fmask, t0, t1, mixed,
                                       lists_of(us));
    };
    // aggregated update of panels p0 .. p0+np-1 on the trailing matrix from
    // block column E (rows and columns >= E nb), tile columns [t0, t1) of
    // 64 * qa; qa = 1 one mixed-class launch, qa = 2 one launch per class
    auto agg = [&](int64_t p0, int np, int64_t E, int64_t t0, int64_t t1, hipStream_t us,
                   int qa) -> int {
        const int64_t e = E * nb;
        if (e >= n || np < 1) return 0;
        mpc_syrk_panels P{};
        P.payload = payload;
        P.scales = scales;
        P.prec = &sm->tile_prec[0][0];
        P.pstride = pstride;
        P.sstride = sstride;
        P.rows0 = e - (p0 + 1) * nb;
        P.slots = nslot;
        P.p0 = (int)(p0 % nslot);
        P.np = np;
        P.nb = nb;
        P.csrc = (first_src && p0 == 0) ? first_src + e + e * lda : nullptr;
        return mpc_launch_syrk_panels(us, dA + e + e * lda, n - e, lda, nb, P, (int64_t)nb * 4, tr,
                                      sm, t0, t1, qa, cmask, fmask, lists_of(us));
    };

    const int Bo = mpc_block_bo(n);
    const int64_t nI = (n + Bo - 1) / Bo;
    float* WT = W ? W + (size_t)nI * Bo * Bo : nullptr;
    // after the panel of step s (on stream ps): build the inverse of the big
    // block that step s completes, on wstream
    auto build_w = [&](int64_t s, hipStream_t ps) -> int {
        if (!W || !((((s + 1) * nb) % Bo) == 0 || s == nt - 1)) return 0;
        const int64_t I = (s * nb) / Bo;
        MPC_HIP(hipEventRecord(h->ev_wready, ps));
        MPC_HIP(hipStreamWaitEvent(h->wstream, h->ev_wready, 0));
        if (I == nI - 1) MPC_HIP(hipEventRecord(h->ev_wpre, h->wstream));
        return mpc_build_block_inverses_range(h->wstream, dA, lda, n, Bo, dinv, W, WT, I, I + 1);
    };

    double tp = 0, tt = 0, tq = 0, ts = 0;
    owned_events<2> tev;
    if ((rc = tev.create())) return rc;
    hipEvent_t e0 = tev.e[0], e1 = tev.e[1];
    MPC_HIP(hipEventRecord(e0, st));
    if (!prof && AG > 1) {
        // One-deep look-ahead with aggregated far updates.  Block columns
        // come in groups of AG.  Inside a group, as below: panel j updates
        // block column j+1 on pstream and the rest of the group's columns on
        // st.  The columns beyond the group are updated once per group with
        // all AG panels (K = AG nb): block column E (the next group's first)
        // on pstream, the rest of the next group on st, everything beyond
        // that on fstream, where it overlaps the next group's panels.
        // Ordering: a writer of a block column waits for its previous writer
        // (ev_bulk: st's last update; ev_far: fstream's last); a panel slot
        // is reused AG panels after the group that read it was aggregated.
        hipStream_t ps = h->pstream, fs = h->fstream;
        MPC_HIP(hipEventRecord(h->ev_col, st));
        MPC_HIP(hipStreamWaitEvent(ps, h->ev_col, 0));
        MPC_HIP(hipStreamWaitEvent(fs, h->ev_col, 0));
        MPC_HIP(hipStreamWaitEvent(h->wstream, h->ev_col, 0));
        if ((rc = panel(0, ps))) return rc;
        if ((rc = build_w(0, ps))) return rc;
        bool far_pending = false;
        // ps has waited for st's last recorded update: a wait on a queue
        // costs several microseconds of the chain even when the event is
        // long complete, so ps waits only for st work it has not seen
        bool bulk_seen = true;
        for (int64_t j = 0; j + 1 < nt; ++j) {
            const int64_t g = j / AG, E = std::min<int64_t>((g + 1) * AG, nt);
            const int64_t bs1 = std::min<int64_t>(nb, n - (j + 1) * nb);
            const int64_t split = (bs1 + MPC_SYRK_TILE - 1) / MPC_SYRK_TILE;
            if (!bulk_seen) {                                              // st's last update
                MPC_HIP(hipStreamWaitEvent(ps, h->ev_bulk, 0));
                bulk_seen = true;
            }
            if (j + 1 < E) {
                // inside the group: block column j+1, then the group's other columns
                if ((rc = update(j, 0, 2 * split, ps, 1))) return rc;
                const int64_t wend = (E - (j + 1)) * nb / MPC_SYRK_TILE;
                if (wend > split) {
                    MPC_HIP(hipEventRecord(h->ev_panel, ps));
                    MPC_HIP(hipStreamWaitEvent(st, h->ev_panel, 0));
                    if ((rc = update(j, split, wend, st, 0))) return rc;
                    MPC_HIP(hipEventRecord(h->ev_bulk, st));
                    bulk_seen = false;
                }
            } else {
                // the group's last panel: aggregated update beyond the group
                const int64_t p0 = g * AG;
                const int np = (int)(j - p0 + 1);
                if (far_pending) MPC_HIP(hipStreamWaitEvent(ps, h->ev_far, 0));
                if ((rc = agg(p0, np, E, 0, nb / 64, ps, 1))) return rc;
                MPC_HIP(hipEventRecord(h->ev_panel, ps));
                MPC_HIP(hipStreamWaitEvent(st, h->ev_panel, 0));
                MPC_HIP(hipStreamWaitEvent(fs, h->ev_panel, 0));
                if ((rc = agg(p0, np, E, nb / MPC_SYRK_TILE, (int64_t)AG * nb / MPC_SYRK_TILE, st, 2)))
                    return rc;
                MPC_HIP(hipEventRecord(h->ev_bulk, st));
                bulk_seen = false;
                if ((rc = agg(p0, np, E, (int64_t)AG * nb / MPC_SYRK_TILE, -1, fs, 2))) return rc;
                MPC_HIP(hipEventRecord(h->ev_far, fs));
                far_pending = true;
            }
            if ((rc = panel(j + 1, ps))) return rc;
            if ((rc = build_w(j + 1, ps))) return rc;
        }
        MPC_HIP(hipEventRecord(h->ev_panel, ps));
        MPC_HIP(hipStreamWaitEvent(st, h->ev_panel, 0));
        if (far_pending) MPC_HIP(hipStreamWaitEvent(st, h->ev_far, 0));
    } else if (!prof) {
        // One-deep look-ahead.  pstream carries the critical chain: the update
        // of block column j+1 by panel j (one mixed-class launch of small
        // tiles), then POTRF, TRSM and quantisation of panel j+1.  st
        // carries the rest of panel j's update (block columns beyond j+1),
        // which overlaps the panel.  Ordering: the bulk update of j waits
        // for the block-column update of j (and so for panel j); the
        // block-column update of j waits for the bulk update of j-1 (the
        // last other writer of block column j+1).  The panel buffers of
        // parity j&1 are rewritten by panel j+2, which pstream runs after
        // the block-column update of j+1, itself after the bulk update of j.
        hipStream_t ps = h->pstream;
        MPC_HIP(hipEventRecord(h->ev_col, st));
        MPC_HIP(hipStreamWaitEvent(ps, h->ev_col, 0));
        MPC_HIP(hipStreamWaitEvent(h->wstream, h->ev_col, 0));
        if ((rc = panel(0, ps))) return rc;
        if ((rc = build_w(0, ps))) return rc;
        for (int64_t j = 0; j + 1 < nt; ++j) {
            const int64_t bs1 = std::min<int64_t>(nb, n - (j + 1) * nb);
            const int64_t split = (bs1 + MPC_SYRK_TILE - 1) / MPC_SYRK_TILE;
            if (j > 0) MPC_HIP(hipStreamWaitEvent(ps, h->ev_bulk, 0));    // bulk j-1
            if ((rc = update(j, 0, 2 * split, ps, 1))) return rc;
            // the bulk starts once the block column is done: the column's
            // small launch then has the chip to itself (sharing it with the
            // bulk's workgroups made it several times slower)
            MPC_HIP(hipEventRecord(h->ev_panel, ps));
            MPC_HIP(hipStreamWaitEvent(st, h->ev_panel, 0));
            if ((rc = update(j, split, -1, st, 0))) return rc;
            MPC_HIP(hipEventRecord(h->ev_bulk, st));
            if ((rc = panel(j + 1, ps))) return rc;
            if ((rc = build_w(j + 1, ps))) return rc;
        }
        MPC_HIP(hipEventRecord(h->ev_panel, ps));
        MPC_HIP(hipStreamWaitEvent(st, h->ev_panel, 0));
    } else {
        // sequential with per-phase events
        owned_events<5> pev;
        if ((rc = pev.create())) return rc;
        hipEvent_t* ev = pev.e;
        for (int64_t j = 0; j < nt; ++j) {
            const int64_t c0 = j * nb, bs = std::min<int64_t>(nb, n - c0), e = c0 + bs, s = n - e;
            const int ntl = (int)((s + tr - 1) / tr), b = (int)(j & 1);
            float* Ajj = dA + c0 + c0 * lda;
            float* A21 = dA + e + c0 * lda;
            MPC_HIP(hipEventRecord(ev[0], st));
            if ((rc = mpc_launch_potrf(st, Ajj, lda, (int)bs, (int)c0, dinv, sm, ntl,
                                       sm->tile_amax[b], sm->tile_maxrs[b], b,
                                       rule == MPC_RULE_FROBENIUS)))
                return rc;
            MPC_HIP(hipEventRecord(ev[1], st));
            if (s > 0) {
                if ((rc = mpc_launch_trsm_stats(st, A21, lda, s, nb, dinv + (c0 / 32) * 1024, Ajj,
                                                rowss + e, tracker + e, tr, sm, sm->tile_amax[b],
                                                sm->tile_maxrs[b], 0, -1,
                                                rule == MPC_RULE_SCHUR ? diag0 + e : nullptr, b)))
                    return rc;
                MPC_HIP(hipEventRecord(ev[2], st));
                if ((rc = mpc_launch_quantize_ex(
                         st, A21, s, nb, lda, tr, nullptr, sm->tile_amax[b], sm->tile_maxrs[b], sm,
                         o->update_prec, ladder, o->n_ladder, eps, sm->tile_prec[b],
                         map ? map + (j + 1) + j * nt : nullptr, payload + b * pstride,
                         (int64_t)nb * 4, scales + b * sstride, &sm->nan_flag, 0, -1, b, tr / nb,
                         nt - (j + 1), rule, rowss + e)))
                    return rc;
                MPC_HIP(hipEventRecord(ev[3], st));
                if ((rc = update(j, 0, -1, st, 0))) return rc;
            } else {
                MPC_HIP(hipEventRecord(ev[2], st));
                MPC_HIP(hipEventRecord(ev[3], st));
            }
            MPC_HIP(hipEventRecord(ev[4], st));
            MPC_HIP(hipEventSynchronize(ev[4]));
            auto elapsed = [](hipEvent_t a, hipEvent_t b) {
                float t = 0.f;
                return hipEventElapsedTime(&t, a, b) == hipSuccess ? t : 0.f;
            };
            tp += elapsed(ev[0], ev[1]);
            tt += elapsed(ev[1], ev[2]);
            tq += elapsed(ev[2], ev[3]);
            ts += elapsed(ev[3], ev[4]);
        }
        if (W && (rc = mpc_build_block_inverses_range(st, dA, lda, n, Bo, dinv, W, WT, 0, nI)))
            return rc;
    }
    if (W && !prof) {            // users of W wait for it (wait_winv)
        MPC_HIP(hipEventRecord(h->ev_wdone, h->wstream));
        h->winv_pending = true;
    }
    MPC_HIP(hipEventRecord(e1, st));
    mpc_small hs;
    MPC_HIP(hipMemcpyAsync(&hs, sm, offsetof(mpc_small, tile_amax), hipMemcpyDeviceToHost, st));
    MPC_HIP(hipStreamSynchronize(st));
    float ms_total = 0;
    (void)hipEventElapsedTime(&ms_total, e0, e1);
    fi->ms_total += ms_total;
    fi->ms_potrf += tp;
    fi->ms_trsm += tt;
    fi->ms_quant += tq;
    fi->ms_syrk += ts;
    fi->min_pivot = (double)__builtin_bit_cast(float, hs.mu_bits);
    fi->info = hs.info;
    // a NaN or Inf takes precedence over the failing column it causes (a
    // retry cannot remove it from A)
    if (hs.nan_flag) {
        mpc_set_error("factor: NaN/Inf in a panel or pivot (first failing column %d)", hs.info);
        return MPC_ERR_NAN;
    }
    return hs.info;
}

static void flops_from_map(const uint8_t* map, int64_t n, int nb, mpc_factor_info* fi) {
    const int64_t nt = (n + nb - 1) / nb;
    const int tr = tile_rows_for(nb);
    for (int k = 0; k < 8; ++k) { fi->syrk_flops[k] = 0; fi->tiles[k] = 0; }
    auto cls = [](int p, int q) {
        if (p == MPC_PREC_FP32 || q == MPC_PREC_FP32) return MPC_PREC_FP32;
        if (p == MPC_PREC_BF16 || q == MPC_PREC_BF16) return MPC_PREC_BF16;
        if (p == MPC_PREC_FP16 || q == MPC_PREC_FP16) return MPC_PREC_FP16;
        return MPC_PREC_E4M3;
    };
    // Per step, the flops of a pair of row tiles (a, b) are 2 ra rb k (a > b)
    // or ra (ra + 1) k (a = b), in the class of the pair's formats.  With
    // S_p = sum of ra over the tiles of format p, the off-diagonal sum is
    // sum_{p,q} S_p S_q - sum_a ra^2 in each class: linear in the tiles.  All
    // terms are integers below 2^53, so the sums are exact in any order.
    for (int64_t j = 0; j < nt; ++j) {
        const int64_t e = std::min<int64_t>((j + 1) * nb, n), s = n - e;
        if (s <= 0) break;
        const int64_t ntl = (s + tr - 1) / tr;
        double S[8] = {}, Q[8] = {}, D[8] = {};
        for (int64_t t = 0; t < ntl; ++t) {
            const int p = map[(j + 1 + t * (tr / nb)) + j * nt];
            if (p >= 8) continue;
            fi->tiles[p]++;
            const double r = (double)std::min<int64_t>(tr, s - t * tr);
            S[p] += r;
            Q[p] += r * r;
            D[p] += r * (r + 1);
        }
        const double k = (double)nb;
        for (int p = 0; p < 8; ++p) {
            if (S[p] == 0) continue;
            for (int q = 0; q < 8; ++q)
                if (S[q] != 0) fi->syrk_flops[cls(p, q)] += k * S[p] * S[q];
            fi->syrk_flops[cls(p, p)] += k * (D[p] - Q[p]);
        }
    }
}

static int factor_impl(mpc_handle* h, float* dA, int64_t n, int64_t lda, const mpc_factor_opts* o,
                       uint8_t* d_map, float* d_tracker, mpc_factor_info* info,
                       const float* restore_from, int64_t restore_ld, bool build_prec = false,
                       const float* deferred_src = nullptr) {
    MPC_CHECK(h && dA, "factor: NULL argument");
    MPC_CHECK(n >= 1 && lda >= n, "factor: n=%lld lda=%lld", (long long)n, (long long)lda);
    mpc_factor_opts def;
    if (!o) { mpc_factor_opts_default(&def); o = &def; }
    int rc = validate_factor_opts(o);
    if (rc) return rc;
    MPC_HIP(hipSetDevice(h->device));
    const int nb = o->nb;
    const int64_t nt = (n + nb - 1) / nb;
    MPC_CHECK(((n + tile_rows_for(nb) - 1) / tile_rows_for(nb)) <= 1024,
              "factor: n too large for the tile statistics (max 1024 tiles)");
    mpc_factor_info fi;
    std::memset(&fi, 0, sizeof fi);
    // internal map / tracker when the caller does not want them
    size_t need = (size_t)nt * nt + 256 + (d_tracker ? 0 : (size_t)(n + 64) * sizeof(float));
    uint8_t* map = d_map;
    float* tracker = d_tracker;
    if (!map || !tracker) {
        if ((rc = mpc_ws_reserve(h->scratch, need))) return rc;
        uint8_t* base = (uint8_t*)h->scratch.ptr;
        if (!tracker) tracker = (float*)base;
        if (!map) map = base + (d_tracker ? 0 : (size_t)(n + 64) * sizeof(float));
    }
    // backup for the fallback path when the caller gave no source to restore from
    const float* src = restore_from;
    int64_t src_ld = restore_ld;
    if (o->fallback && !src) {
        if ((rc = mpc_ws_reserve(h->amat, sizeof(float) * (size_t)n * n))) return rc;
        if ((rc = mpc_launch_copy_f32(h->stream, n, n, dA, lda, (float*)h->amat.ptr, n))) return rc;
        src = (const float*)h->amat.ptr;
        src_ld = n;
    }
    // big-block inverses for a following PCG, built during the factor
    float* W = nullptr;
    const int Bo = mpc_block_bo(n);
    if ((rc = wait_winv(h, h->stream))) return rc;
    if (build_prec) {
        const int64_t nI = (n + Bo - 1) / Bo;
        if ((rc = mpc_ws_reserve(h->winv, sizeof(float) * 2 * (size_t)nI * Bo * Bo))) return rc;
        W = (float*)h->winv.ptr;
    }
    h->winv_owner = nullptr;
    double eps = o->eps_prime;
    mpc_factor_opts cur = *o;
    int attempt = 0, rc2_wait = 0;
    for (;;) {
        // a deferred copy of A serves the first attempt only: a retry starts
        // from A restored in full
        rc = factor_once(h, dA, n, lda, &cur, eps, map, tracker, &fi, W,
                         attempt == 0 ? deferred_src : nullptr);
        if (rc <= 0 || !o->fallback) break;
        // non-positive pivot: restore A, then retry with a smaller eps_prime
        // (twice), then with every update in FP32; an FP32 failure is final
        const bool all_fp32 = cur.update_prec == MPC_PREC_FP32;
        if ((rc2_wait = wait_winv(h, h->stream))) return rc2_wait;   // before A is restored
        if (all_fp32) break;
        int rc2 = mpc_launch_copy_f32(h->stream, n, n, src, src_ld, dA, lda);
        if (rc2) return rc2;
        ++attempt;
        if (cur.update_prec == MPC_PREC_SWITCH && attempt <= 2) eps /= 16.0;
        else cur.update_prec = MPC_PREC_FP32;
    }
    fi.retries = attempt;
    fi.eps_prime_used = eps;
    const double nd = (double)n;
    fi.flops_total = nd * nd * nd / 3.0 + nd * nd / 2.0 + nd / 6.0;
    if (rc >= 0) {
        std::vector<uint8_t> hm((size_t)nt * nt);
        MPC_HIP(hipMemcpyAsync(hm.data(), map, hm.size(), hipMemcpyDeviceToHost, h->stream));
        MPC_HIP(hipStreamSynchronize(h->stream));
        flops_from_map(hm.data(), n, nb, &fi);
    }
    if (info) *info = fi;
    // the dinv blocks now describe this factor, and so do the big-block
    // inverses when they were built along with it
    h->dinv_owner = dA;
    h->dinv_n = n;
    h->dinv_ld = lda;
    if (W && rc == 0) {
        h->winv_owner = dA;
        h->winv_n = n;
        h->winv_ld = lda;
        h->winv_bo = Bo;
    }
    return rc;
}

extern "C" int mpc_factor(mpc_handle* h, float* dA, int64_t n, int64_t lda,
                          const mpc_factor_opts* o, uint8_t* d_prec_map, float* d_updated_diag,
                          mpc_factor_info* info) {
    return factor_impl(h, dA, n, lda, o, d_prec_map, d_updated_diag, info, nullptr, 0);
}

// ---------------------------------------------------------------- PCG
extern "C" void mpc_pcg_opts_default(mpc_pcg_opts* o) {
    if (!o) return;
    o->tol = 1e-10;
    o->maxit = 500;
    o->stop_rule = 0;
    o->x0_mode = 0;
    o->graph = 1;
    o->check_every = 0;
}

namespace {
struct pcg_bufs {
    double *R, *Z, *P, *Q, *W, *Y, *X, *part;
    double *rz, *rz_new, *pq, *rr, *bb;
    int* active;
};

// z = L^-T L^-1 r  (all columns), blocked with the big-block inverses
// (with an inverse build in flight: wait for all but the last big block
// now and for the last one just before its stage)
int apply_prec(mpc_handle* h, const float* L, int64_t ldl, int64_t n, int nb, int nrhs,
               const double* R, double* Z, pcg_bufs& b) {
    (void)nb;
    hipStream_t st = h->stream;
    const int64_t nI = (n + h->winv_bo - 1) / h->winv_bo;
    const float* Wv = (const float*)h->winv.ptr;
    hipEvent_t late = nullptr;
    if (h->winv_pending) {
        MPC_HIP(hipStreamWaitEvent(st, h->ev_wpre, 0));
        late = h->ev_wdone;
        h->winv_pending = false;
    }
    return mpc_apply_block_prec(st, L, ldl, n, h->winv_bo, Wv, Wv + (size_t)nI * h->winv_bo * h->winv_bo,
                                R, b.Y, Z, n, nrhs, b.part, late, nI - 1);
}
}  // namespace

// 32x32 diagonal inverses and big-block inverses of L, cached per factor buffer
// keep_pending: a PCG about to apply the preconditioner waits for an
// inverse build still in flight itself (stage by stage); every other user
// waits here.
static int prepare_prec(mpc_handle* h, const float* dL, int64_t ldl, int64_t n,
                        bool keep_pending = false) {
    hipStream_t st = h->stream;
    int rc;
    const bool cached = h->dinv_owner == dL && h->dinv_n == n && h->dinv_ld == ldl &&
                        h->winv_owner == dL && h->winv_n == n && h->winv_ld == ldl &&
                        h->winv_bo == mpc_block_bo(n);
    if (!(cached && keep_pending) && (rc = wait_winv(h, st))) return rc;
    const int64_t nsb = (n + 31) / 32;
    if (h->dinv_owner != dL || h->dinv_n != n || h->dinv_ld != ldl) {
        if ((rc = mpc_ws_reserve(h->dinv, (size_t)(nsb + 8) * 1024 * sizeof(float)))) return rc;
        if ((rc = mpc_launch_dinv_from_L(st, dL, ldl, n, (float*)h->dinv.ptr))) return rc;
        h->dinv_owner = dL;
        h->dinv_n = n;
        h->dinv_ld = ldl;
        h->winv_owner = nullptr;
    }
    const int Bo = mpc_block_bo(n);
    if (h->winv_owner != dL || h->winv_n != n || h->winv_ld != ldl || h->winv_bo != Bo) {
        const int64_t nI = (n + Bo - 1) / Bo;
        const size_t blk = (size_t)nI * Bo * Bo;
        if ((rc = mpc_ws_reserve(h->winv, sizeof(float) * 2 * blk))) return rc;
        if ((rc = mpc_build_block_inverses(st, dL, ldl, n, Bo, (const float*)h->dinv.ptr,
                                           (float*)h->winv.ptr, (float*)h->winv.ptr + blk)))
            return rc;
        h->winv_owner = dL;
        h->winv_n = n;
        h->winv_ld = ldl;
        h->winv_bo = Bo;
    }
    return 0;
}

// the PCG operator: the FP32 matrix, or an FP64 one (exact products)
struct pcg_operator {
    const float* a32 = nullptr;
    const double* a64 = nullptr;
};
static int op_apply(hipStream_t st, const pcg_operator& op, int64_t lda, int64_t n, int nrhs,
                    const double* X, int64_t ldx, double* Y, int64_t ldy, double* part) {
    return op.a64 ? mpc_launch_symv64(st, op.a64, lda, n, nrhs, X, ldx, Y, ldy, part)
                  : mpc_launch_symv(st, op.a32, lda, n, nrhs, X, ldx, Y, ldy, part);
}

static int pcg_impl(mpc_handle* h, const pcg_operator& op, int64_t lda, const float* dL,
                    int64_t ldl, int64_t n, int nb, int nrhs, const double* dB, int64_t ldb,
                    double* dX, int64_t ldx, const mpc_pcg_opts* o, mpc_pcg_info* info) {
    MPC_CHECK(h && (op.a32 || op.a64) && dL && dB && dX, "pcg: NULL argument");
    MPC_CHECK(n >= 1 && lda >= n && ldl >= n && ldb >= n && ldx >= n, "pcg: bad leading dimension");
    MPC_CHECK(nrhs >= 1 && nrhs <= 4096, "pcg: nrhs must be 1..4096");
    MPC_CHECK(nb == 64 || nb == 128 || nb == 256, "pcg: nb must be 64, 128 or 256");
    mpc_pcg_opts def;
    if (!o) { mpc_pcg_opts_default(&def); o = &def; }
    MPC_CHECK(o->tol > 0 && o->maxit >= 0, "pcg: tol > 0 and maxit >= 0 required");
    MPC_CHECK(o->check_every >= 0, "pcg: check_every must be >= 0");
    MPC_HIP(hipSetDevice(h->device));
    hipStream_t st = h->stream;
    int rc;
    if ((rc = prepare_prec(h, dL, ldl, n, true))) return rc;
    const size_t vec = (size_t)n * nrhs;
    const size_t part = std::max<size_t>(
        std::max<size_t>((size_t)mpc_symv_partial_doubles(n, nrhs),
                         (size_t)mpc_block_prec_partial_doubles(n, mpc_block_bo(n), nrhs)),
        2 * (size_t)nrhs * ((n + 2047) / 2048) + 64);      // (two dots at once)
    const int hmax = o->maxit + 1;
    const size_t nbytes = sizeof(double) * (7 * vec + part + 6 * (size_t)nrhs + 2 * (size_t)hmax + 64) +
                          sizeof(int) * 3 * (size_t)nrhs + 1024;
    if ((rc = mpc_ws_reserve(h->pcg, nbytes))) return rc;
    pcg_bufs b;
    double* base = (double*)h->pcg.ptr;
    b.R = base; b.Z = b.R + vec; b.P = b.Z + vec; b.Q = b.P + vec; b.W = b.Q + vec; b.Y = b.W + vec;
    b.X = b.Y + vec;
    b.part = b.X + vec;
    b.rz = b.part + part; b.rz_new = b.rz + nrhs; b.pq = b.rz_new + nrhs; b.rr = b.pq + nrhs;
    b.bb = b.rr + nrhs;
    double* d_relres = b.bb + nrhs;
    double* d_ahist = d_relres + nrhs;
    double* d_bhist = d_ahist + hmax;
    b.active = (int*)(d_bhist + hmax + 8);
    int* d_cand = b.active + nrhs;
    int* d_iters = d_cand + nrhs;

    owned_events<2> tev;
    if ((rc = tev.create())) return rc;
    hipEvent_t e0 = tev.e[0], e1 = tev.e[1];
    MPC_HIP(hipEventRecord(e0, st));
    hipLaunchKernelGGL(fill_int_kernel, dim3((nrhs + 255) / 256), dim3(256), 0, st, b.active, nrhs, 1);
    // B into a packed copy for the norms
    if ((rc = mpc_launch_dot(st, n, nrhs, dB, ldb, dB, ldb, b.bb, b.part))) return rc;
    // The iterate lives in the workspace (b.X, packed, ld n) and is copied to
    // dX at the end, so that the iteration graph below refers to the
    // library's buffers only and is reused by the next solve in them.
    // x0
    const double* cX = b.X;
    const int64_t ldw = n;
    if (o->x0_mode == 0) {
        if ((rc = mpc_launch_copy2d(st, n, nrhs, dB, ldb, b.R, n))) return rc;
        if ((rc = apply_prec(h, dL, ldl, n, nb, nrhs, b.R, b.X, b))) return rc;   // M^-1 B
    } else if (o->x0_mode == 1) {
        if ((rc = mpc_launch_axpby(st, n, nrhs, 0.0, b.X, ldw, 0.0, b.X, ldw))) return rc;
    } else {
        if ((rc = mpc_launch_copy2d(st, n, nrhs, dX, ldx, b.X, ldw))) return rc;
    }
    // R = B - A X
    if ((rc = op_apply(st, op, lda, n, nrhs, cX, ldw, b.Q, n, b.part))) return rc;
    if ((rc = mpc_launch_copy2d(st, n, nrhs, dB, ldb, b.R, n))) return rc;
    if ((rc = mpc_launch_axpby(st, n, nrhs, -1.0, b.Q, n, 1.0, b.R, n))) return rc;
    if ((rc = apply_prec(h, dL, ldl, n, nb, nrhs, b.R, b.Z, b))) return rc;
    if ((rc = mpc_launch_copy2d(st, n, nrhs, b.Z, n, b.P, n))) return rc;
    if ((rc = mpc_launch_dot2(st, n, nrhs, b.R, n, b.Z, n, b.rz, b.R, n, b.rr, b.part))) return rc;

    std::vector<double> hb(nrhs), hrr(nrhs), hrz(nrhs);
    std::vector<int> act(nrhs, 1), iters(nrhs, 0), cand(nrhs, 0);
    std::vector<double> relres(nrhs, 0.0), alphas, betas;
    // the scalars rz, rz_new, pq, rr, bb are contiguous (nrhs each): one copy
    // (each copy is a separate transfer with its own latency)
    std::vector<double> hsc(5 * (size_t)nrhs);
    MPC_HIP(hipMemcpyAsync(hsc.data(), b.rz, sizeof(double) * 5 * nrhs, hipMemcpyDeviceToHost, st));
    MPC_HIP(hipStreamSynchronize(st));
    std::copy(hsc.begin(), hsc.begin() + nrhs, hrz.begin());
    std::copy(hsc.begin() + 3 * nrhs, hsc.begin() + 4 * nrhs, hrr.begin());
    std::copy(hsc.begin() + 4 * nrhs, hsc.begin() + 5 * nrhs, hb.begin());
    std::vector<double> nb2(nrhs);
    int nactive = 0;
    for (int j = 0; j < nrhs; ++j) {
        nb2[j] = std::sqrt(hb[j]);
        if (nb2[j] == 0.0) nb2[j] = 1.0;
        double res = o->stop_rule == 1 ? std::sqrt(std::fabs(hrz[j])) / nb2[j] : std::sqrt(hrr[j]) / nb2[j];
        relres[j] = res;
        if (res <= o->tol) act[j] = 0; else ++nactive;
    }
    // the flags active, cand, iters are contiguous on the device: one copy
    // each way (hflags is the host staging area of all three)
    std::vector<int> hflags(3 * (size_t)nrhs);
    auto flags_to_device = [&](int count) -> int {
        std::copy(act.begin(), act.end(), hflags.begin());
        std::copy(cand.begin(), cand.end(), hflags.begin() + nrhs);
        std::copy(iters.begin(), iters.end(), hflags.begin() + 2 * nrhs);
        MPC_HIP(hipMemcpyAsync(b.active, hflags.data(), sizeof(int) * count * nrhs,
                               hipMemcpyHostToDevice, st));
        return 0;
    };
    if ((rc = flags_to_device(3))) return rc;

    // One iteration for every active column; convergence is tested on the
    // device (candidates freeze), so iterations run back to back and the host
    // only looks every few of them.
    auto body = [&]() -> int {
        int r;
        if ((r = op_apply(st, op, lda, n, nrhs, b.P, n, b.Q, n, b.part))) return r;
        if ((r = mpc_launch_dot(st, n, nrhs, b.P, n, b.Q, n, b.pq, b.part))) return r;
        if ((r = mpc_launch_pcg_update_xr(st, n, nrhs, b.X, ldw, b.R, n, b.P, n, b.Q, n, b.rz, b.pq,
                                          b.active)))
            return r;
        if ((r = apply_prec(h, dL, ldl, n, nb, nrhs, b.R, b.Z, b))) return r;
        if ((r = mpc_launch_dot2(st, n, nrhs, b.R, n, b.Z, n, b.rz_new, b.R, n, b.rr, b.part)))
            return r;
        if ((r = mpc_launch_pcg_update_p(st, n, nrhs, b.P, n, b.Z, n, b.rz_new, b.rz, b.active)))
            return r;
        return mpc_launch_pcg_status(st, nrhs, b.rz, b.rz_new, b.pq, b.rr, b.bb, o->tol,
                                     o->stop_rule, b.active, d_cand, d_iters, d_relres, d_ahist,
                                     d_bhist, hmax);
    };
    // the iteration as a graph: one launch instead of ~60 (no memset nodes).
    // The instantiated graph is kept in the handle with every value the body
    // passes to its kernels; a call with the same values (repeated solves in
    // the same buffers) replays it without capturing again.
    hipGraphExec_t exec = nullptr;
    if (o->graph && nactive > 0 && o->maxit > 0) {
        const std::vector<uint64_t> key = {
            (uint64_t)(uintptr_t)op.a32, (uint64_t)(uintptr_t)op.a64, (uint64_t)lda,
            (uint64_t)n, (uint64_t)nrhs,
            (uint64_t)(uintptr_t)dL, (uint64_t)ldl, (uint64_t)nb, (uint64_t)(uintptr_t)h->pcg.ptr,
            (uint64_t)h->pcg.bytes, (uint64_t)hmax, __builtin_bit_cast(uint64_t, o->tol),
            (uint64_t)o->stop_rule, (uint64_t)(uintptr_t)h->winv.ptr, (uint64_t)h->winv_bo};
        if (!h->pcg_exec || h->pcg_key != key) {
            if (h->pcg_exec) {
                (void)hipGraphExecDestroy(h->pcg_exec);
                h->pcg_exec = nullptr;
                h->pcg_key.clear();
            }
            hipGraph_t g = nullptr;
            MPC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const int brc = body();
            const hipError_t ce = hipStreamEndCapture(st, &g);
            if (brc) { if (g) (void)hipGraphDestroy(g); return brc; }
            if (ce != hipSuccess) {
                mpc_set_error("pcg: graph capture failed: %s", hipGetErrorString(ce));
                return MPC_ERR_HIP;
            }
            const hipError_t ie = hipGraphInstantiate(&h->pcg_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) {
                h->pcg_exec = nullptr;
                mpc_set_error("pcg: graph instantiation failed: %s", hipGetErrorString(ie));
                return MPC_ERR_HIP;
            }
            h->pcg_key = key;
        }
        exec = h->pcg_exec;
    }
    auto run_body = [&]() -> int {
        if (!exec) return body();
        MPC_HIP(hipGraphLaunch(exec, st));
        return 0;
    };

    int done = 0, next = o->check_every > 0 ? o->check_every : 1;
    double prev_res = 0.0;
    int prev_done = 0;
    for (int j = 0; j < nrhs; ++j) prev_res = std::max(prev_res, act[j] ? relres[j] : 0.0);
    while (nactive > 0 && done < o->maxit) {
        const int nbody = std::min(next, o->maxit - done);
        for (int q = 0; q < nbody; ++q)
            if ((rc = run_body())) return rc;
        done += nbody;
        MPC_HIP(hipMemcpyAsync(hflags.data(), b.active, sizeof(int) * 3 * nrhs,
                               hipMemcpyDeviceToHost, st));
        MPC_HIP(hipMemcpyAsync(relres.data(), d_relres, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
        MPC_HIP(hipStreamSynchronize(st));
        std::copy(hflags.begin(), hflags.begin() + nrhs, act.begin());
        std::copy(hflags.begin() + nrhs, hflags.begin() + 2 * nrhs, cand.begin());
        std::copy(hflags.begin() + 2 * nrhs, hflags.begin() + 3 * nrhs, iters.begin());
        bool any_cand = false;
        for (int j = 0; j < nrhs; ++j) any_cand |= cand[j] != 0;
        if (any_cand) {
            if (o->stop_rule == 0) {
                // confirm on the true residual B - A X; replace R where it lags
                if ((rc = op_apply(st, op, lda, n, nrhs, cX, ldw, b.Q, n, b.part))) return rc;
                if ((rc = mpc_launch_copy2d(st, n, nrhs, dB, ldb, b.W, n))) return rc;
                if ((rc = mpc_launch_axpby(st, n, nrhs, -1.0, b.Q, n, 1.0, b.W, n))) return rc;
                if ((rc = mpc_launch_dot(st, n, nrhs, b.W, n, b.W, n, b.rr, b.part))) return rc;
                MPC_HIP(hipMemcpyAsync(hrr.data(), b.rr, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
                MPC_HIP(hipStreamSynchronize(st));
            }
            for (int j = 0; j < nrhs; ++j) {
                if (!cand[j]) continue;
                cand[j] = 0;
                double tr = o->stop_rule == 0 ? std::sqrt(hrr[j]) / nb2[j] : relres[j];
                relres[j] = tr;
                if (tr <= o->tol) {
                    --nactive;                       // converged: stays frozen
                } else {
                    act[j] = 1;                      // residual replacement, continue
                    MPC_HIP(hipMemcpyAsync(b.R + (size_t)j * n, b.W + (size_t)j * n,
                                           sizeof(double) * n, hipMemcpyDeviceToDevice, st));
                }
            }
            if ((rc = flags_to_device(2))) return rc;        // active and cand
        }
        // next look: where the current convergence rate reaches tol (1..8)
        double res = 0.0;
        for (int j = 0; j < nrhs; ++j) if (act[j]) res = std::max(res, relres[j]);
        if (o->check_every > 0) {
            next = o->check_every;
        } else if (res > 0 && prev_res > 0 && res < prev_res) {
            const double rate = std::log(res / prev_res) / (double)(done - prev_done);
            const double need = std::log(o->tol / res) / rate;
            next = (int)std::min(8.0, std::max(1.0, std::floor(need)));
        } else {
            next = 2;
        }
        prev_res = res;
        prev_done = done;
    }
    if ((rc = mpc_launch_copy2d(st, n, nrhs, b.X, ldw, dX, ldx))) return rc;
    // the alpha and beta histories are contiguous (hmax each): one copy
    const int mh = iters[0] > 0 ? std::min(iters[0], hmax) : 0;
    std::vector<double> hab((size_t)hmax + mh);
    if (mh > 0)
        MPC_HIP(hipMemcpyAsync(hab.data(), d_ahist, sizeof(double) * (hmax + mh),
                               hipMemcpyDeviceToHost, st));
    MPC_HIP(hipEventRecord(e1, st));
    MPC_HIP(hipStreamSynchronize(st));
    if (mh > 0) {
        alphas.assign(hab.begin(), hab.begin() + mh);
        betas.assign(hab.begin() + hmax, hab.begin() + hmax + mh);
    }
    h->last_alpha = alphas;
    h->last_beta = betas;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    mpc_pcg_info pi;
    std::memset(&pi, 0, sizeof pi);
    pi.iters = *std::max_element(iters.begin(), iters.end());
    pi.converged = nrhs - nactive;
    pi.relres = *std::max_element(relres.begin(), relres.end());
    pi.ms_total = ms;
    pi.kappa_est = 1.0;
    if (!alphas.empty()) {
        // Lanczos tridiagonal from CG coefficients; extreme eigenvalues by bisection
        int m = (int)alphas.size();
        std::vector<double> dg(m), off2(m > 1 ? m - 1 : 1);
        for (int i = 0; i < m; ++i) {
            dg[i] = 1.0 / alphas[i] + (i > 0 ? betas[i - 1] / alphas[i - 1] : 0.0);
            if (i > 0) off2[i - 1] = betas[i - 1] / (alphas[i - 1] * alphas[i - 1]);
        }
        auto count = [&](double x) {
            int c = 0;
            double q = dg[0] - x;
            if (q < 0) ++c;
            for (int i = 1; i < m; ++i) {
                if (q == 0.0) q = 1e-300;
                q = dg[i] - x - off2[i - 1] / q;
                if (q < 0) ++c;
            }
            return c;
        };
        double lo = 1e300, hi = -1e300;
        for (int i = 0; i < m; ++i) {
            double rad = (i > 0 ? std::sqrt(off2[i - 1]) : 0) + (i < m - 1 ? std::sqrt(off2[i]) : 0);
            lo = std::min(lo, dg[i] - rad);
            hi = std::max(hi, dg[i] + rad);
        }
        double a = lo, bq = hi;
        for (int it = 0; it < 200; ++it) { double mm = 0.5 * (a + bq); (count(mm) >= 1 ? bq : a) = mm; }
        double emin = 0.5 * (a + bq);
        a = lo; bq = hi;
        for (int it = 0; it < 200; ++it) { double mm = 0.5 * (a + bq); (count(mm) >= m ? bq : a) = mm; }
        double emax = 0.5 * (a + bq);
        pi.kappa_est = emin > 0 ? emax / emin : INFINITY;
    }
    if (info) *info = pi;
    return pi.converged == nrhs ? 0 : MPC_ERR_NOT_CONVERGED;
}

extern "C" int mpc_pcg(mpc_handle* h, const float* dA, int64_t lda, const float* dL, int64_t ldl,
                       int64_t n, int nb, int nrhs, const double* dB, int64_t ldb, double* dX,
                       int64_t ldx, const mpc_pcg_opts* o, mpc_pcg_info* info) {
    pcg_operator op;
    op.a32 = dA;
    return pcg_impl(h, op, lda, dL, ldl, n, nb, nrhs, dB, ldb, dX, ldx, o, info);
}

extern "C" int mpc_pcg64(mpc_handle* h, const double* dA, int64_t lda, const float* dL,
                         int64_t ldl, int64_t n, int nb, int nrhs, const double* dB, int64_t ldb,
                         double* dX, int64_t ldx, const mpc_pcg_opts* o, mpc_pcg_info* info) {
    pcg_operator op;
    op.a64 = dA;
    return pcg_impl(h, op, lda, dL, ldl, n, nb, nrhs, dB, ldb, dX, ldx, o, info);
}

// ---------------------------------------------------------------- iterative refinement and block PCG
namespace {
// workspace of the two solvers below: n x nrhs blocks R, Z, P, Q, X, W and
// the scratch of the products, inside the PCG workspace
struct ext_bufs {
    pcg_bufs b;
    double *G, *Gn, *PQ, *coef, *gpart;
    int* flag;
    int* gtick;
};

int ext_setup(mpc_handle* h, int64_t n, int nrhs, ext_bufs& e) {
    const size_t vec = (size_t)n * nrhs;
    const int Bo = mpc_block_bo(n);
    const size_t part = std::max<size_t>(
        std::max<size_t>((size_t)mpc_symv_partial_doubles(n, nrhs),
                         (size_t)mpc_block_prec_partial_doubles(n, Bo, nrhs)),
        std::max<size_t>(2 * (size_t)nrhs * ((n + 2047) / 2048) + 64,
                         (size_t)mpc_gram_partial_doubles(n, nrhs)));
    const size_t s2 = (size_t)nrhs * nrhs;
    const size_t nbytes = sizeof(double) * (6 * vec + part + 4 * s2 + 8 * (size_t)nrhs + 64) + 256;
    int rc;
    if ((rc = mpc_ws_reserve(h->pcg, nbytes))) return rc;
    double* base = (double*)h->pcg.ptr;
    pcg_bufs& b = e.b;
    b = pcg_bufs{};
    b.R = base; b.Z = b.R + vec; b.P = b.Z + vec; b.Q = b.P + vec; b.X = b.Q + vec; b.W = b.X + vec;
    b.Y = b.W;                                   // the sweeps' work block (W is free then)
    b.part = b.W + vec;
    e.gpart = b.part;
    e.G = b.part + part; e.Gn = e.G + s2; e.PQ = e.Gn + s2; e.coef = e.PQ + s2;
    b.rr = e.coef + s2; b.bb = b.rr + nrhs; b.rz = b.bb + nrhs; b.rz_new = b.rz + nrhs;
    b.pq = b.rz_new + nrhs;
    e.flag = (int*)(b.pq + nrhs + 8);
    e.gtick = mpc_tickets(h, 1);
    if (!e.gtick) return MPC_ERR_NOMEM;
    return 0;
}

// host copy of an nrhs-vector of sums of squares -> relative norms
int rel_norms(hipStream_t st, const double* d, int nrhs, const std::vector<double>& nb,
              std::vector<double>& out) {
    std::vector<double> h(nrhs);
    MPC_HIP(hipMemcpyAsync(h.data(), d, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
    MPC_HIP(hipStreamSynchronize(st));
    out.resize(nrhs);
    for (int j = 0; j < nrhs; ++j) out[j] = std::sqrt(h[j]) / nb[j];
    return 0;
}

int ext_check(mpc_handle* h, const float* dA, int64_t lda, const float* dL, int64_t ldl,
              int64_t n, int nb, int nrhs, const double* dB, int64_t ldb, double* dX,
              int64_t ldx, const mpc_pcg_opts* o, const char* who) {
    MPC_CHECK(h && dA && dL && dB && dX, "%s: NULL argument", who);
    MPC_CHECK(n >= 1 && lda >= n && ldl >= n && ldb >= n && ldx >= n, "%s: bad leading dimension",
              who);
    MPC_CHECK(nb == 64 || nb == 128 || nb == 256, "%s: nb must be 64, 128 or 256", who);
    MPC_CHECK(o->tol > 0 && o->maxit >= 0, "%s: tol > 0 and maxit >= 0 required", who);
    MPC_CHECK(o->x0_mode >= 0 && o->x0_mode <= 2, "%s: x0_mode must be 0, 1 or 2", who);
    (void)nrhs;
    return 0;
}

// x0 by mode into b.X (ld n): 0 M^-1 B, 1 zero, 2 from dX
int ext_x0(mpc_handle* h, const float* dL, int64_t ldl, int64_t n, int nb, int nrhs,
           const double* dB, int64_t ldb, const double* dX, int64_t ldx, int mode, pcg_bufs& b) {
    hipStream_t st = h->stream;
    int rc;
    if (mode == 0) {
        if ((rc = mpc_launch_copy2d(st, n, nrhs, dB, ldb, b.R, n))) return rc;
        return apply_prec(h, dL, ldl, n, nb, nrhs, b.R, b.X, b);
    }
    if (mode == 1) return mpc_launch_axpby(st, n, nrhs, 0.0, b.X, n, 0.0, b.X, n);
    return mpc_launch_copy2d(st, n, nrhs, dX, ldx, b.X, n);
}

// R = B - A X and rr = column sums of squares of R
int ext_residual(hipStream_t st, const pcg_operator& op, int64_t lda, int64_t n, int nrhs,
                 const double* dB, int64_t ldb, pcg_bufs& b) {
    int rc;
    if ((rc = op_apply(st, op, lda, n, nrhs, b.X, n, b.Q, n, b.part))) return rc;
    if ((rc = mpc_launch_copy2d(st, n, nrhs, dB, ldb, b.R, n))) return rc;
    if ((rc = mpc_launch_axpby(st, n, nrhs, -1.0, b.Q, n, 1.0, b.R, n))) return rc;
    return mpc_launch_dot(st, n, nrhs, b.R, n, b.R, n, b.rr, b.part);
}

int ext_norms_b(hipStream_t st, int64_t n, int nrhs, const double* dB, int64_t ldb, pcg_bufs& b,
                std::vector<double>& nb) {
    int rc;
    if ((rc = mpc_launch_dot(st, n, nrhs, dB, ldb, dB, ldb, b.bb, b.part))) return rc;
    std::vector<double> one(nrhs, 1.0);
    if ((rc = rel_norms(st, b.bb, nrhs, one, nb))) return rc;
    for (double& v : nb) if (v == 0.0) v = 1.0;
    return 0;
}
}  // namespace

extern "C" int mpc_refine(mpc_handle* h, const float* dA, int64_t lda, const float* dL,
                          int64_t ldl, int64_t n, int nb, int nrhs, const double* dB, int64_t ldb,
                          double* dX, int64_t ldx, const mpc_pcg_opts* o, mpc_pcg_info* info) {
    mpc_pcg_opts def;
    if (!o) { mpc_pcg_opts_default(&def); o = &def; }
    int rc;
    if ((rc = ext_check(h, dA, lda, dL, ldl, n, nb, nrhs, dB, ldb, dX, ldx, o, "refine"))) return rc;
    MPC_CHECK(nrhs >= 1 && nrhs <= 4096, "refine: nrhs must be 1..4096");
    MPC_HIP(hipSetDevice(h->device));
    hipStream_t st = h->stream;
    if ((rc = prepare_prec(h, dL, ldl, n))) return rc;
    ext_bufs e;
    if ((rc = ext_setup(h, n, nrhs, e))) return rc;
    pcg_bufs& b = e.b;
    pcg_operator op;
    op.a32 = dA;
    owned_events<2> tev;
    if ((rc = tev.create())) return rc;
    MPC_HIP(hipEventRecord(tev.e[0], st));
    std::vector<double> nbv, rel;
    if ((rc = ext_norms_b(st, n, nrhs, dB, ldb, b, nbv))) return rc;
    if ((rc = ext_x0(h, dL, ldl, n, nb, nrhs, dB, ldb, dX, ldx, o->x0_mode, b))) return rc;
    int k = 0;
    bool conv = false;
    for (;;) {
        if ((rc = ext_residual(st, op, lda, n, nrhs, dB, ldb, b))) return rc;
        if ((rc = rel_norms(st, b.rr, nrhs, nbv, rel))) return rc;
        const double res = *std::max_element(rel.begin(), rel.end());
        if (res <= o->tol) { conv = true; break; }
        if (k == o->maxit || !(res < INFINITY)) break;
        // x += M^-1 (b - A x)
        if ((rc = apply_prec(h, dL, ldl, n, nb, nrhs, b.R, b.Z, b))) return rc;
        if ((rc = mpc_launch_axpby(st, n, nrhs, 1.0, b.Z, n, 1.0, b.X, n))) return rc;
        ++k;
    }
    if ((rc = mpc_launch_copy2d(st, n, nrhs, b.X, n, dX, ldx))) return rc;
    MPC_HIP(hipEventRecord(tev.e[1], st));
    MPC_HIP(hipStreamSynchronize(st));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, tev.e[0], tev.e[1]);
    mpc_pcg_info pi;
    std::memset(&pi, 0, sizeof pi);
    pi.iters = k;
    pi.relres = *std::max_element(rel.begin(), rel.end());
    for (double r : rel) pi.converged += r <= o->tol;
    pi.kappa_est = 0.0;
    pi.ms_total = ms;
    if (info) *info = pi;
    return conv ? 0 : MPC_ERR_NOT_CONVERGED;
}

extern "C" int mpc_block_pcg(mpc_handle* h, const float* dA, int64_t lda, const float* dL,
                             int64_t ldl, int64_t n, int nb, int nrhs, const double* dB,
                             int64_t ldb, double* dX, int64_t ldx, const mpc_pcg_opts* o,
                             mpc_pcg_info* info) {
    mpc_pcg_opts def;
    if (!o) { mpc_pcg_opts_default(&def); o = &def; }
    int rc;
    if ((rc = ext_check(h, dA, lda, dL, ldl, n, nb, nrhs, dB, ldb, dX, ldx, o, "block_pcg")))
        return rc;
    MPC_CHECK(nrhs >= 1 && nrhs <= MPC_BLOCK_PCG_MAX, "block_pcg: nrhs must be 1..%d",
              MPC_BLOCK_PCG_MAX);
    MPC_HIP(hipSetDevice(h->device));
    hipStream_t st = h->stream;
    if ((rc = prepare_prec(h, dL, ldl, n))) return rc;
    ext_bufs e;
    if ((rc = ext_setup(h, n, nrhs, e))) return rc;
    pcg_bufs& b = e.b;
    const int s = nrhs;
    pcg_operator op;
    op.a32 = dA;
    owned_events<2> tev;
    if ((rc = tev.create())) return rc;
    MPC_HIP(hipEventRecord(tev.e[0], st));
    std::vector<double> nbv, rel;
    if ((rc = ext_norms_b(st, n, s, dB, ldb, b, nbv))) return rc;
    if ((rc = ext_x0(h, dL, ldl, n, nb, s, dB, ldb, dX, ldx, o->x0_mode, b))) return rc;
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, st, e.flag, 1, 0);
    MPC_LAUNCH_CHECK();
    int k = 0, restarts = 0, hflag = 0;
    bool restart = true, conv = false;
    // R = B - A X, Z = M^-1 R, P = Z, G = Z^T R
    auto start = [&]() -> int {
        int r;
        if ((r = ext_residual(st, op, lda, n, s, dB, ldb, b))) return r;
        if ((r = apply_prec(h, dL, ldl, n, nb, s, b.R, b.Z, b))) return r;
        if ((r = mpc_launch_copy2d(st, n, s, b.Z, n, b.P, n))) return r;
        return mpc_launch_gram(st, n, s, b.Z, n, b.R, n, e.gpart, e.G, e.gtick);
    };
    for (;;) {
        if (restart) {
            if ((rc = start())) return rc;
            restart = false;
        } else if ((rc = mpc_launch_dot(st, n, s, b.R, n, b.R, n, b.rr, b.part))) {
            return rc;
        }
        MPC_HIP(hipMemcpyAsync(&hflag, e.flag, sizeof(int), hipMemcpyDeviceToHost, st));
        if ((rc = rel_norms(st, b.rr, s, nbv, rel))) return rc;
        if (hflag) break;                        // the block recurrence broke down
        const double res = *std::max_element(rel.begin(), rel.end());
        if (res <= o->tol) {
            // confirm on the true residual; a lagging one restarts from X
            if ((rc = ext_residual(st, op, lda, n, s, dB, ldb, b))) return rc;
            if ((rc = rel_norms(st, b.rr, s, nbv, rel))) return rc;
            if (*std::max_element(rel.begin(), rel.end()) <= o->tol) { conv = true; break; }
            if (++restarts > 8) break;
            restart = true;
            continue;
        }
        if (k == o->maxit || !(res < INFINITY)) break;
        ++k;
        // Q = A P; alpha = (P^T Q)^-1 G; X += P alpha; R -= Q alpha
        if ((rc = op_apply(st, op, lda, n, s, b.P, n, b.Q, n, b.part))) return rc;
        if ((rc = mpc_launch_gram(st, n, s, b.P, n, b.Q, n, e.gpart, e.PQ, e.gtick))) return rc;
        if ((rc = mpc_launch_spd_solve(st, s, e.PQ, e.G, e.coef, e.flag))) return rc;
        if ((rc = mpc_launch_block_axpy(st, n, s, b.P, n, e.coef, 1.0, nullptr, 0, b.X, n))) return rc;
        if ((rc = mpc_launch_block_axpy(st, n, s, b.Q, n, e.coef, -1.0, nullptr, 0, b.R, n))) return rc;
        // Z = M^-1 R; beta = G^-1 Z^T R; P = Z + P beta
        if ((rc = apply_prec(h, dL, ldl, n, nb, s, b.R, b.Z, b))) return rc;
        if ((rc = mpc_launch_gram(st, n, s, b.Z, n, b.R, n, e.gpart, e.Gn, e.gtick))) return rc;
        if ((rc = mpc_launch_spd_solve(st, s, e.G, e.Gn, e.coef, e.flag))) return rc;
        if ((rc = mpc_launch_block_axpy(st, n, s, b.P, n, e.coef, 1.0, b.Z, n, b.P, n))) return rc;
        std::swap(e.G, e.Gn);
    }
    if ((rc = mpc_launch_copy2d(st, n, s, b.X, n, dX, ldx))) return rc;
    MPC_HIP(hipEventRecord(tev.e[1], st));
    MPC_HIP(hipStreamSynchronize(st));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, tev.e[0], tev.e[1]);
    mpc_pcg_info pi;
    std::memset(&pi, 0, sizeof pi);
    pi.iters = k;
    pi.kappa_est = 0.0;
    if (hflag && o->maxit > k) {
        // breakdown (P^T A P or Z^T R singular in working precision): the
        // columns finish as independent PCGs from the current iterate
        mpc_pcg_opts o2 = *o;
        o2.x0_mode = 2;
        o2.maxit = o->maxit - k;
        mpc_pcg_info p2;
        const int r2 = pcg_impl(h, op, lda, dL, ldl, n, nb, s, dB, ldb, dX, ldx, &o2, &p2);
        if (r2 && r2 != MPC_ERR_NOT_CONVERGED) return r2;
        pi.iters = k + p2.iters;
        pi.converged = p2.converged;
        pi.relres = p2.relres;
        pi.ms_total = ms + p2.ms_total;
        if (info) *info = pi;
        return r2;
    }
    pi.relres = *std::max_element(rel.begin(), rel.end());
    for (double r : rel) pi.converged += r <= o->tol;
    pi.ms_total = ms;
    if (info) *info = pi;
    return conv ? 0 : MPC_ERR_NOT_CONVERGED;
}

extern "C" int mpc_precond_cg(mpc_handle* h, const float* dA, int64_t lda, int64_t n, int nrhs,
                              const double* dB, int64_t ldb, double* dX, int64_t ldx,
                              const mpc_factor_opts* fo, const mpc_pcg_opts* po,
                              mpc_factor_info* finfo, mpc_pcg_info* pinfo) {
    MPC_CHECK(h && dA && dB && dX, "precond_cg: NULL argument");
    MPC_CHECK(n >= 1 && lda >= n, "precond_cg: bad shape");
    MPC_HIP(hipSetDevice(h->device));
    mpc_factor_opts def;
    if (!fo) { mpc_factor_opts_default(&def); fo = &def; }
    int rc = wait_winv(h, h->stream);              // a build may still read amat
    if (rc) return rc;
    if ((rc = mpc_ws_reserve(h->amat, sizeof(float) * (size_t)n * n))) return rc;
    float* L = (float*)h->amat.ptr;
    // The factor works on a copy of A.  With the aggregated schedule (and
    // lda == n) only the first agg_panels block columns are copied here; the
    // first aggregated update reads the rest from A itself and writes it to
    // the copy, so that part of A is read once instead of copied first.
    if ((rc = validate_factor_opts(fo))) return rc;
    const int AG = agg_panels_of(fo);
    const bool defer = lda == n && AG > 1 && (int64_t)AG * fo->nb < n;
    if ((rc = mpc_launch_copy_f32(h->stream, n, defer ? (int64_t)AG * fo->nb : n, dA, lda, L, n)))
        return rc;
    // the caller's A stays intact, so the fallback restores from it
    rc = factor_impl(h, L, n, n, fo, nullptr, nullptr, finfo, dA, lda, true, defer ? dA : nullptr);
    if (rc != 0) return rc;
    pcg_operator op;
    op.a32 = dA;
    return pcg_impl(h, op, lda, L, n, n, fo->nb, nrhs, dB, ldb, dX, ldx, po, pinfo);
}

// ---------------------------------------------------------------- GP loss and gradient
extern "C" int mpc_rbf_assemble64(mpc_handle* h, const float* dX, int64_t n, int d, int64_t ldx,
                                  const double* ls, int n_ls, double variance, double nugget,
                                  double* dK64, int64_t ldk64, float* dK32, int64_t ldk32) {
    MPC_CHECK(h && dX && ls && (dK64 || dK32), "rbf_assemble64: NULL argument");
    MPC_CHECK(n >= 1 && d >= 1 && ldx >= d, "rbf_assemble64: bad shape");
    MPC_CHECK(!dK64 || ldk64 >= n, "rbf_assemble64: ldk64 < n");
    MPC_CHECK(!dK32 || ldk32 >= n, "rbf_assemble64: ldk32 < n");
    MPC_CHECK(n_ls == 1 || n_ls == d, "rbf_assemble64: n_ls must be 1 or d");
    for (int i = 0; i < n_ls; ++i) MPC_CHECK(ls[i] > 0, "rbf_assemble64: lengthscales must be > 0");
    MPC_HIP(hipSetDevice(h->device));
    std::vector<double> inv(n_ls);
    for (int i = 0; i < n_ls; ++i) inv[i] = 1.0 / ls[i];
    int rc = mpc_ws_reserve(h->scratch, sizeof(double) * n_ls + 256);
    if (rc) return rc;
    MPC_HIP(hipMemcpyAsync(h->scratch.ptr, inv.data(), sizeof(double) * n_ls, hipMemcpyHostToDevice,
                           h->stream));
    rc = mpc_launch_rbf64(h->stream, dX, n, d, ldx, (const double*)h->scratch.ptr, n_ls, variance,
                          nugget, dK64, ldk64, dK32, ldk32);
    if (rc) return rc;
    MPC_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int mpc_probes(mpc_handle* h, double* dZ, int64_t ldz, int64_t n, int nprobe,
                          uint64_t seed, int unit) {
    MPC_CHECK(h && dZ, "probes: NULL argument");
    MPC_CHECK(n >= 1 && nprobe >= 1 && ldz >= n, "probes: bad shape");
    MPC_CHECK(nprobe <= 65535, "probes: at most 65535 probes");
    MPC_HIP(hipSetDevice(h->device));
    return mpc_launch_probes(h->stream, dZ, ldz, n, nprobe, seed, unit);
}

extern "C" void mpc_gp_opts_default(mpc_gp_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->n_probes = 32;
    o->probe_kind = 0;
    o->probe_seed = 20240601ull;
    mpc_factor_opts_default(&o->factor);
    mpc_pcg_opts_default(&o->pcg);
}

extern "C" int mpc_gp_loss_grad(mpc_handle* h, const float* dX, int64_t n, int d, int64_t ldx,
                                const double* ls, int n_ls, double variance, double noise,
                                const double* dB, const mpc_gp_opts* o, double* d_grad_b,
                                double* grad_ls, double* grad_variance, double* grad_noise,
                                mpc_gp_info* info) {
    MPC_CHECK(h && dX && ls && dB, "gp_loss_grad: NULL argument");
    MPC_CHECK(n >= 1 && d >= 1 && d <= MPC_GP_MAXD && ldx >= d,
              "gp_loss_grad: need n >= 1, 1 <= d <= %d, ldx >= d", MPC_GP_MAXD);
    MPC_CHECK(n_ls == 1 || n_ls == d, "gp_loss_grad: n_ls must be 1 or d");
    MPC_CHECK(variance > 0 && noise >= 0, "gp_loss_grad: variance > 0 and noise >= 0 required");
    mpc_gp_opts def;
    if (!o) { mpc_gp_opts_default(&def); o = &def; }
    const int P = o->probe_kind == 1 ? (int)n : o->n_probes;
    MPC_CHECK(P >= 0 && (o->probe_kind == 0 || o->probe_kind == 1), "gp_loss_grad: bad probe options");
    MPC_CHECK(P + 1 <= 4096, "gp_loss_grad: at most 4095 probes");
    MPC_HIP(hipSetDevice(h->device));
    hipStream_t st = h->stream;
    const int R = P + 1;
    int rc;
    // workspaces: A64, the FP32 copy (factor source), the factor L, vectors
    if ((rc = mpc_ws_reserve(h->gp64, sizeof(double) * (size_t)n * n))) return rc;
    if ((rc = mpc_ws_reserve(h->gp32, sizeof(float) * (size_t)n * n))) return rc;
    if ((rc = wait_winv(h, h->stream))) return rc;
    if ((rc = mpc_ws_reserve(h->amat, sizeof(float) * (size_t)n * n))) return rc;
    const int nout = (n_ls == 1 ? 1 : d) + 2;
    const size_t vecR = (size_t)n * R;
    // the partial buffer also serves the dot product b^T alpha
    const size_t gpart = std::max<size_t>((size_t)mpc_gp_partial_doubles(n, nout),
                                          (size_t)(n / 2048 + 72));
    if ((rc = mpc_ws_reserve(h->gpvec, sizeof(double) * (4 * vecR + gpart + nout + 8 + n_ls))))
        return rc;
    double* A64 = (double*)h->gp64.ptr;
    float* A32 = (float*)h->gp32.ptr;
    float* L = (float*)h->amat.ptr;
    double* RHS = (double*)h->gpvec.ptr;     // [b, z_1..z_P]
    double* SOL = RHS + vecR;                // [alpha, A^-1 z_1, ...]
    double* U = SOL + vecR;
    double* V = U + vecR;
    double* gpart_buf = V + vecR;
    double* gout = gpart_buf + gpart;        // nout gradient sums, then logdet
    double* dinvls = gout + nout + 8;

    // 1. A in FP64 (operator) and FP32 (factor source)
    std::vector<double> inv(n_ls);
    for (int i = 0; i < n_ls; ++i) {
        MPC_CHECK(ls[i] > 0, "gp_loss_grad: lengthscales must be > 0");
        inv[i] = 1.0 / ls[i];
    }
    MPC_HIP(hipMemcpyAsync(dinvls, inv.data(), sizeof(double) * n_ls, hipMemcpyHostToDevice, st));
    if ((rc = mpc_launch_rbf64(st, dX, n, d, ldx, dinvls, n_ls, variance, noise, A64, n, A32, n)))
        return rc;
    if ((rc = mpc_launch_copy_f32(st, n, n, A32, n, L, n))) return rc;
    // 2. mixed-precision factor of the FP32 copy (the preconditioner)
    mpc_factor_info fi;
    std::memset(&fi, 0, sizeof fi);
    rc = factor_impl(h, L, n, n, &o->factor, nullptr, nullptr, &fi, A32, n, true);
    if (info) info->factor = fi;
    if (rc != 0) return rc;
    // 3. adjoint solve A [alpha, w_1..w_P] = [b, z_1..z_P] with the FP64 operator
    if ((rc = mpc_launch_copy2d(st, n, 1, dB, n, RHS, n))) return rc;
    if (P > 0 && (rc = mpc_launch_probes(st, RHS + n, n, n, P, o->probe_seed, o->probe_kind)))
        return rc;
    if ((rc = mpc_launch_axpby(st, n, R, 0.0, SOL, n, 0.0, SOL, n))) return rc;
    mpc_pcg_opts po = o->pcg;
    po.x0_mode = 0;
    mpc_pcg_info pi;
    std::memset(&pi, 0, sizeof pi);
    pcg_operator op;
    op.a64 = A64;
    rc = pcg_impl(h, op, n, L, n, n, o->factor.nb, R, RHS, n, SOL, n, &po, &pi);
    if (info) info->pcg = pi;
    if (rc != 0) return rc;
    // 4. U = [-alpha/2, w/2 A^-1 z_p], V = [alpha, z_p]
    const double wp = o->probe_kind == 1 ? 1.0 : (P > 0 ? 1.0 / P : 0.0);
    if ((rc = mpc_launch_axpby(st, n, 1, -0.5, SOL, n, 0.0, U, n))) return rc;
    if (P > 0 && (rc = mpc_launch_axpby(st, n, P, 0.5 * wp, SOL + n, n, 0.0, U + n, n))) return rc;
    if ((rc = mpc_launch_copy2d(st, n, 1, SOL, n, V, n))) return rc;
    if (P > 0 && (rc = mpc_launch_copy2d(st, n, P, RHS + n, n, V + n, n))) return rc;
    // 5. fused contraction against dA/dtheta, the logdet of the factor, b^T alpha
    if ((rc = mpc_launch_gp_grad(st, dX, n, d, ldx, dinvls, n_ls, variance, U, V, n,
                                 P > 0 ? R : 1, gpart_buf, gout)))
        return rc;
    if ((rc = mpc_launch_logdiag(st, L, n, n, gout + nout))) return rc;
    if ((rc = mpc_launch_dot(st, n, 1, dB, n, SOL, n, gout + nout + 1, gpart_buf))) return rc;
    if (d_grad_b && (rc = mpc_launch_copy2d(st, n, 1, SOL, n, d_grad_b, n))) return rc;
    std::vector<double> hg(nout + 2);
    MPC_HIP(hipMemcpyAsync(hg.data(), gout, sizeof(double) * (nout + 2), hipMemcpyDeviceToHost, st));
    MPC_HIP(hipStreamSynchronize(st));
    // d loss / d ls_c = sum m K (dx_c / ls_c)^2 / ls_c ; isotropic: sum over c
    if (grad_ls) {
        if (n_ls == 1) grad_ls[0] = hg[0] * inv[0];
        else for (int c = 0; c < d; ++c) grad_ls[c] = hg[c] * inv[c];
    }
    const int q0 = n_ls == 1 ? 1 : d;
    if (grad_variance) *grad_variance = hg[q0] / variance;
    if (grad_noise) *grad_noise = hg[q0 + 1];
    if (info) {
        info->logdet = 2.0 * hg[nout];
        info->datafit = 0.5 * hg[nout + 1];
        info->loss = info->datafit + 0.5 * info->logdet;
    }
    return 0;
}

// ---------------------------------------------------------------- building blocks
extern "C" int mpc_quantize_panel(mpc_handle* h, const float* dP, int64_t m, int64_t k, int64_t ld,
                                  int tile_rows, const uint8_t* d_prec_of_tile, uint8_t* d_payload,
                                  int64_t row_stride, uint8_t* d_scales) {
    MPC_CHECK(h && dP && d_prec_of_tile && d_payload && d_scales, "quantize_panel: NULL argument");
    MPC_CHECK(m >= 1 && k >= 32 && k % 32 == 0 && ld >= m, "quantize_panel: bad shape");
    MPC_CHECK(tile_rows >= 32 && tile_rows % 32 == 0, "quantize_panel: tile_rows multiple of 32");
    MPC_CHECK(row_stride >= 4 * k, "quantize_panel: row_stride must be >= 4*k");
    MPC_HIP(hipSetDevice(h->device));
    const int64_t nt = (m + tile_rows - 1) / tile_rows;
    int rc = mpc_ws_reserve(h->scratch, sizeof(uint32_t) * nt + sizeof(int) * 4 + 256);
    if (rc) return rc;
    uint32_t* amax = (uint32_t*)h->scratch.ptr;
    int* nanf = (int*)(amax + nt);
    MPC_HIP(hipMemsetAsync(h->scratch.ptr, 0, sizeof(uint32_t) * nt + sizeof(int) * 4, h->stream));
    if ((rc = mpc_launch_tile_amax(h->stream, dP, m, (int)k, ld, tile_rows, amax))) return rc;
    if ((rc = mpc_launch_quantize_ex(h->stream, dP, m, (int)k, ld, tile_rows, d_prec_of_tile, amax,
                                     nullptr, nullptr, 0, nullptr, 0, 0.0, nullptr, nullptr, d_payload,
                                     row_stride, d_scales, nanf)))
        return rc;
    int hn = 0;
    MPC_HIP(hipMemcpyAsync(&hn, nanf, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    MPC_HIP(hipStreamSynchronize(h->stream));
    if (hn) { mpc_set_error("quantize_panel: NaN/Inf or FP16 overflow in the panel"); return MPC_ERR_NAN; }
    return 0;
}

extern "C" int mpc_syrk(mpc_handle* h, float* dC, int64_t m, int64_t ldc, int64_t k,
                        const uint8_t* d_payload, int64_t row_stride, const uint8_t* d_scales,
                        int tile_rows, const uint8_t* d_prec_of_tile) {
    MPC_CHECK(h && dC && d_payload && d_scales && d_prec_of_tile, "syrk: NULL argument");
    MPC_CHECK(m >= 1 && ldc >= m && k >= 64 && k % 64 == 0, "syrk: bad shape (k multiple of 64)");
    MPC_CHECK(tile_rows >= 128 && tile_rows % 128 == 0, "syrk: tile_rows multiple of 128");
    MPC_CHECK(row_stride >= 4 * k, "syrk: row_stride must be >= 4*k");
    MPC_HIP(hipSetDevice(h->device));
    // launch only the compute classes (and FP8 format pairs) that occur
    const int64_t nt = (m + tile_rows - 1) / tile_rows;
    std::vector<uint8_t> pt((size_t)nt);
    MPC_HIP(hipMemcpyAsync(pt.data(), d_prec_of_tile, (size_t)nt, hipMemcpyDeviceToHost, h->stream));
    MPC_HIP(hipStreamSynchronize(h->stream));
    auto cls = [](int p, int q) {
        if (p == MPC_PREC_FP32 || q == MPC_PREC_FP32) return MPC_PREC_FP32;
        if (p == MPC_PREC_BF16 || q == MPC_PREC_BF16) return MPC_PREC_BF16;
        if (p == MPC_PREC_FP16 || q == MPC_PREC_FP16) return MPC_PREC_FP16;
        return MPC_PREC_E4M3;
    };
    unsigned cmask = 0, fmask = 0;
    bool present[8] = {};
    for (int64_t t = 0; t < nt; ++t) {
        MPC_CHECK(pt[t] <= MPC_PREC_MXE5M2, "syrk: tile %lld has precision code %d", (long long)t, pt[t]);
        present[pt[t]] = true;
    }
    for (int p = 0; p < 7; ++p)
        for (int q = 0; q < 7; ++q)
            if (present[p] && present[q]) cmask |= 1u << cls(p, q);
    for (int p = MPC_PREC_E5M2; p <= MPC_PREC_MXE5M2; ++p)
        if (present[p]) fmask |= (p == MPC_PREC_E4M3 || p == MPC_PREC_MXE4M3) ? 1u : 2u;
    return mpc_launch_syrk_classes(h->stream, dC, m, ldc, (int)k, d_payload, row_stride, d_scales,
                                   tile_rows, d_prec_of_tile, nullptr, cmask, fmask, 0, -1);
}

extern "C" int mpc_syrk_masked(mpc_handle* h, float* dC, int64_t m, int64_t ldc, int64_t k,
                               const uint8_t* d_payload, int64_t row_stride,
                               const uint8_t* d_scales, int tile_rows,
                               const uint8_t* d_prec_of_tile, unsigned class_mask,
                               unsigned fp8_mask) {
    MPC_CHECK(h && dC && d_payload && d_scales && d_prec_of_tile, "syrk: NULL argument");
    MPC_CHECK(m >= 1 && ldc >= m && k >= 64 && k % 64 == 0, "syrk: bad shape (k multiple of 64)");
    MPC_CHECK(tile_rows >= 128 && tile_rows % 128 == 0, "syrk: tile_rows multiple of 128");
    MPC_CHECK(row_stride >= 4 * k, "syrk: row_stride must be >= 4*k");
    const unsigned classes = (1u << MPC_PREC_FP32) | (1u << MPC_PREC_BF16) |
                             (1u << MPC_PREC_FP16) | (1u << MPC_PREC_E4M3);
    MPC_CHECK((class_mask & ~classes) == 0 && (fp8_mask & ~3u) == 0,
              "syrk: class_mask 0x%x / fp8_mask 0x%x name no compute class", class_mask, fp8_mask);
    MPC_HIP(hipSetDevice(h->device));
    return mpc_launch_syrk_classes(h->stream, dC, m, ldc, (int)k, d_payload, row_stride, d_scales,
                                   tile_rows, d_prec_of_tile, nullptr, class_mask, fp8_mask, 0,
                                   -1);
}

extern "C" int mpc_apply_precond(mpc_handle* h, const float* dL, int64_t ldl, int64_t n,
                                 int nrhs, const double* dR, int64_t ldr, double* dZ,
                                 int64_t ldz) {
    MPC_CHECK(h && dL && dR && dZ, "apply_precond: NULL argument");
    MPC_CHECK(n >= 1 && ldl >= n && ldr >= n && ldz >= n && nrhs >= 1, "apply_precond: bad shape");
    MPC_HIP(hipSetDevice(h->device));
    int rc = prepare_prec(h, dL, ldl, n);
    if (rc) return rc;
    const size_t vec = (size_t)n * nrhs;
    const size_t part = (size_t)mpc_block_prec_partial_doubles(n, h->winv_bo, nrhs);
    if ((rc = mpc_ws_reserve(h->scratch, sizeof(double) * (3 * vec + part)))) return rc;
    double* W = (double*)h->scratch.ptr;
    double* Y = W + vec;
    double* Z = Y + vec;
    double* P = Z + vec;
    hipStream_t st = h->stream;
    if ((rc = mpc_launch_copy2d(st, n, nrhs, dR, ldr, W, n))) return rc;     // packed, ld n
    const int64_t nI = (n + h->winv_bo - 1) / h->winv_bo;
    const float* Wv = (const float*)h->winv.ptr;
    if ((rc = mpc_apply_block_prec(st, dL, ldl, n, h->winv_bo, Wv,
                                   Wv + (size_t)nI * h->winv_bo * h->winv_bo, W, Y, Z, n, nrhs, P)))
        return rc;
    if ((rc = mpc_launch_copy2d(st, n, nrhs, Z, n, dZ, ldz))) return rc;
    MPC_HIP(hipStreamSynchronize(st));
    return 0;
}

extern "C" int mpc_potrf_block(mpc_handle* h, float* dA, int n, int64_t lda, int* info) {
    MPC_CHECK(h && dA, "potrf_block: NULL argument");
    MPC_CHECK(n >= 1 && n <= 256 && lda >= n, "potrf_block: n must be 1..256");
    MPC_HIP(hipSetDevice(h->device));
    int rc = wait_winv(h, h->stream);
    if (rc) return rc;
    rc = mpc_ws_reserve(h->dinv, (size_t)16 * 1024 * sizeof(float));
    if (rc) return rc;
    mpc_small* sm = (mpc_small*)h->small.ptr;
    hipLaunchKernelGGL(init_small_kernel, dim3(1), dim3(256), 0, h->stream, sm);
    if ((rc = mpc_launch_potrf(h->stream, dA, lda, n, 0, (float*)h->dinv.ptr, sm, 0,
                               sm->tile_amax[0], sm->tile_maxrs[0])))
        return rc;
    int hi = 0;
    MPC_HIP(hipMemcpyAsync(&hi, &sm->info, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    MPC_HIP(hipStreamSynchronize(h->stream));
    h->dinv_owner = nullptr;
    h->winv_owner = nullptr;
    if (info) *info = hi;
    return hi;
}

extern "C" int mpc_pcg_last_coeffs(mpc_handle* h, double* alphas, double* betas, int max,
                                   int* count) {
    MPC_CHECK(h && count && max >= 0, "pcg_last_coeffs: bad argument");
    const int m = std::min<int>(max, (int)h->last_alpha.size());
    MPC_CHECK(m == 0 || (alphas && betas), "pcg_last_coeffs: NULL output arrays");
    for (int i = 0; i < m; ++i) {
        alphas[i] = h->last_alpha[i];
        betas[i] = h->last_beta[i];
    }
    *count = m;
    return 0;
}

extern "C" int mpc_convert_f8(mpc_handle* h, const float* dx, uint8_t* dy, int64_t count, int prec) {
    MPC_CHECK(h && dx && dy, "convert_f8: NULL argument");
    MPC_CHECK(prec == MPC_PREC_E4M3 || prec == MPC_PREC_E5M2, "convert_f8: prec must be E4M3 or E5M2");
    MPC_CHECK(count >= 0, "convert_f8: count");
    MPC_HIP(hipSetDevice(h->device));
    if (count == 0) return 0;
    return mpc_launch_convert_f8(h->stream, dx, dy, count, prec);
}
