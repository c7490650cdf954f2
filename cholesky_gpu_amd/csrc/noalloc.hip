// noalloc.hip -- the panel quantiser without any device allocation.
//
// mpc_quantize_panel keeps its per-call scratch (one tile amax per row tile
// and a NaN flag) in the handle's grow-only workspace, which may call
// hipMalloc.  mpc_quantize_panel_noalloc runs the same kernels with that
// scratch in a file-scope device array: at most MPC_NOALLOC_MAX_TILES tiles.
// Each device holds its own copy of the array (until the process ends), whose
// address is looked up once per device when a handle is created on it; host
// code uses that address, never the array's name.  Every call on a device
// shares the array and clears it first; calls that overlap on one device
// (several host threads or streams) race on it.
#include "mpc_internal.hpp"
#include <mutex>

namespace {

constexpr int kScratchWords = MPC_NOALLOC_MAX_TILES + 64;   // tile amax words + flag
static __device__ __align__(256) uint32_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

__global__ __launch_bounds__(256) void clear_words_kernel(uint32_t* __restrict__ w, int n) {
    for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += (int)(gridDim.x * blockDim.x))
        w[i] = 0u;
}

}  // namespace

// Address of this device's scratch array (the caller has made `device`
// current).  Called by mpc_handle_create; later calls return the cached one.
int mpc_noalloc_init_device(int device) {
    if (device < 0 || device >= kMaxDevices) {
        mpc_set_error("noalloc: device %d beyond the %d supported", device, kMaxDevices);
        return MPC_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return 0;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        mpc_set_error("noalloc: hipGetSymbolAddress failed on device %d: %s", device,
                      hipGetErrorString(e));
        return MPC_ERR_HIP;
    }
    g_ptr[device] = p;
    return 0;
}

extern "C" int mpc_quantize_panel_noalloc(mpc_handle* h, const float* dP, int64_t m, int64_t k,
                                          int64_t ld, int tile_rows,
                                          const uint8_t* d_prec_of_tile, uint8_t* d_payload,
                                          int64_t row_stride, uint8_t* d_scales) {
    MPC_CHECK(h && dP && d_prec_of_tile && d_payload && d_scales,
              "quantize_panel_noalloc: NULL argument");
    MPC_CHECK(m >= 1 && k >= 32 && k % 32 == 0 && ld >= m, "quantize_panel_noalloc: bad shape");
    MPC_CHECK(tile_rows >= 32 && tile_rows % 32 == 0,
              "quantize_panel_noalloc: tile_rows multiple of 32");
    MPC_CHECK(row_stride >= 4 * k, "quantize_panel_noalloc: row_stride must be >= 4*k");
    const int64_t nt = (m + tile_rows - 1) / tile_rows;
    MPC_CHECK(nt <= MPC_NOALLOC_MAX_TILES,
              "quantize_panel_noalloc: %lld row tiles exceed the static scratch (%d tiles)",
              (long long)nt, MPC_NOALLOC_MAX_TILES);
    MPC_HIP(hipSetDevice(h->device));
    uint32_t* amax;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        amax = (uint32_t*)(h->device < kMaxDevices ? g_ptr[h->device] : nullptr);
    }
    MPC_CHECK(amax, "quantize_panel_noalloc: scratch of device %d not set up", h->device);
    int* nanf = (int*)(amax + MPC_NOALLOC_MAX_TILES);
    // a fill kernel rather than a memset node: callers may capture the stream
    hipLaunchKernelGGL(clear_words_kernel, dim3(1), dim3(256), 0, h->stream, amax, (int)nt);
    hipLaunchKernelGGL(clear_words_kernel, dim3(1), dim3(64), 0, h->stream,
                       (uint32_t*)nanf, 1);
    MPC_LAUNCH_CHECK();
    int rc;
    if ((rc = mpc_launch_tile_amax(h->stream, dP, m, (int)k, ld, tile_rows, amax))) return rc;
    if ((rc = mpc_launch_quantize_ex(h->stream, dP, m, (int)k, ld, tile_rows, d_prec_of_tile, amax,
                                     nullptr, nullptr, 0, nullptr, 0, 0.0, nullptr, nullptr,
                                     d_payload, row_stride, d_scales, nanf)))
        return rc;
    int hn = 0;
    MPC_HIP(hipMemcpyAsync(&hn, nanf, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    MPC_HIP(hipStreamSynchronize(h->stream));
    if (hn) {
        mpc_set_error("quantize_panel_noalloc: NaN/Inf in the panel");
        return MPC_ERR_NAN;
    }
    return 0;
}
