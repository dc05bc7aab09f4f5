// RMSE on caller device buffers without any allocation: the kernels of
// als_rmse with their per-block partial sums in a fixed device array of the
// library instead of engine memory. For callers that cannot call hipMalloc
// (code under HIP graph capture, tight serving loops).
#include "als.h"
#include "als_launch.hpp"

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

namespace {

// RMSE_BLOCKS per-block partial sums (8 KiB). Each GPU holds its own copy
// for the life of the process. In host code this name is only a stand-in
// for the symbol: the device address comes from hipGetSymbolAddress.
static __device__ __align__(256) double scratch[als::RMSE_BLOCKS];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double *g_ptr[MAX_DEVICES] = {};  // scratch's address on each device, once looked up
thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// The address of this device's copy of `scratch`, looked up once per
// device with that device current (the caller's current device).
int scratch_ptr(double **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return fail(ALS_E_HIP, "no current HIP device");
    if (dev >= MAX_DEVICES) return fail(ALS_E_UNSUPPORTED, "device ordinal beyond the table of scratch addresses");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (st != hipSuccess || !p)
            return fail(ALS_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(st));
        g_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_ptr[dev];
    return ALS_OK;
}

}  // namespace

extern "C" int als_rmse_dev_noalloc(const int *d_row, const int *d_col, const float *d_val, int64_t n,
                                    const float *d_U, const float *d_V, int f, int ld, double *d_out,
                                    void *hip_stream) {
    if (n < 0 || n > (int64_t)INT32_MAX)
        return fail(ALS_E_INVALID, "n must be in 0 .. 2^31 - 1 (the engine's rating-count limit)");
    if (f < 1 || ld < f || ld % 4 != 0 || ld > (1 << 20))
        return fail(ALS_E_INVALID, "need 1 <= f <= ld, ld a multiple of 4 (16-byte row slices), ld <= 2^20");
    if (!d_out || (n > 0 && (!d_row || !d_col || !d_val || !d_U || !d_V)))
        return fail(ALS_E_INVALID, "null device pointer");
    if ((reinterpret_cast<uintptr_t>(d_U) | reinterpret_cast<uintptr_t>(d_V)) % 16 != 0)
        return fail(ALS_E_INVALID, "factor matrices must be 16-byte aligned");
    double *ptr = nullptr;
    int rc = scratch_ptr(&ptr);
    if (rc) return rc;
    als::RmseArgs a{};
    a.row = d_row;
    a.col = d_col;
    a.val = d_val;
    a.n = n;
    a.U = d_U;
    a.V = d_V;
    a.f = f;
    a.ld = ld;
    a.partials = ptr;  // every partial is written by its block: nothing to clear
    const hipError_t st = als::launch_rmse(a, als::RMSE_BLOCKS, d_out, true, static_cast<hipStream_t>(hip_stream));
    if (st != hipSuccess) return fail(ALS_E_HIP, std::string("RMSE launch: ") + hipGetErrorString(st));
    return ALS_OK;
}

extern "C" const char *als_noalloc_last_error(void) { return g_err.c_str(); }
