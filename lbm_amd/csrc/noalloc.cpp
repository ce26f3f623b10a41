// noalloc.cpp -- lbm_mass_noalloc: the deterministic FP64 sum of lbm_mass over
// a caller's device array, for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops).  It launches the same
// two kernels as lbm_mass (block sums of a grid-stride loop, then one block
// adding them in fixed order), so its result is bit-identical to lbm_mass on
// the same values.  It calls no hipMalloc, hipFree or other allocating
// routine and issues no memset or copy: two kernel launches on the caller's
// stream.
//
// The block sums live in `scratch`, a file-scope device array of
// lbm::RED_BLOCKS doubles (8 KiB) that every GPU the library runs on holds
// until the process ends.  Host code never hands the symbol itself to a
// kernel: its address on each device comes from hipGetSymbolAddress, taken
// once per device with that device current (by lbm_create, else by the first
// lbm_mass_noalloc call there), and only that address is used.  All calls on
// one GPU share its array: calls that overlap on one GPU, on any streams,
// race.  Nothing in it needs clearing between calls (each call writes every
// block sum it reads).
#include "../../include/lbm_hip.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "kernels.hpp"

namespace {

constexpr int MAX_DEVICES = 64;
constexpr int64_t MAX_CELLS = (int64_t(1) << 31) - 1;  // the grid limit of lbm_create

static __device__ __align__(256) double scratch[lbm::RED_BLOCKS];

std::mutex g_mu;
double* g_scratch[MAX_DEVICES] = {};

} // namespace

namespace lbm {

// Address of `scratch` on device `dev` (current), resolved once per device.
// Returns nullptr and sets *err on failure.
double* noalloc_scratch(int dev, hipError_t* err)
{
    *err = hipSuccess;
    if (dev < 0 || dev >= MAX_DEVICES) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            *err = e;
            return nullptr;
        }
        g_scratch[dev] = static_cast<double*>(ptr);
    }
    return g_scratch[dev];
}

} // namespace lbm

extern "C" int lbm_mass_noalloc(const double* d_field, int64_t n, double* d_sum, void* stream)
{
    auto fail = [](int code, const std::string& msg) {
        lbm::set_global_error(msg);  // lbm_last_error(NULL)
        return code;
    };
    if (!d_field || !d_sum) return fail(LBM_ERR_ARG, "lbm_mass_noalloc: null device pointer");
    if (n < 1) return fail(LBM_ERR_ARG, "lbm_mass_noalloc: n must be at least 1");
    if (n > MAX_CELLS) return fail(LBM_ERR_ARG, "lbm_mass_noalloc: n exceeds 2^31-1 values");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("HIP failure: ") + hipGetErrorString(e) + " in hipGetDevice");
    double* parts = lbm::noalloc_scratch(dev, &e);
    if (!parts)
        return fail(LBM_ERR_HIP, std::string("HIP failure: ") + hipGetErrorString(e) + " in hipGetSymbolAddress(scratch)");
    // Only g.n is read by the sum kernels; at most RED_BLOCKS block sums.
    const lbm::Grid g{static_cast<int>(n), 1, 1, n};
    lbm::f64::sum_field_into(static_cast<hipStream_t>(stream), g, d_field, parts, d_sum);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("HIP failure: ") + hipGetErrorString(e) + " in launch");
    return LBM_OK;
}
