// gocean_noalloc.hip -- an entry point that allocates nothing.
//
// gocean_max_dt keeps its partial minima in the caller's buffer (out_dev[2..],
// up to 1024 floats).  gocean_max_dt_static keeps them in a file-scope device
// array instead, so a caller passes only the two output floats: for code under
// HIP graph capture, serving loops and harnesses that cannot allocate.  Each
// GPU the library runs on holds its own copy of the array (4 KiB) until the
// process ends; calls that overlap on one GPU share it and race.  Nothing in
// it needs clearing: each call's first launch writes every partial its second
// launch reads.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/gocean.h"
#include "gocean_internal.hpp"

namespace gocean {

// partial minima of the CFL reduction: one float per reduction workgroup; the
// reduction grid has at most kMaxDtPartials workgroups, so any domain fits
static __device__ __align__(256) float scratch[kMaxDtPartials];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// device address of `scratch` on the current device, looked up on the first
// call on that device (with that device current) and kept
static int scratch_address(float** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error((int)e, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(GOCEAN_EINVAL, "device %d beyond the %d devices this library tracks", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return set_error((int)e, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace gocean

using namespace gocean;

extern "C" int gocean_max_dt_static(const gocean_domain_t* d, const float* eta, const float* hu,
                                    const float* hv, const float* Hm, float courant, float* out_dev,
                                    void* stream) {
    clear_error();
    if (int e = validate_domain(d)) return e;
    if (!eta || !hu || !hv || !Hm || !out_dev) return set_error(GOCEAN_EINVAL, "NULL pointer");
    if (max_dt_partials(d) > kMaxDtPartials)
        return set_error(GOCEAN_EINVAL, "the reduction needs %d partial minima, the static array holds %d",
                         max_dt_partials(d), kMaxDtPartials);
    float* part = nullptr;
    if (int e = scratch_address(&part)) return e;
    return max_dt_into(d, eta, hu, hv, Hm, courant, part, out_dev, (hipStream_t)stream);
}
