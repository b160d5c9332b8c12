// cgs_noalloc.hip -- cgs_search_noalloc: the plain search with its device
// scratch in a static array instead of the on-demand workspace, for callers
// that cannot allocate (code under HIP graph capture, tight serving loops).
//
// `scratch` below is a file-scope __device__ array.  In host code its name is
// only a host-side stand-in for the symbol: each device holds its own copy at
// its own address, which hipGetSymbolAddress returns once per device (with
// that device current) on the entry point's first call there.  Kernels and
// copies get that address, never the array's name.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "cgs.h"
#include "cgs_host.hpp"

namespace cgs {

static __device__ __align__(256) unsigned char scratch[CGS_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up on the first call there.
cgs_status device_scratch(void **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return set_error(CGS_ERR_NO_DEVICE, "no HIP device");
    if (dev >= kMaxDevices) return set_error(CGS_ERR_UNSUPPORTED, "search_noalloc: device ordinal beyond 63");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(CGS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return CGS_OK;
}

}  // namespace
}  // namespace cgs

using namespace cgs;

extern "C" cgs_status cgs_search_noalloc(const cgs_grid *grid, const cgs_image *moved, const cgs_image *fixed,
                                         const cgs_search_opts *opts, float *result, cgs_argmin *argmin_out,
                                         void *stream) {
    clear_error();
    void *p = nullptr;
    const cgs_status st = device_scratch(&p);
    if (st != CGS_OK) return st;
    return run_search_in(p, CGS_NOALLOC_SCRATCH_BYTES, grid, moved, fixed, opts, result, argmin_out,
                         (hipStream_t) stream);
}
