// cgs_host.hpp -- host-side helpers shared by the C-ABI translation units:
// argument validation, the per-device workspace, error reporting and the
// host restatement of CudaGrid's bookkeeping.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "cgs.h"
#include "cgs_device.hpp"

namespace cgs {

cgs_status set_error(cgs_status st, const std::string &msg);
void clear_error();

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess)                                                           \
            return ::cgs::set_error(CGS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// Device memory the library owns for one GPU: the zero-padded moving image,
// its bilinear tap table, the (cos, sin) table, the argmin key and small
// pinned host mirrors.  Buffers grow on demand and live until the process
// ends.  Calls on one device serialise on mu.
struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    cgs_status ensure(size_t bytes, bool pinned_host = false);
};

struct Workspace {
    std::mutex mu;
    DevBuf mov, quads, sincos, sincos_host, hist, key, key_host;
    DevBuf prep, prep_host;     // proposal-network input descriptors
    DevBuf cov, cov_host;       // covariance partial sums
    DevBuf fixp, fixq;          // line-traversal costs: zero-padded fixed image, its tap table
    DevBuf alt_tab, alt_tab_host;   // calcMIAlt's row values, fixed-area table and lambda sequence from 0
    DevBuf alt_fb0;                 // calcMIAlt: fixed-image bins of the rows that start at lambda = 0
    int alt_tab_fw = 0, alt_tab_fh = 0;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    // HIP events around one launch on the caller's stream; elapsed_ms()
    // stores the duration for cgs_last_kernel_ms() (after a stream sync).
    hipError_t record_start(hipStream_t s);
    hipError_t record_stop(hipStream_t s);
    hipError_t elapsed_ms();
};

void set_last_kernel_ms(float ms);

cgs_status workspace_for_current_device(Workspace **out);

inline bool is_alt_cost(int cost) { return cost >= CGS_COST_SQD_ALT && cost <= CGS_COST_MI_ALT; }
inline bool is_mi_cost(int cost) {
    return cost == CGS_COST_MI || cost == CGS_COST_MI_PLAIN || cost == CGS_COST_MI_STREAM_ALT || cost == CGS_COST_MI_ALT;
}

// The line-traversal search (cgs_alt.hip).  Called by cgs_search with the
// options validated and ws->mu held.
cgs_status run_alt_search(Workspace *ws, const cgs_grid *grid, const cgs_image *moved, const cgs_image *fixed,
                          const cgs_search_opts *o, float *result, cgs_argmin *out, hipStream_t stream,
                          uint32_t *hist_dump_host);

// Which kernel runs a plain search, and the scratch it stages: the moving
// image's tap table (fast and MI kernels) or zero-padded copy (exact kernel),
// the per-level (cos, sin) table of axis 0, the histogram dump.
enum KernelKind { kKernelExact = 0, kKernelFast = 1, kKernelMI = 2 };
struct SearchSizes {
    KernelKind kind = kKernelExact;
    size_t nmov = 0, ntot = 0;   // moving-image elements, and with the zero pad
    size_t tap_bytes = 0;        // tap table (fast / MI kernels)
    size_t mov_bytes = 0;        // padded copy (exact kernel)
    int levels0 = 0;             // (cos, sin) entries
    size_t hist_bytes = 0;       // one joint histogram
};
SearchSizes search_sizes(const cgs_grid *grid, const cgs_image *moved, const cgs_search_opts *o);

// Axis-0 levels the allocation-free search stages (cos, sin) pairs for.
constexpr int kNoAllocMaxLevels0 = 1024;

// One plain (not line-traversal) search with every device scratch buffer
// carved from `scratch` (scratch_bytes long, 256-byte aligned) and nothing
// allocated; synchronises the stream.  The body of cgs_search_noalloc.
cgs_status run_search_in(void *scratch, size_t scratch_bytes, const cgs_grid *grid, const cgs_image *moved,
                         const cgs_image *fixed, const cgs_search_opts *o, float *result, cgs_argmin *out,
                         hipStream_t stream);

cgs_status validate_image(const cgs_image *im, const char *name);
cgs_status validate_search(const cgs_grid *grid, const cgs_image *moved, const cgs_image *fixed,
                           const cgs_search_opts *o);
void fill_grid_dev(const cgs_grid *grid, GridDev *g);
// Work-index axis order: order[0] is the fastest axis (a permutation of
// 0..dim-1).  Returns false (identity order kept) when the grid is not a full
// mixed-radix box or order is not a permutation.
bool set_work_order(GridDev *g, const int *order);
// The work order the tiled kernels use for a parametrisation (false: grid
// order); CGS_WORK_ORDER ("4,0,5,1,2,3,6,7", or "grid") overrides it.
bool work_order(int dim, int param_mode, int *order);
// Grid points per workgroup of the fast kernel when the caller leaves it to
// the library: 6 when the fastest work axis has a multiple of 3 levels, else
// 8 (CGS_FAST_GPT = 4, 6 or 8 overrides it).
int default_fast_gpt(const GridDev &g);
cgs_argmin decode_key(unsigned long long key, int tie_mode);
uint32_t reference_rank_inverse(uint32_t rank);

}  // namespace cgs
