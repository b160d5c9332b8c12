// cgs_host.cpp -- host-side C-ABI pieces: CudaGrid bookkeeping, argument
// validation, the per-device workspace and error reporting.
#include "cgs_host.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>

namespace cgs {

static thread_local std::string g_last_error;

cgs_status set_error(cgs_status st, const std::string &msg) {
    g_last_error = msg;
    return st;
}

void clear_error() { g_last_error.clear(); }

static bool is_finite(float v) { return std::isfinite(v); }

static thread_local float g_last_kernel_ms = -1.0f;

void set_last_kernel_ms(float ms) { g_last_kernel_ms = ms; }

hipError_t Workspace::record_start(hipStream_t s) {
    if (!ev_start) {
        hipError_t e = hipEventCreate(&ev_start);
        if (e != hipSuccess) return e;
        e = hipEventCreate(&ev_stop);
        if (e != hipSuccess) return e;
    }
    return hipEventRecord(ev_start, s);
}

hipError_t Workspace::record_stop(hipStream_t s) { return hipEventRecord(ev_stop, s); }

hipError_t Workspace::elapsed_ms() {
    float ms = -1.0f;
    hipError_t e = hipEventElapsedTime(&ms, ev_start, ev_stop);
    g_last_kernel_ms = ms;
    return e;
}

cgs_status DevBuf::ensure(size_t bytes, bool pinned_host) {
    if (bytes <= cap && ptr) return CGS_OK;
    if (ptr) {
        if (pinned_host) hipHostFree(ptr);
        else hipFree(ptr);
    }
    ptr = nullptr;
    cap = 0;
    const size_t want = bytes + (bytes >> 2) + 256;
    const hipError_t e = pinned_host ? hipHostMalloc(&ptr, want, hipHostMallocDefault) : hipMalloc(&ptr, want);
    if (e != hipSuccess) {
        ptr = nullptr;
        return set_error(CGS_ERR_HIP, std::string("workspace allocation: ") + hipGetErrorString(e));
    }
    cap = want;
    return CGS_OK;
}

cgs_status workspace_for_current_device(Workspace **out) {
    static std::mutex mu;
    static std::map<int, Workspace *> spaces;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return set_error(CGS_ERR_NO_DEVICE, "no HIP device");
    std::lock_guard<std::mutex> lock(mu);
    Workspace *&w = spaces[dev];
    if (!w) w = new Workspace();   // lives until the process ends
    *out = w;
    return CGS_OK;
}

cgs_status validate_image(const cgs_image *im, const char *name) {
    if (!im) return set_error(CGS_ERR_INVALID_ARGUMENT, std::string(name) + ": null image");
    if (!im->data) return set_error(CGS_ERR_INVALID_ARGUMENT, std::string(name) + ": null data");
    if (im->width <= 0 || im->height <= 0 || im->channels <= 0 || im->channels > 4)
        return set_error(CGS_ERR_INVALID_ARGUMENT, std::string(name) + ": bad dimensions");
    if (im->pixtype != CGS_PIX_U8 && im->pixtype != CGS_PIX_F32)
        return set_error(CGS_ERR_INVALID_ARGUMENT, std::string(name) + ": bad pixel type");
    // the truncated float linear index and the bilinear taps stay exact
    // only below 2^24 elements
    if ((int64_t) im->width * im->height * im->channels >= (1 << 24) - 4 * (im->width + 1))
        return set_error(CGS_ERR_TOO_LARGE, std::string(name) + ": image too large (>= 2^24 elements)");
    return CGS_OK;
}

cgs_status validate_search(const cgs_grid *grid, const cgs_image *moved, const cgs_image *fixed,
                           const cgs_search_opts *o) {
    if (!grid || !o) return set_error(CGS_ERR_INVALID_ARGUMENT, "null grid/opts");
    cgs_status st;
    if ((st = validate_image(moved, "moved")) != CGS_OK) return st;
    if ((st = validate_image(fixed, "fixed")) != CGS_OK) return st;
    if (moved->channels != fixed->channels)
        return set_error(CGS_ERR_INVALID_ARGUMENT, "moved/fixed channel counts differ");
    if (o->cost < CGS_COST_SAD || o->cost > CGS_COST_MI_ALT) return set_error(CGS_ERR_INVALID_ARGUMENT, "bad cost");
    if (o->sampler != CGS_SAMPLER_NEAREST && o->sampler != CGS_SAMPLER_BILINEAR)
        return set_error(CGS_ERR_INVALID_ARGUMENT, "bad sampler");
    if (o->param_mode < CGS_PARAM_CENTRED || o->param_mode > CGS_PARAM_NORM)
        return set_error(CGS_ERR_INVALID_ARGUMENT, "bad param mode");
    if (o->tie_mode != CGS_TIE_LOWEST && o->tie_mode != CGS_TIE_REFERENCE)
        return set_error(CGS_ERR_INVALID_ARGUMENT, "bad tie mode");
    if (is_mi_cost(o->cost) && (o->mi_bins < 8 || o->mi_bins > 64))
        return set_error(CGS_ERR_INVALID_ARGUMENT, "mi_bins must be in [8, 64]");
    if (is_alt_cost(o->cost)) {
        // the line-traversal functors sample through valueAt (bilinear) and
        // fix their own parametrisation
        if (o->sampler != CGS_SAMPLER_BILINEAR)
            return set_error(CGS_ERR_UNSUPPORTED, "line-traversal costs sample bilinearly only");
        if (o->param_mode != (o->cost == CGS_COST_MI_ALT ? CGS_PARAM_NORM : CGS_PARAM_CENTRED))
            return set_error(CGS_ERR_INVALID_ARGUMENT,
                             "line-traversal costs: calcMIAlt takes CGS_PARAM_NORM, the others CGS_PARAM_CENTRED");
    } else if (o->param_mode == CGS_PARAM_NORM) {
        return set_error(CGS_ERR_INVALID_ARGUMENT, "CGS_PARAM_NORM belongs to calcMIAlt only");
    }
    if (o->tile_gpt != 0 && o->tile_gpt != 2 && o->tile_gpt != 4 && o->tile_gpt != 6 && o->tile_gpt != 8)
        return set_error(CGS_ERR_INVALID_ARGUMENT, "tile_gpt must be 0, 2, 4, 6 or 8");
    const int need = o->param_mode == CGS_PARAM_TRANSLATION ? 2 : 8;
    if (grid->dim != need)
        return set_error(CGS_ERR_INVALID_ARGUMENT, "grid dimension does not match the parameter mode");
    if (grid->num_elements <= 0) return set_error(CGS_ERR_INVALID_ARGUMENT, "empty grid");
    for (int a = 0; a < grid->dim; ++a)   // sample indices stay exact in binary32
        if (grid->offset[a] < 0 || (int64_t) grid->offset[a] + (int64_t) grid->count[a] > (1 << 24))
            return set_error(CGS_ERR_INVALID_ARGUMENT, "grid offset outside 0..2^24");
    return CGS_OK;
}

void fill_grid_dev(const cgs_grid *grid, GridDev *g) {
    std::memset(g, 0, sizeof(*g));
    g->dim = grid->dim;
    g->n = (int32_t) grid->num_elements;
    for (int a = 0; a < kMaxDim; ++a) {
        g->start[a] = a < grid->dim ? grid->start[a] : 0.0f;
        g->res[a] = a < grid->dim ? grid->resolution[a] : 0.0f;
        g->incr[a] = a < grid->dim ? grid->increments[a] : 1;
        g->off[a] = a < grid->dim ? grid->offset[a] : 0;
        g->wincr[a] = g->incr[a];
    }
    // samples per axis as the mixed-radix decode sees them (the last axis
    // takes what is left of n); identity work order
    for (int a = 0; a < kMaxDim; ++a) {
        if (a + 1 < g->dim) g->count[a] = g->incr[a + 1] / g->incr[a];
        else if (a + 1 == g->dim) g->count[a] = (g->n + g->incr[a] - 1) / g->incr[a];
        else g->count[a] = 1;
    }
}

bool work_order(int dim, int param_mode, int *order) {
    const char *e = std::getenv("CGS_WORK_ORDER");
    if (e) {   // "grid" (or anything that is not dim digits) keeps the grid order
        int k = 0;
        for (const char *c = e; *c && k < dim; ++c)
            if (*c >= '0' && *c <= '9') order[k++] = *c - '0';
        return k == dim;
    }
    // the centred and non-centred parametrisations: x-translation (axis 4)
    // fastest, then rotation, then y-translation; workgroups then hold points
    // whose warps differ by horizontal shifts and share tap cache lines
    if (dim != kMaxDim || (param_mode != CGS_PARAM_CENTRED && param_mode != CGS_PARAM_NONCENTRED)) return false;
    static const int kOrder[kMaxDim] = {4, 0, 5, 1, 2, 3, 6, 7};
    for (int k = 0; k < dim; ++k) order[k] = kOrder[k];
    return true;
}

int default_fast_gpt(const GridDev &g) {
    const char *e = std::getenv("CGS_FAST_GPT");
    if (e) {
        const int v = std::atoi(e);
        if (v == 4 || v == 6 || v == 8) return v;
    }
    if (!g.permuted) return 8;
    int fastest = 0;   // the axis with work stride 1
    for (int a = 0; a < g.dim; ++a)
        if (g.wincr[a] == 1 && g.count[a] > 1) fastest = a;
    return g.count[fastest] % 3 == 0 ? 6 : 8;
}

bool set_work_order(GridDev *g, const int *order) {
    int64_t prod = 1;
    for (int a = 0; a < g->dim; ++a) prod *= g->count[a];
    if (prod != g->n) return false;   // not a full mixed-radix box: keep the identity order
    int seen = 0;
    for (int k = 0; k < g->dim; ++k) {
        if (order[k] < 0 || order[k] >= g->dim || (seen >> order[k]) & 1) return false;
        seen |= 1 << order[k];
    }
    int32_t stride = 1;
    for (int k = 0; k < g->dim; ++k) {
        g->wincr[order[k]] = stride;
        stride *= g->count[order[k]];
    }
    g->permuted = 1;
    return true;
}

static uint32_t bitrev9(uint32_t t) {
    uint32_t r = 0;
    for (int b = 0; b < 9; ++b)
        if (t & (1u << b)) r |= 1u << (8 - b);
    return r;
}

uint32_t reference_rank_inverse(uint32_t rank) {
    return (rank & ~1023u) | bitrev9((rank >> 1) & 511u) | ((rank & 1u) << 9);
}

cgs_argmin decode_key(unsigned long long key, int tie_mode) {
    cgs_argmin r;
    if (key == ~0ull) {
        r.value = INFINITY;
        r.index = -1;
        return r;
    }
    const uint32_t hi = (uint32_t) (key >> 32), lo = (uint32_t) key;
    const uint32_t bits = (hi & 0x80000000u) ? (hi & 0x7FFFFFFFu) : ~hi;
    std::memcpy(&r.value, &bits, sizeof(float));
    r.index = (int32_t) (tie_mode == CGS_TIE_REFERENCE ? reference_rank_inverse(lo) : lo);
    return r;
}

}  // namespace cgs

using namespace cgs;

extern "C" {

const char *cgs_last_error(void) { return g_last_error.c_str(); }

float cgs_last_kernel_ms(void) { return g_last_kernel_ms; }

// CudaGrid::setStartPoint/setEndPoint/setNumSamples -> updateProcess
// (src/gpu/cudaGridSearch.cuh:198-215, 256-270): resolution (end-start) /
// (count-1) in binary32; an axis whose span is below 1e-7*count collapses to
// one sample; a one-sample axis gets resolution 0 (the reference reads
// end[-1]-start[-1] there).  numElements (:89-98): int32 running product,
// each step (float)total * count truncated.
cgs_status cgs_grid_init(cgs_grid *g, int32_t dim, const float *start, const float *end, const float *count) {
    clear_error();
    if (!g || !start || !end || !count) return set_error(CGS_ERR_INVALID_ARGUMENT, "grid_init: null pointer");
    if (dim < 1 || dim > CGS_MAX_DIM) return set_error(CGS_ERR_INVALID_ARGUMENT, "grid_init: dim must be 1..8");
    std::memset(g, 0, sizeof(*g));
    g->dim = dim;
    for (int a = 0; a < dim; ++a) {
        if (!is_finite(start[a]) || !is_finite(end[a]) || !is_finite(count[a]) || count[a] < 1.0f)
            return set_error(CGS_ERR_INVALID_ARGUMENT, "grid_init: non-finite bound or count < 1");
        g->start[a] = start[a];
        g->end[a] = end[a];
        float c = count[a];
        const float diff = end[a] - start[a];
        if ((double) std::fabs(diff) < 1.0e-7 * (double) c) c = 1.0f;
        g->count[a] = c;
        g->resolution[a] = (c == 1.0f) ? 0.0f : diff / (c - 1.0f);
    }
    int32_t total = 1;
    for (int a = 0; a < dim; ++a) {
        const float p = (float) total * g->count[a];
        if (!(p < 2147483648.0f)) return set_error(CGS_ERR_GRID_TOO_LARGE, "grid has 2^31 or more points");
        total = (int32_t) p;
    }
    g->num_elements = total;
    g->increments[0] = 1;
    for (int a = 1; a < dim; ++a) g->increments[a] = (int32_t) (g->count[a - 1] * (float) g->increments[a - 1]);
    return CGS_OK;
}

cgs_status cgs_grid_point(const cgs_grid *g, int64_t index, float *out) {
    clear_error();
    if (!g || !out) return set_error(CGS_ERR_INVALID_ARGUMENT, "grid_point: null pointer");
    if (index < 0 || index >= g->num_elements) return set_error(CGS_ERR_INVALID_ARGUMENT, "grid_point: index out of range");
    int32_t rem = (int32_t) index;
    int32_t idx[CGS_MAX_DIM];
    for (int a = g->dim - 1; a >= 0; --a) {
        idx[a] = rem / g->increments[a];
        rem -= idx[a] * g->increments[a];
    }
    for (int a = 0; a < g->dim; ++a)
        out[a] = std::fma((float) (idx[a] + g->offset[a]), g->resolution[a], g->start[a]);
    return CGS_OK;
}

cgs_status cgs_grid_subbox(const cgs_grid *g, const int32_t *lo, const int32_t *count, cgs_grid *out) {
    clear_error();
    if (!g || !lo || !count || !out) return set_error(CGS_ERR_INVALID_ARGUMENT, "grid_subbox: null pointer");
    if (g->dim < 1 || g->dim > CGS_MAX_DIM) return set_error(CGS_ERR_INVALID_ARGUMENT, "grid_subbox: bad grid");
    cgs_grid r = *g;
    int64_t total = 1;
    for (int a = 0; a < g->dim; ++a) {
        const int32_t n = (int32_t) g->count[a];
        if (lo[a] < 0 || count[a] < 1 || (int64_t) lo[a] + count[a] > n)
            return set_error(CGS_ERR_INVALID_ARGUMENT, "grid_subbox: box outside the grid");
        r.offset[a] = g->offset[a] + lo[a];
        r.count[a] = (float) count[a];
        r.start[a] = g->start[a];
        r.resolution[a] = g->resolution[a];
        r.end[a] = std::fma((float) (r.offset[a] + count[a] - 1), g->resolution[a], g->start[a]);
        total *= count[a];
    }
    r.num_elements = total;
    r.increments[0] = 1;
    for (int a = 1; a < g->dim; ++a) r.increments[a] = r.increments[a - 1] * count[a - 1];
    *out = r;
    return CGS_OK;
}

}  // extern "C"
