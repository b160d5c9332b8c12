// cgs_batch.hip -- batched ragged grid search: one persistent kernel over
// (pair, grid-tile) work items of uneven size and of every cost kind,
// optionally replayed from a captured hipGraph.
//
// Every work item is one workgroup's tile of grid points of one pair, run by
// the same per-workgroup bodies as the single-pair kernels (fast_block /
// mi_block in cgs_kernels.hpp) on per-pair descriptors kept in device memory,
// so each pair's costs and argmin are identical to cgs_search on that pair.
// Workgroups claim items_per_claim items at a time from one device counter
// and exit when the counter passes the last item, so the ragged mix of image
// sizes, grid sizes and costs load-balances itself; the MI items, the most
// expensive, are numbered first so that the tail is made of cheap ones.  The
// moving image is staged per pair in its kernel's tap table (the fast
// kernel's float differences, the MI kernel's raw float taps, or for uint8 MI
// pairs the 4-byte packed taps), the fixed image as float (exact for uint8).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "cgs.h"
#include "cgs_device.hpp"
#include "cgs_host.hpp"
#include "cgs_kernels.hpp"

namespace cgs {

// The per-workgroup body a descriptor runs: the fast sweep with the SAD / NCC /
// SQD sums (fast_block's cost template argument is the kind itself), or the MI
// sweep on float taps or on uint8 packed taps.
enum BatchKind : int32_t { kBatchSAD = 0, kBatchNCC = 1, kBatchSQD = 2, kBatchMI = 3, kBatchMIU8 = 4 };
constexpr int kBatchKinds = 5;
// Item numbering order of the kinds: MI first (the most work per item).
constexpr int kKindOrder[kBatchKinds] = {kBatchMI, kBatchMIU8, kBatchNCC, kBatchSQD, kBatchSAD};

struct BatchDesc {
    SearchParams P;         // as for a single-pair search, pointing at the batch's staging
    int32_t kind;           // BatchKind
    uint32_t item_begin;    // first work item of the pair
};

// Converts an image to float (exact for uint8).
template <typename T>
__global__ void to_float_kernel(const T *src, float *dst, long n) {
    for (long i = (long) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long) gridDim.x * blockDim.x)
        dst[i] = (float) src[i];
}

// Tap tables as in cgs_search.hip: DIFF selects the fast kernel's
// (p, dx, dy, dxy) form, otherwise the MI kernel's raw (tl, tr, bl, br).
template <typename T, bool DIFF>
__global__ void batch_taps_kernel(const T *src, float4 *dst, int W, int H) {
    const long n = (long) W * H;
    const long total = n + 2L * W + 2;
    for (long j = (long) blockIdx.x * blockDim.x + threadIdx.x; j < total; j += (long) gridDim.x * blockDim.x) {
        const long L = j - (W + 1);
        auto P = [&](long i) { return (i >= 0 && i < n) ? (float) src[i] : 0.0f; };
        const float tl = P(L), tr = P(L + 1), bl = P(L + W), br = P(L + W + 1);
        dst[j] = DIFF ? make_float4(tl, tr - tl, bl - tl, (br - bl) - (tr - tl)) : make_float4(tl, tr, bl, br);
    }
}

__global__ void batch_clear_kernel(unsigned long long *keys, int n, uint32_t *counter) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) keys[i] = ~0ull;
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0u;
}

// The one persistent kernel: GF grid points per fast item, GM per MI item.
// descs: every descriptor, sorted by item_begin (items numbered over all
// descriptors in kKindOrder).  The item's kind is wave-uniform, so a
// workgroup runs one body per item with no divergence.
template <int GF, int GM>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4, 8))) void batch_kernel(
    const BatchDesc *descs, int npairs, uint32_t total, uint32_t *counter, int per_claim) {
    __shared__ union {
        BlockShared<GF> f;
        BlockShared<GM> m;
    } S;
    __shared__ uint32_t s_claim;
    extern __shared__ uint32_t s_hist[];   // MI: GM histograms of the batch's largest bin count
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s_claim = atomicAdd(counter, (uint32_t) per_claim);
        __syncthreads();
        // uniform by construction; readfirstlane tells the compiler so, which
        // keeps the descriptor and the tap table's buffer resource in SGPRs
        const uint32_t first = (uint32_t) __builtin_amdgcn_readfirstlane((int) s_claim);
        __syncthreads();   // every thread has read s_claim before the next claim
        if (first >= total) break;
        const uint32_t last = min(first + (uint32_t) per_claim, total);
        for (uint32_t item = first; item < last; ++item) {
            // the pair owning the item: last descriptor with item_begin <= item
            int lo = 0, hi = npairs - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (descs[mid].item_begin <= item) lo = mid;
                else hi = mid - 1;
            }
            const SearchParams &P = descs[lo].P;
            const int32_t tile = (int32_t) (item - descs[lo].item_begin);
            switch (__builtin_amdgcn_readfirstlane(descs[lo].kind)) {
            case kBatchMIU8: mi_block<uint8_t, float, GM>(P, tile * GM, S.m, s_hist); break;
            case kBatchMI: mi_block<float, float, GM>(P, tile * GM, S.m, s_hist); break;
            case kBatchNCC: fast_block<kBatchNCC, float, GF>(P, tile * GF, S.f); break;
            case kBatchSQD: fast_block<kBatchSQD, float, GF>(P, tile * GF, S.f); break;
            default: fast_block<kBatchSAD, float, GF>(P, tile * GF, S.f); break;
            }
        }
    }
}

}  // namespace cgs

using namespace cgs;

struct cgs_batch_s {
    std::vector<cgs_batch_pair> pairs;
    cgs_tile_config tiles{};
    int cus = 0;
    uint32_t total_items = 0;
    bool has_mi = false;     // some descriptor of the current launch runs the MI sweep
    int nslots = 0;          // descriptors (= keys) of the current launch
    std::vector<int> slot_pair;                                       // pair of each key slot
    int64_t total_points = 0, total_pe = 0;
    // device staging
    DevBuf descs, keys, keys_host, counter;
    std::vector<DevBuf> taps, fixf, sincos;
    std::vector<BatchDesc> host_descs;
    // captured graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
    void drop_graph() {
        if (exec) (void) hipGraphExecDestroy(exec);
        if (graph) (void) hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_stream = nullptr;
    }
    ~cgs_batch_s() {
        drop_graph();
        auto rel = [](DevBuf &b, bool host) {
            if (b.ptr) (void) (host ? hipHostFree(b.ptr) : hipFree(b.ptr));
            b.ptr = nullptr;
        };
        rel(descs, false); rel(keys, false); rel(counter, false); rel(keys_host, true);
        for (auto &b : taps) rel(b, false);
        for (auto &b : fixf) rel(b, false);
        for (auto &b : sincos) rel(b, false);
    }
};

namespace {

bool is_mi_kind(int kind) { return kind == kBatchMI || kind == kBatchMIU8; }

int batch_kind(const cgs_batch_pair &p) {
    switch (p.opts.cost) {
    case CGS_COST_NCC: case CGS_COST_NCC_STREAM: return kBatchNCC;
    case CGS_COST_SQD: case CGS_COST_SQD_STREAM: return kBatchSQD;
    case CGS_COST_MI: case CGS_COST_MI_PLAIN: return p.moved.pixtype == CGS_PIX_U8 ? kBatchMIU8 : kBatchMI;
    default: return kBatchSAD;
    }
}

template <int GF, int GM>
hipError_t launch_batch(cgs_batch *b, hipStream_t s) {
    // the largest MI pair's histograms (bin count and copies both vary by pair)
    size_t words = 0;
    if (b->has_mi)
        for (const cgs_batch_pair &p : b->pairs)
            if (is_mi_kind(batch_kind(p))) words = std::max(words, mi_hist_words(GM, p.opts.mi_bins));
    const size_t lds = words * sizeof(uint32_t);
    const int blocks = std::max(1, b->cus * b->tiles.wg_per_cu);
    hipLaunchKernelGGL((batch_kernel<GF, GM>), dim3(blocks), dim3(kThreads), lds, s, (const BatchDesc *) b->descs.ptr,
                       (int) b->host_descs.size(), b->total_items, (uint32_t *) b->counter.ptr,
                       b->tiles.items_per_claim);
    return hipGetLastError();
}

hipError_t launch_batch_any(cgs_batch *b, hipStream_t s) {
    if (b->host_descs.empty() || b->total_items == 0) return hipSuccess;
    const int m = b->tiles.mi_gpt;
    if (b->tiles.fast_gpt == 6)
        return m == 8 ? launch_batch<6, 8>(b, s) : m == 4 ? launch_batch<6, 4>(b, s) : launch_batch<6, 2>(b, s);
    return m == 8 ? launch_batch<4, 8>(b, s) : m == 4 ? launch_batch<4, 4>(b, s) : launch_batch<4, 2>(b, s);
}

cgs_status parse_tiles(const cgs_tile_config *tiles, cgs_tile_config *out) {
    cgs_tile_config t{6, 4, 4, 8};
    if (tiles) {
        if (tiles->fast_gpt) t.fast_gpt = tiles->fast_gpt;
        if (tiles->mi_gpt) t.mi_gpt = tiles->mi_gpt;
        if (tiles->wg_per_cu) t.wg_per_cu = tiles->wg_per_cu;
        if (tiles->items_per_claim) t.items_per_claim = tiles->items_per_claim;
    }
    if ((t.fast_gpt != 4 && t.fast_gpt != 6) || (t.mi_gpt != 2 && t.mi_gpt != 4 && t.mi_gpt != 8) || t.wg_per_cu < 1 ||
        t.wg_per_cu > 8 || t.items_per_claim < 1 || t.items_per_claim > 64)
        return set_error(CGS_ERR_INVALID_ARGUMENT, "batch: bad tile configuration");
    *out = t;
    return CGS_OK;
}

// Validates the pairs, records them and allocates their staging: float tap
// table and float fixed image per pair, and the (cos, sin) table of each
// pair grid's axis-0 levels.
cgs_status stage_alloc(cgs_batch *b, const cgs_batch_pair *pairs, int32_t n, const cgs_tile_config &t) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return set_error(CGS_ERR_NO_DEVICE, "no HIP device");
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    cgs_status st;
    for (int32_t i = 0; i < n; ++i) {
        const cgs_batch_pair &p = pairs[i];
        if ((st = validate_search(&p.grid, &p.moved, &p.fixed, &p.opts)) != CGS_OK) return st;
        if (is_alt_cost(p.opts.cost))
            return set_error(CGS_ERR_UNSUPPORTED, "line-traversal costs run through cgs_search, not the batch");
        if (p.opts.sampler != CGS_SAMPLER_BILINEAR || p.moved.channels != 1 || p.moved.width > 65536 ||
            p.moved.height > 65536)
            return set_error(CGS_ERR_UNSUPPORTED, "batch pairs must be one-channel bilinear searches");
        if (is_mi_kind(batch_kind(p)) && p.opts.mi_bins > 32 && t.mi_gpt > 2)
            return set_error(CGS_ERR_UNSUPPORTED, "MI with more than 32 bins needs mi_gpt 2");
    }
    b->drop_graph();
    b->pairs.assign(pairs, pairs + n);
    b->tiles = t;
    b->cus = cus;
    b->taps.resize(n);
    b->fixf.resize(n);
    b->sincos.resize(n);
    if ((st = b->keys.ensure(sizeof(unsigned long long) * n)) != CGS_OK ||
        (st = b->keys_host.ensure(sizeof(unsigned long long) * n, true)) != CGS_OK ||
        (st = b->counter.ensure(64)) != CGS_OK ||
        (st = b->descs.ensure(sizeof(BatchDesc) * n)) != CGS_OK)
        return st;
    std::vector<float> sc;
    for (int32_t i = 0; i < n; ++i) {
        const cgs_batch_pair &p = pairs[i];
        const int W = p.moved.width, H = p.moved.height;
        const size_t ntaps = (size_t) W * H + 2 * (size_t) W + 2;
        const size_t tap_bytes = batch_kind(p) == kBatchMIU8 ? 4 : sizeof(float4);
        if ((st = b->taps[i].ensure(ntaps * tap_bytes)) != CGS_OK) return st;
        if ((st = b->fixf[i].ensure((size_t) p.fixed.width * p.fixed.height * sizeof(float))) != CGS_OK) return st;
        const bool has_theta = p.opts.param_mode == CGS_PARAM_CENTRED || p.opts.param_mode == CGS_PARAM_NONCENTRED;
        const int levels0 = has_theta ? (int) p.grid.count[0] : 0;
        if (levels0 > 0) {
            if ((st = b->sincos[i].ensure(sizeof(float) * 2 * levels0)) != CGS_OK) return st;
            sc.resize(2 * levels0);
            for (int l = 0; l < levels0; ++l) {
                const float theta = fmaf((float) (p.grid.offset[0] + l), p.grid.resolution[0], p.grid.start[0]);
                sc[2 * l] = (float) std::cos((double) theta);
                sc[2 * l + 1] = (float) std::sin((double) theta);
            }
            HIP_TRY(hipMemcpy(b->sincos[i].ptr, sc.data(), sizeof(float) * sc.size(), hipMemcpyHostToDevice));
        }
    }
    return CGS_OK;
}

// Per-pair staging kernels: tap table and float fixed image.
cgs_status enqueue_staging(cgs_batch *b, hipStream_t s) {
    const unsigned pb = 256;
    for (size_t i = 0; i < b->pairs.size(); ++i) {
        const cgs_batch_pair &p = b->pairs[i];
        const int kind = batch_kind(p);
        const bool mi = is_mi_kind(kind);
        const int W = p.moved.width, H = p.moved.height;
        float4 *taps = (float4 *) b->taps[i].ptr;
        if (kind == kBatchMIU8) {   // the plain MI kernel's packed uint8 taps
            hipLaunchKernelGGL(tap_u8_kernel, dim3(pb), dim3(256), 0, s, (const uint8_t *) p.moved.data,
                               (uint32_t *) taps, W, H);
        } else if (p.moved.pixtype == CGS_PIX_U8) {
            hipLaunchKernelGGL((batch_taps_kernel<uint8_t, true>), dim3(pb), dim3(256), 0, s,
                               (const uint8_t *) p.moved.data, taps, W, H);
        } else {
            if (mi) hipLaunchKernelGGL((batch_taps_kernel<float, false>), dim3(pb), dim3(256), 0, s,
                                       (const float *) p.moved.data, taps, W, H);
            else hipLaunchKernelGGL((batch_taps_kernel<float, true>), dim3(pb), dim3(256), 0, s,
                                    (const float *) p.moved.data, taps, W, H);
        }
        const long nf = (long) p.fixed.width * p.fixed.height;
        if (p.fixed.pixtype == CGS_PIX_U8)
            hipLaunchKernelGGL(to_float_kernel<uint8_t>, dim3(pb), dim3(256), 0, s, (const uint8_t *) p.fixed.data,
                               (float *) b->fixf[i].ptr, nf);
        else
            hipLaunchKernelGGL(to_float_kernel<float>, dim3(pb), dim3(256), 0, s, (const float *) p.fixed.data,
                               (float *) b->fixf[i].ptr, nf);
    }
    HIP_TRY(hipGetLastError());
    return CGS_OK;
}

// Descriptor of pair i searching g (the pair's grid or a box of it); its
// key slot is its position in the vector handed to set_descs.
BatchDesc describe(const cgs_batch *b, int i, const cgs_grid &g, float *result) {
    const cgs_batch_pair &p = b->pairs[i];
    BatchDesc D;
    std::memset(&D, 0, sizeof(D));
    fill_grid_dev(&g, &D.P.grid);
    int order[kMaxDim];
    if (work_order(g.dim, p.opts.param_mode, order)) set_work_order(&D.P.grid, order);
    const bool has_theta = p.opts.param_mode == CGS_PARAM_CENTRED || p.opts.param_mode == CGS_PARAM_NONCENTRED;
    D.P.sincos = has_theta ? (const float2 *) b->sincos[i].ptr + (g.offset[0] - p.grid.offset[0]) : nullptr;
    D.P.param_mode = p.opts.param_mode;
    D.P.cost = p.opts.cost;
    D.P.quads = (const float4 *) b->taps[i].ptr;
    D.P.mw = p.moved.width;
    D.P.mh = p.moved.height;
    D.P.mc = 1;
    D.P.fix = b->fixf[i].ptr;
    D.P.fw = p.fixed.width;
    D.P.fh = p.fixed.height;
    D.P.cx = (float) p.moved.width / 2;
    D.P.cy = (float) p.moved.height / 2;
    D.P.result = result;
    D.P.key = nullptr;   // set by set_descs once the key buffer is sized
    D.P.tie_mode = p.opts.tie_mode;
    D.P.bins = p.opts.mi_bins;
    D.P.ceil_taps = p.moved.pixtype == CGS_PIX_F32;
    D.kind = batch_kind(p);
    return D;
}

// Orders the slot-ordered descriptors by kind (kKindOrder; slot order kept
// inside a kind), numbers their work items over all of them and uploads them.
cgs_status set_descs(cgs_batch *b, std::vector<BatchDesc> &by_slot, const std::vector<int> &slot_pair) {
    // every slot has its own descriptor and key
    const size_t slots = std::max<size_t>(1, by_slot.size());
    cgs_status st;
    if ((st = b->descs.ensure(sizeof(BatchDesc) * slots)) != CGS_OK ||
        (st = b->keys.ensure(sizeof(unsigned long long) * slots)) != CGS_OK ||
        (st = b->keys_host.ensure(sizeof(unsigned long long) * slots, true)) != CGS_OK)
        return st;
    int64_t points = 0, pe = 0;
    for (size_t k = 0; k < by_slot.size(); ++k) {
        by_slot[k].P.key = (unsigned long long *) b->keys.ptr + k;
        points += by_slot[k].P.grid.n;
        pe += (int64_t) by_slot[k].P.grid.n * by_slot[k].P.fw * by_slot[k].P.fh;
    }
    uint64_t items = 0;
    b->host_descs.clear();
    b->has_mi = false;
    for (int kind : kKindOrder) {
        for (BatchDesc D : by_slot) {
            if (D.kind != kind) continue;
            const int gpt = is_mi_kind(kind) ? b->tiles.mi_gpt : b->tiles.fast_gpt;
            D.item_begin = (uint32_t) items;
            items += (uint64_t) ((D.P.grid.n + gpt - 1) / gpt);
            if (items >= 0xFFFFFFF0ull) return set_error(CGS_ERR_TOO_LARGE, "batch has 2^32 or more work items");
            b->has_mi = b->has_mi || is_mi_kind(kind);
            b->host_descs.push_back(D);
        }
    }
    if (!b->host_descs.empty())
        HIP_TRY(hipMemcpy(b->descs.ptr, b->host_descs.data(), sizeof(BatchDesc) * b->host_descs.size(),
                          hipMemcpyHostToDevice));
    b->nslots = (int) by_slot.size();
    b->slot_pair = slot_pair;
    b->total_items = (uint32_t) items;
    b->total_points = points;
    b->total_pe = pe;
    return CGS_OK;
}

// Clears the keys of the current slots and the work counters, launches the kernels.
cgs_status enqueue_search(cgs_batch *b, hipStream_t s) {
    hipLaunchKernelGGL(batch_clear_kernel, dim3(16), dim3(256), 0, s, (unsigned long long *) b->keys.ptr, b->nslots,
                       (uint32_t *) b->counter.ptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(launch_batch_any(b, s));
    return CGS_OK;
}

cgs_status read_keys(cgs_batch *b, hipStream_t s, std::vector<cgs_argmin> &out) {
    const size_t n = (size_t) b->nslots;
    HIP_TRY(hipMemcpyAsync(b->keys_host.ptr, b->keys.ptr, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const unsigned long long *k = (const unsigned long long *) b->keys_host.ptr;
    out.resize(n);
    for (size_t i = 0; i < n; ++i) out[i] = decode_key(k[i], b->pairs[b->slot_pair[i]].opts.tie_mode);
    return CGS_OK;
}

cgs_status enqueue_run(cgs_batch *b, hipStream_t s) {
    cgs_status st = enqueue_staging(b, s);
    return st != CGS_OK ? st : enqueue_search(b, s);
}

}  // namespace

extern "C" {

cgs_status cgs_batch_create(cgs_batch **out) {
    clear_error();
    if (!out) return set_error(CGS_ERR_INVALID_ARGUMENT, "null out");
    *out = new cgs_batch_s();
    return CGS_OK;
}

cgs_status cgs_batch_destroy(cgs_batch *b) {
    delete b;
    return CGS_OK;
}

cgs_status cgs_batch_plan(cgs_batch *b, const cgs_batch_pair *pairs, int32_t n, const cgs_tile_config *tiles) {
    clear_error();
    if (!b || !pairs || n <= 0) return set_error(CGS_ERR_INVALID_ARGUMENT, "batch_plan: bad arguments");
    cgs_tile_config t;
    cgs_status st = parse_tiles(tiles, &t);
    if (st != CGS_OK) return st;
    if ((st = stage_alloc(b, pairs, n, t)) != CGS_OK) return st;
    std::vector<BatchDesc> by_slot(n);
    std::vector<int> slot_pair(n);
    for (int32_t i = 0; i < n; ++i) {
        by_slot[i] = describe(b, i, pairs[i].grid, pairs[i].result);   // keys in the caller's pair order
        slot_pair[i] = i;
    }
    return set_descs(b, by_slot, slot_pair);
}

cgs_status cgs_batch_run(cgs_batch *b, int32_t use_graph, void *stream) {
    clear_error();
    if (!b || b->pairs.empty() || b->nslots != (int) b->pairs.size())
        return set_error(CGS_ERR_INVALID_ARGUMENT, "batch_run: not planned");
    hipStream_t s = (hipStream_t) stream;
    if (!use_graph) return enqueue_run(b, s);
    if (b->exec && b->graph_stream != s) b->drop_graph();
    if (!b->exec) {
        if (!s) return set_error(CGS_ERR_INVALID_ARGUMENT, "graph capture needs a non-null stream");
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const cgs_status st = enqueue_run(b, s);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (st != CGS_OK) { if (g) (void) hipGraphDestroy(g); return st; }
        if (e != hipSuccess) return set_error(CGS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        b->graph = g;
        HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
        b->graph_stream = s;
    }
    HIP_TRY(hipGraphLaunch(b->exec, s));
    return CGS_OK;
}

cgs_status cgs_batch_results(cgs_batch *b, cgs_argmin *out, void *stream) {
    clear_error();
    if (!b || !out || b->pairs.empty() || b->nslots != (int) b->pairs.size())
        return set_error(CGS_ERR_INVALID_ARGUMENT, "batch_results: bad arguments");
    std::vector<cgs_argmin> r;
    cgs_status st = read_keys(b, (hipStream_t) stream, r);
    if (st != CGS_OK) return st;
    for (size_t i = 0; i < r.size(); ++i) out[i] = r[i];
    return CGS_OK;
}

cgs_status cgs_batch_info(const cgs_batch *b, int64_t *items, int64_t *points, int64_t *pixel_evals) {
    clear_error();
    if (!b) return set_error(CGS_ERR_INVALID_ARGUMENT, "null batch");
    if (items) *items = b->total_items;
    if (points) *points = b->total_points;
    if (pixel_evals) *pixel_evals = b->total_pe;
    return CGS_OK;
}

// Local refinement: every (pair, start) descent searches a (2r+1)^D box of
// the pair's grid around its centre; a descent whose box minimum lies on a
// box face that is not a grid face re-centres there and searches again, up
// to max_steps boxes.  A descent that reaches a centre another start of the
// same pair already holds stops (it would repeat that start's boxes).  Each
// pair keeps the lowest local minimum over its starts (lower index on
// ties).  All descents still moving share one batched launch per step; the
// staging (tap tables, float images) is done once.
cgs_status cgs_batch_refine(cgs_batch *b, const cgs_batch_pair *pairs, int32_t n, const int32_t *centres,
                            const cgs_refine_opts *ro, cgs_refine_result *out, void *stream) {
    clear_error();
    if (!b || !pairs || n <= 0 || !centres || !ro || !out)
        return set_error(CGS_ERR_INVALID_ARGUMENT, "batch_refine: bad arguments");
    if (ro->radius < 1 || ro->radius > 16 || ro->max_steps < 1 || ro->max_steps > 1024 || std::isnan(ro->threshold) ||
        ro->starts < 1 || ro->starts > 64)
        return set_error(CGS_ERR_INVALID_ARGUMENT,
                         "batch_refine: radius 1..16, max_steps 1..1024, starts 1..64, threshold not NaN");
    hipStream_t s = (hipStream_t) stream;
    cgs_tile_config t;
    cgs_status st = parse_tiles(nullptr, &t);
    if (st != CGS_OK) return st;
    if ((st = stage_alloc(b, pairs, n, t)) != CGS_OK) return st;
    if ((st = enqueue_staging(b, s)) != CGS_OK) return st;
    const int S = ro->starts;
    enum { kMoving, kLocalMin, kOnFace, kMerged };
    struct Descent {
        int32_t c[CGS_MAX_DIM];
        int state;
        int32_t index;      // grid index of its last box minimum
        float value;
    };
    std::vector<Descent> D((size_t) n * S);
    for (int32_t i = 0; i < n; ++i) {
        const cgs_grid &g = pairs[i].grid;
        for (int k = 0; k < S; ++k) {
            Descent &d = D[(size_t) i * S + k];
            for (int a = 0; a < g.dim; ++a)
                d.c[a] = std::min(std::max(centres[((size_t) i * S + k) * g.dim + a], 0), (int32_t) g.count[a] - 1);
            d.state = kMoving;
            d.index = -1;
            d.value = INFINITY;
        }
        out[i] = cgs_refine_result{-1, INFINITY, 0, CGS_REFINE_LOCAL, -1, INFINITY, 0};
    }
    auto same_centre = [](const Descent &x, const Descent &y, int dim) {
        for (int a = 0; a < dim; ++a)
            if (x.c[a] != y.c[a]) return false;
        return true;
    };
    std::vector<BatchDesc> descs;
    std::vector<int> slot_pair, slot_descent;
    std::vector<cgs_grid> boxes((size_t) n * S);
    std::vector<cgs_argmin> r;
    for (int step = 1; step <= ro->max_steps; ++step) {
        descs.clear();
        slot_pair.clear();
        slot_descent.clear();
        for (int32_t i = 0; i < n; ++i) {
            const cgs_grid &g = pairs[i].grid;
            for (int k = 0; k < S; ++k) {
                Descent &d = D[(size_t) i * S + k];
                if (d.state != kMoving) continue;
                bool dup = false;
                for (int q = 0; q < k && !dup; ++q) {
                    const Descent &e = D[(size_t) i * S + q];
                    dup = (e.state == kMoving || e.state == kLocalMin) && same_centre(d, e, g.dim);
                }
                if (dup) {
                    d.state = kMerged;
                    continue;
                }
                int32_t lo[CGS_MAX_DIM], cnt[CGS_MAX_DIM];
                for (int a = 0; a < g.dim; ++a) {
                    lo[a] = std::max(0, d.c[a] - ro->radius);
                    const int32_t hi = std::min((int32_t) g.count[a] - 1, d.c[a] + ro->radius);
                    cnt[a] = hi - lo[a] + 1;
                }
                cgs_grid &box = boxes[(size_t) i * S + k];
                if ((st = cgs_grid_subbox(&g, lo, cnt, &box)) != CGS_OK) return st;
                descs.push_back(describe(b, i, box, nullptr));
                slot_pair.push_back(i);
                slot_descent.push_back(i * S + k);
            }
        }
        if (descs.empty()) break;
        if ((st = set_descs(b, descs, slot_pair)) != CGS_OK) return st;
        if ((st = enqueue_search(b, s)) != CGS_OK) return st;
        if ((st = read_keys(b, s, r)) != CGS_OK) return st;
        for (size_t k = 0; k < r.size(); ++k) {
            const int32_t i = slot_pair[k];
            Descent &d = D[slot_descent[k]];
            const cgs_grid &g = pairs[i].grid, &box = boxes[slot_descent[k]];
            out[i].steps += 1;
            out[i].points += box.num_elements;
            if (r[k].index < 0) {   // no finite cost in the box
                d.state = kOnFace;
                continue;
            }
            int32_t rem = r[k].index, full = 0;
            bool face = false;
            for (int a = g.dim - 1; a >= 0; --a) {
                const int32_t j = rem / box.increments[a];
                rem -= j * box.increments[a];
                const int32_t lo = box.offset[a] - g.offset[a], c = lo + j;
                const int32_t hi = lo + (int32_t) box.count[a] - 1;
                if ((c == lo && lo > 0) || (c == hi && hi < (int32_t) g.count[a] - 1)) face = true;
                d.c[a] = c;
                full += c * g.increments[a];
            }
            d.index = full;
            d.value = r[k].value;
            if (!face) d.state = kLocalMin;
            else if (step == ro->max_steps) d.state = kOnFace;
        }
    }
    descs.clear();
    slot_pair.clear();
    for (int32_t i = 0; i < n; ++i) {
        cgs_refine_result &o = out[i];
        bool any_local = false;
        for (int k = 0; k < S; ++k) {
            const Descent &d = D[(size_t) i * S + k];
            if (d.state != kLocalMin) continue;
            if (!any_local || d.value < o.local_value || (d.value == o.local_value && d.index < o.local_index)) {
                o.local_value = d.value;
                o.local_index = d.index;
            }
            any_local = true;
        }
        if (!any_local) {
            o.outcome = CGS_REFINE_FALLBACK_BOUNDARY;
            for (int k = 0; k < S; ++k) {   // report the best box minimum seen
                const Descent &d = D[(size_t) i * S + k];
                if (d.index >= 0 && (o.local_index < 0 || d.value < o.local_value)) {
                    o.local_value = d.value;
                    o.local_index = d.index;
                }
            }
        } else {
            o.outcome = o.local_value <= ro->threshold ? CGS_REFINE_LOCAL : CGS_REFINE_FALLBACK_THRESHOLD;
        }
        if (o.outcome == CGS_REFINE_LOCAL) {
            o.index = o.local_index;
            o.value = o.local_value;
        } else if (ro->fallback) {
            descs.push_back(describe(b, i, pairs[i].grid, nullptr));
            slot_pair.push_back(i);
        }
    }
    if (!descs.empty()) {
        if ((st = set_descs(b, descs, slot_pair)) != CGS_OK) return st;
        if ((st = enqueue_search(b, s)) != CGS_OK) return st;
        if ((st = read_keys(b, s, r)) != CGS_OK) return st;
        for (size_t k = 0; k < r.size(); ++k) {
            cgs_refine_result &o = out[slot_pair[k]];
            o.index = r[k].index;
            o.value = r[k].value;
            o.points += pairs[slot_pair[k]].grid.num_elements;
        }
    }
    b->nslots = 0;   // the batch holds no plan for cgs_batch_run after a refinement
    return CGS_OK;
}

}  // extern "C"
