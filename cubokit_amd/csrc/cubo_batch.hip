// cubo_batch.hip — C ABI of batched mask stacks (include/cubo_hip.h, "batches").
//
// Many resident uint16 cubes of uneven scan and detector size, each with its own stack of up to 64 masks,
// run in ONE persistent launch: every block takes (cube, tile) items from a work counter until none is
// left. The counter reset (a one-thread kernel, never a memset) and the persistent kernel, which writes every
// cube's results into mapped pinned host memory as its tiles finish (over PCIe, while the rest still streams),
// are captured once in a hipGraph and replayed per run.
//
// The launch shape (waves per block, resident blocks per compute unit, item order) is chosen by an
// autotuner that times every candidate on the device and keeps the fastest per batch signature for the
// life of the process.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/cubo_hip.h"
#include "cube.hpp"
#include "errors.hpp"
#include "host_math.hpp"
#include "maskstack_kernels.hpp"
#include "stack_prep.hpp"

using cubo::DeviceGuard;
using cubo::fail;
using cubo::hip_fail;

namespace {

struct Shape {
    int waves = 8;          // 8 (256-position tiles) or 4 (128-position tiles)
    int blocks_per_cu = 1;
    int order = 0;          // 0: given order; 1: largest detector first (longest items first); 2: interleaved
                            // so that result writes to host memory overlap the cube reads
    bool operator<(const Shape &o) const {
        return std::tie(waves, blocks_per_cu, order) < std::tie(o.waves, o.blocks_per_cu, o.order);
    }
};

std::mutex g_tune_mutex;
std::map<std::string, Shape> g_tuned;    // batch signature -> chosen shape

}  // namespace

struct cubo_hip_batch {
    int device;
    int n;
    std::vector<cubo_hip_cube *> cubes;
    std::vector<int> k;
    std::vector<int64_t> npos, ndet;
    std::vector<void *> wb;              // per cube: device FP16 hi/lo weights
    std::vector<double *> inv;           // per cube: device 1/scale
    std::vector<size_t> yoff;            // per cube: offset in the pinned result buffer (doubles)
    std::vector<std::array<int, 2>> tile_lo;
    double *yhost = nullptr;             // pinned and mapped: the kernel writes every cube's [npos][k] here
    double *ymap = nullptr;              // its device address
    size_t ytotal = 0;
    cubo::StackLaunch *launches = nullptr;   // device [n], in item order
    cubo::BatchItem *items = nullptr;        // device: the items in the order of the shape
    int64_t max_items = 0;                   // its capacity: the tiles of 128-position items
    unsigned long long *counter = nullptr;
    hipStream_t stream = nullptr;
    Shape shape;
    int64_t total_tiles = 0;
    std::string signature;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    cubo::LaunchTimer timer;     // device time of the graph replays, until queried
    bool ran = false;            // the pinned buffer holds the results of a run
};

namespace {

void drop_graph(cubo_hip_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void free_batch(cubo_hip_batch *b) {
    if (!b) return;
    DeviceGuard g(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    b->timer.release();
    for (void *p : b->wb) if (p) (void)hipFree(p);
    for (double *p : b->inv) if (p) (void)hipFree(p);
    if (b->launches) (void)hipFree(b->launches);
    if (b->items) (void)hipFree(b->items);
    if (b->counter) (void)hipFree(b->counter);
    if (b->yhost) (void)hipHostFree(b->yhost);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

// Rates that set the interleaved order: the stack kernel reads the cubes from HBM at about 6.5 TB/s and
// writes results into mapped host memory at about 52 GB/s (both measured on MI355X). A tile whose results
// take longer to write than its positions take to read is write-heavy.
constexpr double kReadBytesPerS = 6.5e12;
constexpr double kHostWriteBytesPerS = 52e9;

// The items of a shape: every cube's tiles, cube by cube in the given order or largest detector first, or
// interleaved: write-heavy and read-heavy tiles merged (each kind keeping cube order) so that the time the
// items so far spend writing to host memory keeps level with the time they spend reading, and the two overlap.
std::vector<cubo::BatchItem> plan_items(const cubo_hip_batch *b, const Shape &s) {
    const int64_t bm = 32 * s.waves;
    std::vector<int> order(b->n);
    std::iota(order.begin(), order.end(), 0);
    if (s.order == 1)
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->ndet[x] > b->ndet[y]; });
    std::vector<cubo::BatchItem> items, heavy, light;
    std::vector<double> heavy_t[2], light_t[2];      // per item: [0] read seconds, [1] write seconds
    for (int i : order) {
        const int64_t tiles = (b->npos[i] + bm - 1) / bm;
        for (int64_t t = 0; t < tiles; ++t) {
            const cubo::BatchItem it{i, (int32_t)t};
            if (s.order != 2) { items.push_back(it); continue; }
            const double rows = (double)std::min(bm, b->npos[i] - t * bm);
            const double rd = rows * (double)b->ndet[i] * 2 / kReadBytesPerS;
            const double wr = rows * (double)b->k[i] * 8 / kHostWriteBytesPerS;
            if (wr > rd) { heavy.push_back(it); heavy_t[0].push_back(rd); heavy_t[1].push_back(wr); }
            else { light.push_back(it); light_t[0].push_back(rd); light_t[1].push_back(wr); }
        }
    }
    if (s.order != 2) return items;
    double read_s = 0, write_s = 0;
    size_t h = 0, l = 0;
    while (h < heavy.size() || l < light.size()) {
        const bool take_heavy = l == light.size() || (h < heavy.size() && write_s <= read_s);
        if (take_heavy) { items.push_back(heavy[h]); read_s += heavy_t[0][h]; write_s += heavy_t[1][h]; ++h; }
        else { items.push_back(light[l]); read_s += light_t[0][l]; write_s += light_t[1][l]; ++l; }
    }
    return items;
}

// Per-cube descriptors and the item order for a shape; uploads them and sets total_tiles.
hipError_t upload_plan(cubo_hip_batch *b, const Shape &s) {
    std::vector<cubo::StackLaunch> desc(b->n);
    for (int i = 0; i < b->n; ++i) {
        cubo::StackLaunch &a = desc[i];
        a.cube = (const uint16_t *)b->cubes[i]->data;
        a.npos = b->npos[i];
        a.ndet = b->ndet[i];
        a.nchunks = (int)((b->ndet[i] + 63) / 64);
        a.k = b->k[i];
        a.k0 = 0;
        a.kn = b->k[i];
        a.wb = b->wb[i];
        a.inv_scale = b->inv[i];
        a.tile_lo[0] = b->tile_lo[i][0];
        a.tile_lo[1] = b->tile_lo[i][1];
        a.y = b->ymap + b->yoff[i];
    }
    const std::vector<cubo::BatchItem> items = plan_items(b, s);
    const int64_t t = (int64_t)items.size();
    if (t >= INT32_MAX || t > b->max_items) return hipErrorInvalidValue;   // the kernel numbers items in 32 bits
    b->total_tiles = t;
    hipError_t e = hipMemcpyAsync(b->launches, desc.data(), desc.size() * sizeof(desc[0]), hipMemcpyHostToDevice,
                                  b->stream);
    if (e == hipSuccess && t > 0)
        e = hipMemcpyAsync(b->items, items.data(), items.size() * sizeof(items[0]), hipMemcpyHostToDevice,
                           b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    return e;
}

hipError_t launch(cubo_hip_batch *b, const Shape &s) {
    const int blocks = cubo::compute_units(b->device) * s.blocks_per_cu;
    return cubo::launch_batch_stack(b->launches, b->items, b->total_tiles, b->counter, s.waves, blocks, b->stream);
}

int build_graph(cubo_hip_batch *b) {
    drop_graph(b);
    if (hipError_t e = upload_plan(b, b->shape); e != hipSuccess) return hip_fail(e, "uploading the batch plan");
    CUBO_HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    // the kernel writes the results straight into the mapped host buffer: no copy nodes
    hipError_t e = launch(b, b->shape);
    hipGraph_t graph = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->stream, &graph);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return hip_fail(e != hipSuccess ? e : e2, "capturing the batch graph");
    }
    b->graph = graph;
    CUBO_HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    return CUBO_HIP_OK;
}

}  // namespace

extern "C" {

int cubo_hip_batch_create(int32_t n, cubo_hip_cube *const *cubes, const float *const *masks, const int32_t *k,
                          cubo_hip_batch **out) {
    if (!cubes || !masks || !k || !out || n <= 0) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "empty batch");
    *out = nullptr;
    const int dev = cubes[0] ? cubes[0]->device : -1;
    for (int i = 0; i < n; ++i) {
        const cubo_hip_cube *c = cubes[i];
        if (!c || !masks[i]) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "batch item %d is NULL", i);
        if (c->device != dev) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "a batch runs on one device");
        if (int st = cubo::require_resident(c, "a batch")) return st;
        if (c->dtype != CUBO_HIP_U16 || c->ndet % 8 != 0)
            return fail(CUBO_HIP_ERR_UNSUPPORTED, "batch item %d: the batched stack needs a uint16 cube whose "
                        "pattern length is a multiple of 8", i);
        if (k[i] < 1 || k[i] > 64) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "batch item %d: %d masks (1..64)", i, k[i]);
        for (int64_t j = 0; j < (int64_t)k[i] * c->ndet; ++j)
            if (!std::isfinite(masks[i][j])) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "a mask needs finite weights");
    }
    auto *b = new (std::nothrow) cubo_hip_batch();
    if (!b) return fail(CUBO_HIP_ERR_OUT_OF_MEMORY, "host allocation");
    b->device = dev;
    b->n = n;
    DeviceGuard g(dev);
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { free_batch(b); return hip_fail(e, "hipStreamCreate"); }
    for (int i = 0; i < n; ++i) {
        cubo_hip_cube *c = cubes[i];
        b->cubes.push_back(c);
        b->k.push_back(k[i]);
        b->npos.push_back(c->npos);
        b->ndet.push_back(c->ndet);
        b->signature += std::to_string(c->npos) + "x" + std::to_string(c->ndet) + "x" + std::to_string(k[i]) + ";";
        // at most 32 masks: the one-tile layout, and the kernel multiplies one 32-mask tile; else two tiles
        cubo::SplitWeights sw = cubo::split_weights(masks[i], k[i], c->ndet);
        (void)hipStreamSynchronize(c->stream);       // the cube's own uploads are done
        void *wb = nullptr;
        double *inv = nullptr;
        e = hipMalloc(&wb, sw.wb.size() * sizeof(_Float16));
        if (e == hipSuccess) e = hipMemcpy(wb, sw.wb.data(), sw.wb.size() * sizeof(_Float16), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc((void **)&inv, k[i] * sizeof(double));
        if (e == hipSuccess) e = hipMemcpy(inv, sw.inv.data(), k[i] * sizeof(double), hipMemcpyHostToDevice);
        b->wb.push_back(wb);
        b->inv.push_back(inv);
        b->tile_lo.push_back({sw.tile_lo[0], sw.tile_lo[1]});
        b->yoff.push_back(b->ytotal);
        b->ytotal += (size_t)c->npos * k[i];
        if (e != hipSuccess) { free_batch(b); return hip_fail(e, "preparing the batch weights"); }
    }
    e = hipMalloc((void **)&b->launches, n * sizeof(cubo::StackLaunch));
    for (int i = 0; i < n; ++i) b->max_items += (b->npos[i] + 127) / 128;
    if (e == hipSuccess) e = hipMalloc((void **)&b->items, std::max<int64_t>(b->max_items, 1) * sizeof(cubo::BatchItem));
    if (e == hipSuccess) e = hipMalloc((void **)&b->counter, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipHostMalloc((void **)&b->yhost, std::max<size_t>(b->ytotal, 1) * sizeof(double),
                                           hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void **)&b->ymap, b->yhost, 0);
    if (e != hipSuccess) { free_batch(b); return hip_fail(e, "allocating the batch"); }
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        auto it = g_tuned.find(b->signature);
        if (it != g_tuned.end()) b->shape = it->second;
        else b->shape = Shape{8, cubo::batch_blocks_per_cu(8), 0};
    }
    if (int st = build_graph(b)) { free_batch(b); return st; }
    *out = b;
    return CUBO_HIP_OK;
}

int cubo_hip_batch_tune(cubo_hip_batch *b, int32_t *waves, int32_t *blocks_per_cu, int32_t *order) {
    if (!b) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL batch");
    DeviceGuard g(b->device);
    std::vector<Shape> cand;
    for (int w : {8, 4})
        for (int bpc = 1; bpc <= cubo::batch_blocks_per_cu(w); ++bpc)
            for (int o : {0, 1, 2}) cand.push_back(Shape{w, bpc, o});
    hipEvent_t t0 = nullptr, t1 = nullptr;
    CUBO_HIP_TRY(hipEventCreate(&t0), "hipEventCreate");
    if (hipError_t e = hipEventCreate(&t1); e != hipSuccess) {
        (void)hipEventDestroy(t0);
        return hip_fail(e, "hipEventCreate");
    }
    float best = 1e30f;
    Shape best_s = b->shape;
    int status = CUBO_HIP_OK;
    for (const Shape &s : cand) {
        hipError_t e = upload_plan(b, s);
        if (e == hipSuccess) e = launch(b, s);                  // warm-up
        float total = 0;
        for (int rep = 0; rep < 3 && e == hipSuccess; ++rep) {
            e = hipEventRecord(t0, b->stream);
            if (e == hipSuccess) e = launch(b, s);
            if (e == hipSuccess) e = hipEventRecord(t1, b->stream);
            if (e == hipSuccess) e = hipEventSynchronize(t1);
            float ms = 0;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
            total += ms;
        }
        if (e != hipSuccess) { status = hip_fail(e, "timing a batch shape"); break; }
        if (total < best) { best = total; best_s = s; }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (status) {
        // the candidates overwrote the device item list the graph was captured with: put the plan of the kept
        // shape back (a failure here leaves no graph, and cubo_hip_batch_run refuses to run)
        (void)hipStreamSynchronize(b->stream);
        if (build_graph(b) != CUBO_HIP_OK) drop_graph(b);
        return status;
    }
    b->shape = best_s;
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        g_tuned[b->signature] = best_s;
    }
    if (int st = build_graph(b)) return st;
    if (waves) *waves = best_s.waves;
    if (blocks_per_cu) *blocks_per_cu = best_s.blocks_per_cu;
    if (order) *order = best_s.order;
    return CUBO_HIP_OK;
}

int cubo_hip_batch_run(cubo_hip_batch *b, double *const *y) {
    if (!b) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL batch");
    if (!b->exec) return fail(CUBO_HIP_ERR_INCOMPLETE, "the batch has no graph (a failed tune could not restore it)");
    DeviceGuard g(b->device);
    CUBO_HIP_TRY(b->timer.start(b->stream), "hipEventRecord");
    CUBO_HIP_TRY(hipGraphLaunch(b->exec, b->stream), "hipGraphLaunch");
    CUBO_HIP_TRY(b->timer.stop(b->stream), "hipEventRecord");
    CUBO_HIP_TRY(hipStreamSynchronize(b->stream), "waiting for the batch");
    b->ran = true;
    if (y) {
        // the pinned results (one contiguous range, cube after cube) into the caller's arrays, in pieces on
        // several host threads
        const size_t total = b->ytotal * sizeof(double);
        cubo::parallel_pieces(total, cubo::reader_threads(), [&](size_t off, size_t len) {
            for (int i = 0; i < b->n; ++i) {
                const size_t lo = b->yoff[i] * sizeof(double), hi = lo + (size_t)b->npos[i] * b->k[i] * sizeof(double);
                const size_t s0 = std::max(lo, off), s1 = std::min(hi, off + len);
                if (y[i] && s0 < s1) memcpy((char *)y[i] + (s0 - lo), (const char *)b->yhost + s0, s1 - s0);
            }
            return true;
        });
    }
    return CUBO_HIP_OK;
}

int cubo_hip_batch_time(cubo_hip_batch *b, double *total_ms, int64_t *runs) {
    if (!b || !total_ms || !runs) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    DeviceGuard g(b->device);
    CUBO_HIP_TRY(b->timer.total(total_ms, runs), "timing the replays");
    return CUBO_HIP_OK;
}

int cubo_hip_batch_view(const cubo_hip_batch *b, int32_t i, const double **y) {
    if (!b || !y) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (i < 0 || i >= b->n) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "batch item %d of %d", i, b->n);
    if (!b->ran) return fail(CUBO_HIP_ERR_INCOMPLETE, "the batch has not run");
    *y = b->yhost + b->yoff[i];
    return CUBO_HIP_OK;
}

int cubo_hip_batch_destroy(cubo_hip_batch *b) {
    free_batch(b);
    return CUBO_HIP_OK;
}

}  // extern "C"
