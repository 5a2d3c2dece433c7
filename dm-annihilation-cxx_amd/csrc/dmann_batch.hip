// Batched maps: many observers (each with its own nside) rendered from one
// particle set, the whole render -- map clears, the footprint work, the FP64
// finalisation of every map -- captured once in a hipGraph and replayed.
// Two schedules: by default each map runs the single-map pipeline (pre-pass,
// sky-covering kernel, sort, grouped kernel) on buffers shared by the maps;
// dmann_batch_set_tuning with chunk >= 1 selects one persistent launch over
// (map, particle chunk) work items with per-pixel atomics.  No counterpart in
// the reference, which renders one observer per run.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <new>
#include <vector>

#include "dmann.h"
#include "dmann_geom.h"
#include "dmann_kernels.h"

using namespace dmann;

#ifndef DMANN_GROUP_BPC
#define DMANN_GROUP_BPC 8
#endif

struct dmann_batch {
    int device = 0;
    int acc = kAccF32;
    int nmaps = 0;
    hipStream_t stream = nullptr;   // the batch's own stream (graph capture and replay)
    hipStream_t caller = nullptr;   // the caller's stream: renders wait for it and it waits for them
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    std::vector<DevParams> P;
    std::vector<MapRef> maps;
    std::vector<double *> out;
    std::vector<int64_t> npix;
    DevParams *d_P = nullptr;
    MapRef *d_maps = nullptr;
    unsigned long long *d_work = nullptr;
    unsigned long long *d_stats = nullptr;
    int chunk = 0; // 0: per-map pipeline; >= 1: persistent launch, particles per work item
    int blocks = 0;
    int cus = 256;
    // per-map pipeline: each map's parameters (single-map thresholds) and
    // the buffers the maps share
    std::vector<DevParams> Pmap;
    DiscRec *recs = nullptr;
    int64_t recs_cap = 0;
    unsigned *counts = nullptr;
    LaunchCfg cfg{};
    GroupCfg gc{};
    int nside_c_max = 1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *graph_parts = nullptr;
    int64_t graph_n = -1;
    long long replays = 0;
    char msg[512] = {0};
};

static char g_batch_msg[512] = "";

static int bfail(dmann_batch *b, int code, const char *fmt, ...)
{
    char *dst = b ? b->msg : g_batch_msg;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(dst, 512, fmt, ap);
    va_end(ap);
    return code;
}

#define BCHK(b, call)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return bfail((b), DMANN_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

static void drop_graph(dmann_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_parts = nullptr;
    b->graph_n = -1;
}

static void batch_free(dmann_batch *b)
{
    drop_graph(b);
    for (auto &m : b->maps) {
        (void)hipFree(m.f32);
        (void)hipFree(m.lo);
        (void)hipFree(m.f64);
    }
    for (double *o : b->out) (void)hipFree(o);
    (void)hipFree(b->d_P);
    (void)hipFree(b->d_maps);
    (void)hipFree(b->d_work);
    (void)hipFree(b->d_stats);
    (void)hipFree(b->recs);
    (void)hipFree(b->counts);
    (void)hipFree(b->gc.keys);
    (void)hipFree(b->gc.idx);
    (void)hipFree(b->gc.hist);
    (void)hipFree(b->gc.sums);
    (void)hipFree(b->gc.work);
    (void)hipFree(b->gc.tab_f);
    (void)hipFree(b->gc.tab_i);
    if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
    if (b->join_ev) (void)hipEventDestroy(b->join_ev);
    b->fork_ev = b->join_ev = nullptr;
    if (b->stream) (void)hipStreamDestroy(b->stream);
    b->maps.clear();
    b->out.clear();
}

// Coarse sort-cell resolution of a map in the per-map pipeline (the
// single-map default, at most kMaxSortNside).
static int sort_nside(int nside)
{
    const int nc = default_cell_nside(nside);
    return nc < kMaxSortNside ? nc : kMaxSortNside;
}

// Per-map pipeline buffers for n particles (records, sort keys and index
// grow with n; the sort histogram is sized for the finest map), allocated
// before any capture.
static hipError_t ensure_pipeline(dmann_batch *b, int64_t n)
{
    hipError_t e;
    if (!b->counts && (e = hipMalloc(&b->counts, 4 * sizeof(unsigned))) != hipSuccess) return e;
    const int64_t cap = n > 1 ? n : 1;
    if (b->recs_cap < cap) {
        (void)hipFree(b->recs);
        (void)hipFree(b->gc.keys);
        (void)hipFree(b->gc.idx);
        b->recs = nullptr;
        b->gc.keys = b->gc.idx = nullptr;
        b->recs_cap = 0;
        if ((e = hipMalloc(&b->recs, cap * sizeof(DiscRec))) != hipSuccess) return e;
        if (b->acc != kAccF64) {
            if ((e = hipMalloc(&b->gc.keys, cap * sizeof(unsigned))) != hipSuccess) return e;
            if ((e = hipMalloc(&b->gc.idx, cap * sizeof(unsigned))) != hipSuccess) return e;
        }
        b->recs_cap = cap;
    }
    if (b->acc != kAccF64 && !b->gc.hist) {
        const int64_t nbins = 12LL * b->nside_c_max * b->nside_c_max * kSizeBuckets * 16;
        if ((e = hipMalloc(&b->gc.hist, nbins * sizeof(unsigned))) != hipSuccess) return e;
        if ((e = hipMalloc(&b->gc.sums, ((nbins + 1023) / 1024) * sizeof(unsigned))) != hipSuccess) return e;
        if ((e = hipMalloc(&b->gc.work, sizeof(unsigned))) != hipSuccess) return e;
        const size_t rows = (size_t)b->gc.group_blocks * (size_t)b->gc.k * 64;
        if ((e = hipMalloc(&b->gc.tab_f, rows * sizeof(float4))) != hipSuccess) return e;
        if ((e = hipMalloc(&b->gc.tab_i, rows * sizeof(int4))) != hipSuccess) return e;
    }
    return hipSuccess;
}

// Enqueues one full render of the batch on its stream.
static hipError_t enqueue(dmann_batch *b, const Particle *parts, int64_t n)
{
    hipError_t e;
    for (int m = 0; m < b->nmaps; m++)
        if ((e = launch_clear(b->acc, b->maps[m], b->npix[m], b->d_stats + (size_t)m * kNumStats, b->stream)) !=
            hipSuccess)
            return e;
    if (n > 0 && b->chunk == 0) {
        // the single-map pipeline, map after map, on the shared buffers
        for (int m = 0; m < b->nmaps; m++) {
            GroupCfg gc = b->gc;
            gc.nside_c = sort_nside(b->Pmap[m].nside);
            const int ka = auto_group_k(n, 12LL * b->Pmap[m].nside * b->Pmap[m].nside);
            if (gc.k > ka) gc.k = ka;
            if ((e = launch_render(b->acc, parts, n, b->Pmap[m], b->recs, b->counts, b->maps[m],
                                   b->d_stats + (size_t)m * kNumStats, b->cfg, b->stream, nullptr,
                                   b->acc == kAccF64 ? nullptr : &gc)) != hipSuccess)
                return e;
        }
    } else if (n > 0 &&
        (e = launch_batch(b->acc, parts, n, b->d_P, b->d_maps, b->nmaps, b->chunk, b->d_work, b->d_stats, b->blocks,
                          b->stream)) != hipSuccess)
        return e;
    for (int m = 0; m < b->nmaps; m++)
        if ((e = launch_finalize(b->acc, b->maps[m], b->npix[m], b->P[m].dOmega, b->out[m], b->stream)) != hipSuccess)
            return e;
    return hipSuccess;
}

extern "C" {

int dmann_batch_create(dmann_batch **out, const dmann_params *params, int nmaps, int device)
{
    if (!out || !params || nmaps < 1) return bfail(nullptr, DMANN_ERR_ARG, "bad arguments");
    *out = nullptr;
    const int prec = params[0].precision;
    if (prec < DMANN_PREC_FP32 || prec > DMANN_PREC_FP64)
        return bfail(nullptr, DMANN_ERR_ARG,
                     prec == DMANN_PREC_FIXED ? "the reproducible precision is available for single maps only"
                                              : "unknown precision");
    for (int m = 0; m < nmaps; m++) {
        if (params[m].nside < 1 || params[m].nside > kMaxNside)
            return bfail(nullptr, DMANN_ERR_ARG, "map %d: nside %d outside 1..%d", m, params[m].nside, kMaxNside);
        if (params[m].precision != prec)
            return bfail(nullptr, DMANN_ERR_ARG, "map %d: all maps of a batch share one precision", m);
        if (params[m].flux_kind < DMANN_FLUX_ANN_NOSOM || params[m].flux_kind > DMANN_FLUX_ANN_ADIABATIC)
            // (the halo-core plug-in needs a context's core table)
            return bfail(nullptr, DMANN_ERR_ARG, "map %d: unknown flux plug-in %d", m, params[m].flux_kind);
        if (!(params[m].footprint_pad >= 0.0 && params[m].footprint_pad < kPi))
            return bfail(nullptr, DMANN_ERR_ARG, "map %d: footprint_pad outside [0, pi)", m);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return bfail(nullptr, DMANN_ERR_HIP, "no HIP device");
    if (device < 0 || device >= ndev) return bfail(nullptr, DMANN_ERR_ARG, "device %d of %d", device, ndev);
    dmann_batch *b = new (std::nothrow) dmann_batch();
    if (!b) return bfail(nullptr, DMANN_ERR_ARG, "out of host memory");
    b->device = device;
    b->acc = prec == DMANN_PREC_FP32 ? kAccF32 : prec == DMANN_PREC_COMPENSATED ? kAccComp : kAccF64;
    b->nmaps = nmaps;
    auto bail = [&](const char *what, hipError_t err) {
        snprintf(g_batch_msg, sizeof g_batch_msg, "%s: %s", what, hipGetErrorString(err));
        batch_free(b);
        delete b;
        return DMANN_ERR_HIP;
    };
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return bail("hipSetDevice", e);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        b->cus = prop.multiProcessorCount;
    b->blocks = b->cus * 8;
    b->cfg.prepass_blocks = b->cus * 64;
    b->cfg.medium_blocks = b->cus * 8;
    b->cfg.large_blocks = b->cus * 4;
    b->cfg.sky_blocks = b->cus * 12;
    b->gc.k = kDefaultGroupK;
    b->gc.sort_blocks = b->cus * 16;
    b->gc.group_blocks = b->cus * DMANN_GROUP_BPC;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) return bail("stream", e);
    if ((e = hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming)) != hipSuccess) return bail("event", e);
    if ((e = hipEventCreateWithFlags(&b->join_ev, hipEventDisableTiming)) != hipSuccess) return bail("event", e);
    for (int m = 0; m < nmaps; m++) {
        const dmann_params &p = params[m];
        DevParams P;
        fill_devparams(&p, P, 1 << 16);
        b->P.push_back(P);
        DevParams Pm;
        fill_devparams(&p, Pm, 1 << 22); // the single-map thresholds
        b->Pmap.push_back(Pm);
        const int nc = sort_nside(p.nside);
        if (nc > b->nside_c_max) b->nside_c_max = nc;
        b->npix.push_back(P.npix);
        MapRef r{nullptr, nullptr, nullptr};
        double *o = nullptr;
        b->maps.push_back(r);
        b->out.push_back(o);
        MapRef &mr = b->maps.back();
        if (b->acc == kAccF64) {
            if ((e = hipMalloc(&mr.f64, P.npix * sizeof(double))) != hipSuccess) return bail("hipMalloc map", e);
        } else {
            if ((e = hipMalloc(&mr.f32, P.npix * sizeof(float))) != hipSuccess) return bail("hipMalloc map", e);
            if (b->acc == kAccComp && (e = hipMalloc(&mr.lo, P.npix * sizeof(float))) != hipSuccess)
                return bail("hipMalloc map lo", e);
        }
        if ((e = hipMalloc(&b->out.back(), P.npix * sizeof(double))) != hipSuccess) return bail("hipMalloc out", e);
    }
    if ((e = hipMalloc(&b->d_P, nmaps * sizeof(DevParams))) != hipSuccess) return bail("hipMalloc params", e);
    if ((e = hipMalloc(&b->d_maps, nmaps * sizeof(MapRef))) != hipSuccess) return bail("hipMalloc maps", e);
    if ((e = hipMalloc(&b->d_work, sizeof(unsigned long long))) != hipSuccess) return bail("hipMalloc work", e);
    if ((e = hipMalloc(&b->d_stats, (size_t)nmaps * kNumStats * sizeof(unsigned long long))) != hipSuccess)
        return bail("hipMalloc stats", e);
    if ((e = hipMemcpy(b->d_P, b->P.data(), nmaps * sizeof(DevParams), hipMemcpyHostToDevice)) != hipSuccess)
        return bail("copy params", e);
    if ((e = hipMemcpy(b->d_maps, b->maps.data(), nmaps * sizeof(MapRef), hipMemcpyHostToDevice)) != hipSuccess)
        return bail("copy maps", e);
    *out = b;
    return DMANN_OK;
}

int dmann_batch_set_tuning(dmann_batch *b, int chunk, int blocks_per_cu)
{
    if (!b) return DMANN_ERR_ARG;
    if (chunk < 0 || chunk > (1 << 20) || blocks_per_cu < 1 || blocks_per_cu > 64)
        return bfail(b, DMANN_ERR_ARG, "tuning out of range (chunk 0..2^20, blocks per CU 1..64)");
    b->chunk = chunk;
    b->blocks = b->cus * blocks_per_cu;
    drop_graph(b);
    return DMANN_OK;
}

int dmann_batch_render_device(dmann_batch *b, const dmann_particle *dev_parts, int64_t n, int use_graph)
{
    if (!b) return DMANN_ERR_ARG;
    if (n < 0 || (n > 0 && !dev_parts)) return bfail(b, DMANN_ERR_ARG, "bad particle arguments");
    BCHK(b, hipSetDevice(b->device));
    const Particle *parts = reinterpret_cast<const Particle *>(dev_parts);
    if (b->chunk == 0 && n > b->recs_cap) drop_graph(b); // its buffers are reallocated
    if (b->chunk == 0) BCHK(b, ensure_pipeline(b, n));
    // the render reads particles the caller's stream may still be writing,
    // and the caller's later work reads the maps: fork from and join back
    // into the caller's stream
    BCHK(b, hipEventRecord(b->fork_ev, b->caller));
    BCHK(b, hipStreamWaitEvent(b->stream, b->fork_ev, 0));
    if (!use_graph) {
        BCHK(b, enqueue(b, parts, n));
        BCHK(b, hipEventRecord(b->join_ev, b->stream));
        BCHK(b, hipStreamWaitEvent(b->caller, b->join_ev, 0));
        return DMANN_OK;
    }
    if (!b->exec || b->graph_parts != (const void *)dev_parts || b->graph_n != n) {
        drop_graph(b);
        BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = enqueue(b, parts, n);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(b->stream, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return bfail(b, DMANN_ERR_HIP, "graph capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
        }
        b->graph = g;
        BCHK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->graph_parts = dev_parts;
        b->graph_n = n;
    }
    BCHK(b, hipGraphLaunch(b->exec, b->stream));
    b->replays++;
    BCHK(b, hipEventRecord(b->join_ev, b->stream));
    BCHK(b, hipStreamWaitEvent(b->caller, b->join_ev, 0));
    return DMANN_OK;
}

int dmann_batch_set_stream(dmann_batch *b, void *stream)
{
    if (!b) return DMANN_ERR_ARG;
    b->caller = (hipStream_t)stream;
    return DMANN_OK;
}

int dmann_batch_synchronize(dmann_batch *b)
{
    if (!b) return DMANN_ERR_ARG;
    BCHK(b, hipSetDevice(b->device));
    BCHK(b, hipStreamSynchronize(b->stream));
    std::vector<unsigned long long> h((size_t)b->nmaps * kNumStats);
    BCHK(b, hipMemcpy(h.data(), b->d_stats, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (int m = 0; m < b->nmaps; m++)
        if (h[(size_t)m * kNumStats + kStatBad])
            return bfail(b, DMANN_ERR_DOMAIN,
                         "map %d: %llu particle(s) with undefined direction (d == 0 or |cos theta| > 1)", m,
                         h[(size_t)m * kNumStats + kStatBad]);
    return DMANN_OK;
}

int64_t dmann_batch_npix(const dmann_batch *b, int i) { return (b && i >= 0 && i < b->nmaps) ? b->npix[i] : -1; }

int dmann_batch_map_device(dmann_batch *b, int i, double **dev_map)
{
    if (!b || !dev_map || i < 0 || i >= b->nmaps) return bfail(b, DMANN_ERR_ARG, "bad map index");
    *dev_map = b->out[i];
    return DMANN_OK;
}

int dmann_batch_get_map(dmann_batch *b, int i, double *host_map)
{
    if (!b || !host_map || i < 0 || i >= b->nmaps) return bfail(b, DMANN_ERR_ARG, "bad map index");
    BCHK(b, hipSetDevice(b->device));
    BCHK(b, hipStreamSynchronize(b->stream));
    BCHK(b, hipMemcpy(host_map, b->out[i], b->npix[i] * sizeof(double), hipMemcpyDeviceToHost));
    return DMANN_OK;
}

int dmann_batch_get_stats(dmann_batch *b, int i, dmann_stats *st)
{
    if (!b || !st || i < 0 || i >= b->nmaps) return bfail(b, DMANN_ERR_ARG, "bad map index");
    BCHK(b, hipSetDevice(b->device));
    BCHK(b, hipStreamSynchronize(b->stream));
    unsigned long long h[kNumStats];
    BCHK(b, hipMemcpy(h, b->d_stats + (size_t)i * kNumStats, sizeof h, hipMemcpyDeviceToHost));
    st->kept = (int64_t)h[kStatKept];
    st->subpixel = (int64_t)h[kStatSub];
    st->fallback = (int64_t)h[kStatFallback];
    st->extended = (int64_t)h[kStatExt];
    st->contributions = (int64_t)h[kStatContrib];
    // (the persistent launch has no sky-covering class)
    st->large = b->chunk == 0 ? (int64_t)h[kStatLarge] : 0;
    st->large_contributions = b->chunk == 0 ? (int64_t)h[kStatLargeContrib] : 0;
    return DMANN_OK;
}

long long dmann_batch_graph_replays(const dmann_batch *b) { return b ? b->replays : -1; }

void dmann_batch_destroy(dmann_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_free(b);
    delete b;
}

const char *dmann_batch_last_error(const dmann_batch *b) { return b ? b->msg : g_batch_msg; }

} // extern "C"
