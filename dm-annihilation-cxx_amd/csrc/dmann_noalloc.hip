// Allocation-free render: dmann_render_map_noalloc (declared in
// include/dmann.h).  For callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc): the caller passes the particles and the output map, and every
// other buffer the kernels need -- the per-chunk particle records, the class
// counters and the statistics -- lives in `scratch`, a file-scope device
// array of fixed size that each GPU holds until the process ends.
//
// The kernels are the context path's own (pre-pass, sky-covering kernel,
// per-particle wavefront kernel, finalisation); the map accumulates in FP64
// directly in the caller's output (FP32 weights for the FP32 and compensated
// precisions, FP64 weights for FP64) and is divided by dOmega in place.
// Particles go through the scratch records in chunks of kNoAllocChunk.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "dmann.h"
#include "dmann_geom.h"
#include "dmann_kernels.h"

using namespace dmann;

namespace {

// particles per chunk: one DiscRec (48 bytes) each in the scratch records
constexpr int64_t kNoAllocChunk = 1 << 20;
// header: 4 class counters (u32) at 0, kNumStats u64 counters at 64
constexpr size_t kHeaderBytes = 256;
constexpr size_t kScratchBytes = kHeaderBytes + (size_t)kNoAllocChunk * sizeof(DiscRec);
constexpr int kMaxDevices = 64;

} // namespace

// The scratch array.  In host code this name is a stand-in, not a device
// address: host code only ever uses the per-device address that
// hipGetSymbolAddress returns (g_scratch below).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_mu;
unsigned char *g_scratch[kMaxDevices] = {};
int g_cus[kMaxDevices] = {};

// Per-device set-up, once per device with that device current: the
// device's address of `scratch` and its CU count.
hipError_t scratch_for_device(int dev, unsigned char **ptr, int *cus)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1) n = 256;
        g_cus[dev] = n;
        g_scratch[dev] = static_cast<unsigned char *>(p);
    }
    *ptr = g_scratch[dev];
    *cus = g_cus[dev];
    return hipSuccess;
}

// the message goes where dmann_last_error(NULL) reads it
int nfail(int code, const char *msg, hipError_t e = hipSuccess)
{
    char buf[512];
    if (e != hipSuccess) snprintf(buf, sizeof buf, "%s: %s", msg, hipGetErrorString(e));
    else snprintf(buf, sizeof buf, "%s", msg);
    set_global_error(buf);
    return code;
}

} // namespace

extern "C" {

int dmann_render_map_noalloc(const dmann_params *p, const dmann_particle *dev_parts, int64_t n, double *dev_map,
                             void *stream, dmann_stats *host_stats)
{
    if (!p || !dev_map || n < 0 || (n > 0 && !dev_parts))
        return nfail(DMANN_ERR_ARG, "dmann_render_map_noalloc: null pointer or negative particle count");
    if (p->nside < 1 || p->nside > kMaxNside) return nfail(DMANN_ERR_ARG, "dmann_render_map_noalloc: nside out of range");
    if (p->precision < DMANN_PREC_FP32 || p->precision > DMANN_PREC_FP64)
        return nfail(DMANN_ERR_ARG, "dmann_render_map_noalloc: unknown precision (the reproducible precision "
                                    "needs a context)");
    // (the halo-core plug-in needs a context's core table)
    if (p->flux_kind < DMANN_FLUX_ANN_NOSOM || p->flux_kind > DMANN_FLUX_ANN_ADIABATIC)
        return nfail(DMANN_ERR_ARG, "dmann_render_map_noalloc: flux plug-in not available without a context");
    if (!(p->footprint_pad >= 0.0 && p->footprint_pad < kPi))
        return nfail(DMANN_ERR_ARG, "dmann_render_map_noalloc: footprint_pad outside [0, pi)");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return nfail(DMANN_ERR_HIP, "hipGetDevice", e);
    if (dev < 0 || dev >= kMaxDevices) return nfail(DMANN_ERR_ARG, "dmann_render_map_noalloc: device index too large");
    unsigned char *base = nullptr;
    int cus = 256;
    if ((e = scratch_for_device(dev, &base, &cus)) != hipSuccess)
        return nfail(DMANN_ERR_HIP, "hipGetSymbolAddress(scratch)", e);
    unsigned *counts = reinterpret_cast<unsigned *>(base);
    unsigned long long *stats = reinterpret_cast<unsigned long long *>(base + 64);
    DiscRec *recs = reinterpret_cast<DiscRec *>(base + kHeaderBytes);

    const hipStream_t st = (hipStream_t)stream;
    DevParams P;
    fill_devparams(p, P, 1 << 22);
    const int acc = p->precision == DMANN_PREC_FP64 ? kAccF64 : kAccW64;
    const int64_t npix = npix_of(p->nside);
    MapRef map{nullptr, nullptr, dev_map};
    LaunchCfg cfg;
    cfg.prepass_blocks = cus * 64;
    cfg.medium_blocks = cus * 8;
    cfg.large_blocks = cus * 4;
    cfg.sky_blocks = 0; // the gather needs the grouping's key buffer
    // clears the output map and the statistics (fill kernels, no memset)
    if ((e = launch_clear(acc, map, npix, stats, st)) != hipSuccess) return nfail(DMANN_ERR_HIP, "clear", e);
    const Particle *parts = reinterpret_cast<const Particle *>(dev_parts);
    for (int64_t i = 0; i < n; i += kNoAllocChunk) {
        const int64_t m = n - i < kNoAllocChunk ? n - i : kNoAllocChunk;
        if ((e = launch_render(acc, parts + i, m, P, recs, counts, map, stats, cfg, st, nullptr, nullptr)) !=
            hipSuccess)
            return nfail(DMANN_ERR_HIP, "render", e);
    }
    if ((e = launch_finalize(acc, map, npix, P.dOmega, dev_map, st)) != hipSuccess)
        return nfail(DMANN_ERR_HIP, "finalize", e);
    if (!host_stats) return DMANN_OK;
    unsigned long long h[kNumStats];
    if ((e = hipMemcpyAsync(h, stats, sizeof h, hipMemcpyDeviceToHost, st)) != hipSuccess)
        return nfail(DMANN_ERR_HIP, "read statistics", e);
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return nfail(DMANN_ERR_HIP, "synchronize", e);
    host_stats->kept = (int64_t)h[kStatKept];
    host_stats->subpixel = (int64_t)h[kStatSub];
    host_stats->fallback = (int64_t)h[kStatFallback];
    host_stats->extended = (int64_t)h[kStatExt];
    host_stats->contributions = (int64_t)h[kStatContrib];
    host_stats->large = (int64_t)h[kStatLarge];
    host_stats->large_contributions = (int64_t)h[kStatLargeContrib];
    if (h[kStatBad])
        return nfail(DMANN_ERR_DOMAIN, "dmann_render_map_noalloc: particle(s) with undefined direction "
                                       "(d == 0 or |cos theta| > 1)");
    return DMANN_OK;
}

} // extern "C"
