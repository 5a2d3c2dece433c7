// C ABI of libdmann.so (declared in include/dmann.h).
//
// Replaces the module-global device API of the reference render path
// (src/chealpixrender.cu:28-125) with an opaque, re-entrant context.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "dmann.h"
#include "dmann_footprint.h"
#include "dmann_geom.h"
#include "dmann_kernels.h"

// resident single-wave blocks per CU of the grouped kernel (extra blocks
// find the group counter exhausted and exit)
#ifndef DMANN_GROUP_BPC
#define DMANN_GROUP_BPC 8
#endif

using namespace dmann;

struct dmann_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int acc = kAccF32;
    DevParams P{};
    int64_t max_parts = 0;
    int64_t rec_cap = 0; // records the render buffers hold (>= max_parts)
    int64_t npix = 0;
    DiscRec *d_recs = nullptr;
    unsigned *d_counts = nullptr;
    unsigned long long *d_stats = nullptr;
    MapRef map{nullptr, nullptr, nullptr, nullptr, 1.0};
    double *d_out = nullptr; // finalised FP64 map for host reads
    // the reproducible form: its buffers, and the map's fixed-point scale
    // 2^fix_E, set at the first render after a reset from the kept flux
    // (fix_total, summed over the renders since) and lowered by shifting
    // the map when a later render would overflow it
    DetCfg det{};
    int fix_E = 0;
    bool fix_set = false;
    double fix_total = 0.0;
    LaunchCfg cfg{};
    GroupCfg gc{};
    int cus = 0;
    bool timing = false;
    std::vector<hipEvent_t> events; // 5 per timed render call
    hipEvent_t order_ev = nullptr;  // orders a new stream after the old one
    // host input: dmann_render gathers particles in two pinned staging
    // buffers and renders kStage of them per launch, the copy of one buffer
    // (on copy_stream) overlapping the render of the other
    dmann_particle *pin[2] = {nullptr, nullptr};
    Particle *dstage[2] = {nullptr, nullptr};
    hipEvent_t copied[2] = {nullptr, nullptr};   // H2D of a buffer done
    hipEvent_t consumed[2] = {nullptr, nullptr}; // render that read a buffer done
    hipStream_t copy_stream = nullptr;
    int cur = 0;      // buffer being filled
    int64_t fill = 0; // particles in it
    // halo cores of the ANN_CORE_NOSOM plug-in (device copies)
    float4 *d_halos = nullptr;
    int *d_cell_start = nullptr;
    int *d_cell_list = nullptr;
    bool halos_set = false;
    char msg[512] = {0};
};

static void drop_events(dmann_ctx *c)
{
    for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
    c->events.clear();
}

static char g_create_msg[512] = "";

static int fail(dmann_ctx *c, int code, const char *fmt, ...)
{
    char *dst = c ? c->msg : g_create_msg;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(dst, 512, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(ctx, call)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return fail((ctx), DMANN_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

static void free_group(dmann_ctx *c)
{
    (void)hipFree(c->gc.keys);
    (void)hipFree(c->gc.idx);
    (void)hipFree(c->gc.keys_out);
    (void)hipFree(c->gc.iota);
    (void)hipFree(c->gc.sort_tmp);
    (void)hipFree(c->gc.work);
    (void)hipFree(c->gc.tab_f);
    (void)hipFree(c->gc.tab_i);
    c->gc.keys = c->gc.idx = c->gc.keys_out = c->gc.iota = c->gc.work = nullptr;
    c->gc.sort_tmp = nullptr;
    c->gc.sort_tmp_bytes = 0;
    c->gc.hist = c->gc.sums = nullptr;
    c->gc.tab_f = nullptr;
    c->gc.tab_i = nullptr;
}

// Sort-cell buffers for grouped rendering (k > 0); k == 0 frees them.  On
// an allocation failure everything is freed and the context falls back to
// the per-particle kernel (k = 0), so it never holds k > 0 with missing
// buffers.
static hipError_t alloc_group(dmann_ctx *c, int k, int nside_c)
{
    free_group(c);
    c->gc.k = 0;
    c->gc.nside_c = nside_c;
    if (k <= 0 || c->acc == kAccF64) return hipSuccess;
    const int64_t cap = c->rec_cap ? c->rec_cap : 1;
    const size_t rows = (size_t)c->gc.group_blocks * (size_t)k * 64;
    const size_t bytes = group_sort_bytes(cap);
    if (bytes == 0) return hipErrorInvalidValue;
    hipError_t e;
    if ((e = hipMalloc(&c->gc.keys, cap * sizeof(unsigned))) != hipSuccess ||
        (e = hipMalloc(&c->gc.idx, cap * sizeof(unsigned))) != hipSuccess ||
        (e = hipMalloc(&c->gc.keys_out, cap * sizeof(unsigned))) != hipSuccess ||
        (e = hipMalloc(&c->gc.iota, cap * sizeof(unsigned))) != hipSuccess ||
        (e = hipMalloc(&c->gc.sort_tmp, bytes)) != hipSuccess ||
        (e = hipMalloc(&c->gc.work, sizeof(unsigned))) != hipSuccess ||
        (e = hipMalloc(&c->gc.tab_f, rows * sizeof(float4))) != hipSuccess ||
        (e = hipMalloc(&c->gc.tab_i, rows * sizeof(int4))) != hipSuccess) {
        free_group(c);
        return e;
    }
    c->gc.sort_tmp_bytes = bytes;
    c->gc.k = k;
    return hipSuccess;
}

// True when grouped rendering is configured with all of its buffers.
static bool group_ready(const GroupCfg &gc)
{
    return gc.k > 0 && gc.keys && gc.idx && gc.keys_out && gc.iota && gc.sort_tmp && gc.work && gc.tab_f && gc.tab_i;
}

static void free_halos(dmann_ctx *c)
{
    (void)hipFree(c->d_halos);
    (void)hipFree(c->d_cell_start);
    (void)hipFree(c->d_cell_list);
    c->d_halos = nullptr;
    c->d_cell_start = c->d_cell_list = nullptr;
    c->P.halos = nullptr;
    c->P.cell_start = c->P.cell_list = nullptr;
    c->P.nhalos = 0;
    c->halos_set = false;
}

static void free_stage(dmann_ctx *c)
{
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    for (int b = 0; b < 2; b++) {
        (void)hipHostFree(c->pin[b]);
        (void)hipFree(c->dstage[b]);
        if (c->copied[b]) (void)hipEventDestroy(c->copied[b]);
        if (c->consumed[b]) (void)hipEventDestroy(c->consumed[b]);
        c->pin[b] = nullptr;
        c->dstage[b] = nullptr;
        c->copied[b] = c->consumed[b] = nullptr;
    }
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    c->copy_stream = nullptr;
    c->fill = 0;
}

static void free_det(dmann_ctx *c)
{
    (void)hipFree(c->det.tile_med);
    (void)hipFree(c->det.tile_large);
    (void)hipFree(c->det.tile_sums);
    (void)hipFree(c->det.flux_part);
    (void)hipFree(c->det.flux_total);
    c->det = DetCfg{};
}

// Buffers of the reproducible form for rec_cap records.
static hipError_t alloc_det(dmann_ctx *c)
{
    free_det(c);
    const int64_t cap = c->rec_cap ? c->rec_cap : 1;
    const int64_t tiles = (cap + 511) / 512 + 1;
    hipError_t e;
    if ((e = hipMalloc(&c->det.tile_med, tiles * sizeof(unsigned))) != hipSuccess ||
        (e = hipMalloc(&c->det.tile_large, tiles * sizeof(unsigned))) != hipSuccess ||
        (e = hipMalloc(&c->det.tile_sums, ((tiles + 1023) / 1024) * sizeof(unsigned))) != hipSuccess ||
        (e = hipMalloc(&c->det.flux_part, (c->cfg.prepass_blocks > 0 ? c->cfg.prepass_blocks : 1) * sizeof(double))) !=
            hipSuccess ||
        (e = hipMalloc(&c->det.flux_total, sizeof(double))) != hipSuccess) {
        free_det(c);
        return e;
    }
    return hipSuccess;
}

static void free_all(dmann_ctx *c)
{
    free_stage(c);
    free_det(c);
    free_group(c);
    free_halos(c);
    hipFree(c->d_recs);
    hipFree(c->d_counts);
    hipFree(c->d_stats);
    hipFree(c->map.f32);
    hipFree(c->map.lo);
    hipFree(c->map.f64);
    hipFree(c->map.fix);
    hipFree(c->d_out);
    c->d_recs = nullptr;
    c->d_counts = nullptr;
    c->d_stats = nullptr;
    c->map = MapRef{nullptr, nullptr, nullptr, nullptr, 1.0};
    c->d_out = nullptr;
}

namespace dmann {
void set_global_error(const char *msg) { snprintf(g_create_msg, sizeof g_create_msg, "%s", msg); }

void fill_devparams(const dmann_params *p, DevParams &P, int large_pix)
{
    P = DevParams{};
    for (int i = 0; i < 3; i++) P.opos[i] = p->opos[i];
    for (int i = 0; i < 9; i++) P.rot[i] = p->rotmatrix[i];
    P.flux_factor = p->flux_factor;
    P.mass_cut = p->mass_cut;
    P.nside = p->nside;
    P.npix = (int)npix_of(p->nside);
    P.dOmega = 4 * kPi / (double)P.npix;              // Map::setNside, datatypes.h:248-253
    P.theta0 = acos(1.0 - P.dOmega / (2.0 * 3.141592)); // src/genmap.cpp:38
    P.large_pix = large_pix;
    P.flux_kind = p->flux_kind;
    P.decay_factor = p->decay_factor;
    P.vel_to_cgs = p->vel_to_cgs;
    P.c_in_cgs = p->c_in_cgs;
    for (int i = 0; i < 3; i++) P.cpos[i] = p->cpos[i];
    P.footprint_pad = p->footprint_pad;
    P.subpixel_sort = 1;
}
} // namespace dmann

// Particles per host-input launch: small reader chunks (the reference's
// mem_buffer, 2048 in src/VL2_info.ini:57) are gathered into launches of
// this size, larger calls are split.
constexpr int64_t kStage = 1 << 22;

// Pinned and device staging buffers, and render buffers for kStage records.
static int ensure_stage(dmann_ctx *c)
{
    if (c->pin[0]) return DMANN_OK;
    if (c->rec_cap < kStage) {
        const int k = c->gc.k, nside_c = c->gc.nside_c;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        (void)hipFree(c->d_recs);
        c->d_recs = nullptr;
        c->rec_cap = 0;
        HIPCHK(c, hipMalloc(&c->d_recs, kStage * sizeof(DiscRec)));
        c->rec_cap = kStage;
        if (k > 0) HIPCHK(c, alloc_group(c, k, nside_c));
        if (c->acc == kAccFix) HIPCHK(c, alloc_det(c));
    }
    for (int b = 0; b < 2; b++) {
        HIPCHK(c, hipHostMalloc((void **)&c->pin[b], kStage * sizeof(dmann_particle), hipHostMallocDefault));
        HIPCHK(c, hipMalloc(&c->dstage[b], kStage * sizeof(Particle)));
        HIPCHK(c, hipEventCreateWithFlags(&c->copied[b], hipEventDisableTiming));
        HIPCHK(c, hipEventCreateWithFlags(&c->consumed[b], hipEventDisableTiming));
    }
    HIPCHK(c, hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    return DMANN_OK;
}

// The fixed-point scale of the reproducible form before a render of n
// particles: their kept flux (a bound on any pixel's share) is summed on
// the device in a fixed order and read back; the first render after a
// reset sets the scale 2^E so that the flux so far stays below 2^58 (room
// for 16 times more in later renders), and a later render that would take
// the total past 2^61 shifts the map down first.
static int fix_prepare(dmann_ctx *c, const Particle *parts, int64_t n)
{
    const int blocks = c->cfg.prepass_blocks > 0 ? c->cfg.prepass_blocks : 1;
    HIPCHK(c, launch_flux_total(parts, n, c->P, c->det, blocks, c->stream));
    double t = 0.0;
    HIPCHK(c, hipMemcpyAsync(&t, c->det.flux_total, sizeof t, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (!(t >= 0.0) || t > 1e300) return fail(c, DMANN_ERR_DOMAIN, "kept flux %g cannot be held in fixed point", t);
    c->fix_total += t;
    if (c->fix_total <= 0.0) return DMANN_OK; // nothing deposited yet
    int ex = 0;
    (void)frexp(c->fix_total, &ex); // fix_total < 2^ex
    if (!c->fix_set) {
        c->fix_E = 58 - ex;
        c->fix_set = true;
    } else if (ex + c->fix_E > 61) {
        const int k = ex + c->fix_E - 58;
        HIPCHK(c, launch_fix_shift(c->map.fix, c->npix, k, c->stream));
        c->fix_E -= k;
    }
    c->map.fix_scale = ldexp(1.0, c->fix_E);
    return DMANN_OK;
}

// Launches one render of n device-resident particles on the context's
// stream (the reproducible form first fixes the map's scale).
static int render_launch(dmann_ctx *c, const Particle *parts, int64_t n, hipEvent_t *ev)
{
    if (c->acc == kAccFix) {
        const int r = fix_prepare(c, parts, n);
        if (r != DMANN_OK) return r;
    }
    GroupCfg gc = c->gc;
    if (gc.k_auto && gc.k > 0) {
        const int ka = auto_group_k(n, c->npix);
        if (ka < gc.k) gc.k = ka;
    }
    HIPCHK(c, launch_render(c->acc, parts, n, c->P, c->d_recs, c->d_counts, c->map, c->d_stats, c->cfg, c->stream, ev,
                            &gc, c->acc == kAccFix ? &c->det : nullptr));
    return DMANN_OK;
}

// Renders the particles gathered in the current staging buffer.
static int stage_flush(dmann_ctx *c)
{
    if (c->fill == 0) return DMANN_OK;
    const int b = c->cur;
    // the copy may overwrite the device buffer once the render that last
    // read it is done; the render waits for its copy
    HIPCHK(c, hipStreamWaitEvent(c->copy_stream, c->consumed[b], 0));
    HIPCHK(c, hipMemcpyAsync(c->dstage[b], c->pin[b], c->fill * sizeof(Particle), hipMemcpyHostToDevice,
                             c->copy_stream));
    HIPCHK(c, hipEventRecord(c->copied[b], c->copy_stream));
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->copied[b], 0));
    hipEvent_t *ev = nullptr;
    if (c->timing) {
        for (int i = 0; i < 5; i++) {
            hipEvent_t e;
            HIPCHK(c, hipEventCreate(&e));
            c->events.push_back(e);
        }
        ev = &c->events[c->events.size() - 5];
    }
    const int r = render_launch(c, c->dstage[b], c->fill, ev);
    if (r != DMANN_OK) return r;
    HIPCHK(c, hipEventRecord(c->consumed[b], c->stream));
    c->cur ^= 1;
    c->fill = 0;
    return DMANN_OK;
}

#define FLUSH(c)                                                                                   \
    do {                                                                                           \
        const int f_ = stage_flush(c);                                                             \
        if (f_ != DMANN_OK) return f_;                                                             \
    } while (0)

extern "C" {

int dmann_create(dmann_ctx **out, const dmann_params *p, int64_t max_parts, int device)
{
    if (!out || !p) return fail(nullptr, DMANN_ERR_ARG, "null argument");
    *out = nullptr;
    if (p->nside < 1 || p->nside > kMaxNside)
        return fail(nullptr, DMANN_ERR_ARG, "nside %d outside 1..%d", p->nside, kMaxNside);
    if (p->flux_kind < DMANN_FLUX_ANN_NOSOM || p->flux_kind > DMANN_FLUX_ANN_CORE_NOSOM)
        return fail(nullptr, DMANN_ERR_ARG, "unknown flux plug-in %d", p->flux_kind);
    if (!(p->footprint_pad >= 0.0 && p->footprint_pad < kPi))
        return fail(nullptr, DMANN_ERR_ARG, "footprint_pad %g outside [0, pi)", p->footprint_pad);
    if (max_parts < 0 || max_parts > (int64_t)INT32_MAX * 2)
        return fail(nullptr, DMANN_ERR_ARG, "max_parts %lld outside 0..%lld", (long long)max_parts,
                    (long long)INT32_MAX * 2);
    if (p->precision < DMANN_PREC_FP32 || p->precision > DMANN_PREC_FIXED)
        return fail(nullptr, DMANN_ERR_ARG, "unknown precision %d", p->precision);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(nullptr, DMANN_ERR_HIP, "no HIP device");
    if (device < 0 || device >= ndev) return fail(nullptr, DMANN_ERR_ARG, "device %d of %d", device, ndev);

    dmann_ctx *c = new (std::nothrow) dmann_ctx();
    if (!c) return fail(nullptr, DMANN_ERR_ARG, "out of host memory");
    c->device = device;
    c->acc = p->precision == DMANN_PREC_FP32          ? kAccF32
             : p->precision == DMANN_PREC_COMPENSATED ? kAccComp
             : p->precision == DMANN_PREC_FIXED       ? kAccFix
                                                      : kAccF64;
    c->max_parts = max_parts;
    c->rec_cap = max_parts;
    c->npix = npix_of(p->nside);
    fill_devparams(p, c->P, 1 << 22);

    hipError_t e = hipSetDevice(device);
    int dev_cus = 256;
    if (e == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
            dev_cus = prop.multiProcessorCount;
    }
    c->cus = dev_cus;
    c->cfg.prepass_blocks = dev_cus * 64;
    c->cfg.medium_blocks = dev_cus * 8;
    c->cfg.large_blocks = dev_cus * 4;
    c->cfg.sky_blocks = dev_cus * 12;
    auto bail = [&](const char *what, hipError_t err) {
        snprintf(g_create_msg, sizeof g_create_msg, "%s: %s", what, hipGetErrorString(err));
        free_all(c);
        delete c;
        return DMANN_ERR_HIP;
    };
    if (e != hipSuccess) return bail("hipSetDevice", e);
    const int64_t np = c->npix;
    if (c->acc == kAccF64) {
        if ((e = hipMalloc(&c->map.f64, np * sizeof(double))) != hipSuccess) return bail("hipMalloc map", e);
    } else if (c->acc == kAccFix) {
        if ((e = hipMalloc(&c->map.fix, np * sizeof(unsigned long long))) != hipSuccess)
            return bail("hipMalloc map", e);
    } else {
        if ((e = hipMalloc(&c->map.f32, np * sizeof(float))) != hipSuccess) return bail("hipMalloc map", e);
        if (c->acc == kAccComp && (e = hipMalloc(&c->map.lo, np * sizeof(float))) != hipSuccess)
            return bail("hipMalloc map lo", e);
    }
    if ((e = hipMalloc(&c->d_recs, (c->rec_cap ? c->rec_cap : 1) * sizeof(DiscRec))) != hipSuccess)
        return bail("hipMalloc records", e);
    if ((e = hipMalloc(&c->d_counts, 4 * sizeof(unsigned))) != hipSuccess) return bail("hipMalloc counts", e);
    if ((e = hipMalloc(&c->d_stats, kNumStats * sizeof(unsigned long long))) != hipSuccess)
        return bail("hipMalloc stats", e);
    c->gc.sort_blocks = dev_cus * 16;
    c->gc.group_blocks = dev_cus * DMANN_GROUP_BPC;
    if ((e = alloc_group(c, kDefaultGroupK, default_cell_nside(p->nside))) != hipSuccess) return bail("hipMalloc sort cells", e);
    c->gc.k_auto = true;
    if (c->acc == kAccFix && (e = alloc_det(c)) != hipSuccess) return bail("hipMalloc reproducible-form buffers", e);
    if ((e = launch_clear(c->acc, c->map, np, c->d_stats, c->stream)) != hipSuccess) return bail("clear", e);
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return bail("clear sync", e);
    *out = c;
    return DMANN_OK;
}

int dmann_set_tuning(dmann_ctx *c, int medium_blocks_per_cu, int large_blocks_per_cu, int large_pix)
{
    if (!c) return DMANN_ERR_ARG;
    if (medium_blocks_per_cu < 1 || medium_blocks_per_cu > 64 || large_blocks_per_cu < 1 ||
        large_blocks_per_cu > 64 || large_pix < 16)
        return fail(c, DMANN_ERR_ARG, "tuning out of range");
    int cus = c->cfg.medium_blocks / 8 > 0 ? c->cfg.medium_blocks / 8 : 1;
    if (c->cus > 0) cus = c->cus;
    c->cfg.medium_blocks = cus * medium_blocks_per_cu;
    c->cfg.large_blocks = cus * large_blocks_per_cu;
    c->P.large_pix = large_pix;
    return DMANN_OK;
}

int dmann_set_subpixel_sort(dmann_ctx *c, int on)
{
    if (!c) return DMANN_ERR_ARG;
    if (on != 0 && on != 1) return fail(c, DMANN_ERR_ARG, "subpixel sort flag %d is not 0 or 1", on);
    c->P.subpixel_sort = on;
    return DMANN_OK;
}

int dmann_set_sky_gather(dmann_ctx *c, int blocks_per_cu)
{
    if (!c) return DMANN_ERR_ARG;
    if (blocks_per_cu < 0 || blocks_per_cu > 32)
        return fail(c, DMANN_ERR_ARG, "sky gather workgroups per CU %d outside 0..32", blocks_per_cu);
    c->cfg.sky_blocks = (c->cus > 0 ? c->cus : 1) * blocks_per_cu;
    return DMANN_OK;
}

int dmann_set_grouping(dmann_ctx *c, int k, int nside_c)
{
    if (!c) return DMANN_ERR_ARG;
    if (k < 0 || k > 64) return fail(c, DMANN_ERR_ARG, "group size %d outside 0..64", k);
    if (nside_c < 1 || nside_c > c->P.nside || nside_c > kMaxNestNside)
        return fail(c, DMANN_ERR_ARG, "sort cell nside %d outside 1..min(nside, %d)", nside_c, kMaxNestNside);
    HIPCHK(c, hipSetDevice(c->device));
    FLUSH(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, alloc_group(c, k, nside_c));
    c->gc.k_auto = false;
    return DMANN_OK;
}

int dmann_set_halo_cores(dmann_ctx *c, const float *halos, int64_t n)
{
    if (!c) return DMANN_ERR_ARG;
    if (n < 0 || (n > 0 && !halos) || n > (int64_t)INT32_MAX / 8)
        return fail(c, DMANN_ERR_ARG, "bad halo table (%lld rows)", (long long)n);
    HIPCHK(c, hipSetDevice(c->device));
    FLUSH(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_halos(c);
    c->halos_set = true;
    if (n == 0) return DMANN_OK;
    // rows with the core radius of HaloCore::getCorrection (halocore.cpp:56-59)
    std::vector<float4> rows((size_t)n);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, rmax = 0;
    for (int64_t i = 0; i < n; i++) {
        const float *h = halos + 4 * i;
        const float rc = (float)(h[3] / (0.73 * 40000.0));
        rows[i] = make_float4(h[0], h[1], h[2], rc);
        const double ext = fabs((double)rc) * (1 + 1e-6) + 1e-6 * (1 + fabs(h[0]) + fabs(h[1]) + fabs(h[2]));
        for (int a = 0; a < 3; a++) {
            lo[a] = fmin(lo[a], (double)h[a] - ext);
            hi[a] = fmax(hi[a], (double)h[a] + ext);
        }
        rmax = fmax(rmax, ext);
        if (!std::isfinite(h[0]) || !std::isfinite(h[1]) || !std::isfinite(h[2]) || !std::isfinite(rc))
            return fail(c, DMANN_ERR_ARG, "halo row %lld is not finite", (long long)i);
    }
    // cells at least one core diameter wide, at most 128 per axis
    double span = 0;
    for (int a = 0; a < 3; a++) span = fmax(span, hi[a] - lo[a]);
    const double cell = fmax(fmax(2 * rmax, span / 128.0), 1e-30);
    int dim[3];
    for (int a = 0; a < 3; a++) dim[a] = (int)fmin(128.0, floor((hi[a] - lo[a]) / cell) + 1);
    const int64_t ncell = (int64_t)dim[0] * dim[1] * dim[2];
    std::vector<int> start((size_t)ncell + 1, 0);
    auto cells_of = [&](int64_t i, int lo_c[3], int hi_c[3]) {
        const float4 &r = rows[i];
        const double ext = fabs((double)r.w) * (1 + 1e-6) + 1e-6 * (1 + fabs(r.x) + fabs(r.y) + fabs(r.z));
        const double cc[3] = {r.x, r.y, r.z};
        for (int a = 0; a < 3; a++) {
            lo_c[a] = (int)fmax(0.0, floor((cc[a] - ext - lo[a]) / cell));
            hi_c[a] = (int)fmin((double)dim[a] - 1, floor((cc[a] + ext - lo[a]) / cell));
        }
    };
    for (int64_t i = 0; i < n; i++) {
        int a0[3], a1[3];
        cells_of(i, a0, a1);
        for (int z = a0[2]; z <= a1[2]; z++)
            for (int y = a0[1]; y <= a1[1]; y++)
                for (int x = a0[0]; x <= a1[0]; x++) start[((int64_t)z * dim[1] + y) * dim[0] + x + 1]++;
    }
    for (int64_t k = 0; k < ncell; k++) start[k + 1] += start[k];
    std::vector<int> list((size_t)start[ncell]);
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int64_t i = 0; i < n; i++) {
        int a0[3], a1[3];
        cells_of(i, a0, a1);
        for (int z = a0[2]; z <= a1[2]; z++)
            for (int y = a0[1]; y <= a1[1]; y++)
                for (int x = a0[0]; x <= a1[0]; x++) list[fill[((int64_t)z * dim[1] + y) * dim[0] + x]++] = (int)i;
    }
    HIPCHK(c, hipMalloc(&c->d_halos, n * sizeof(float4)));
    HIPCHK(c, hipMalloc(&c->d_cell_start, (ncell + 1) * sizeof(int)));
    HIPCHK(c, hipMalloc(&c->d_cell_list, (list.size() ? list.size() : 1) * sizeof(int)));
    HIPCHK(c, hipMemcpy(c->d_halos, rows.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_cell_start, start.data(), (ncell + 1) * sizeof(int), hipMemcpyHostToDevice));
    if (!list.empty())
        HIPCHK(c, hipMemcpy(c->d_cell_list, list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice));
    c->P.halos = c->d_halos;
    c->P.cell_start = c->d_cell_start;
    c->P.cell_list = c->d_cell_list;
    for (int a = 0; a < 3; a++) {
        c->P.grid_lo[a] = lo[a];
        c->P.grid_dim[a] = dim[a];
    }
    c->P.grid_inv = 1.0 / cell;
    c->P.nhalos = (int)n;
    return DMANN_OK;
}

int dmann_set_stream(dmann_ctx *c, void *stream)
{
    if (!c) return DMANN_ERR_ARG;
    const hipStream_t s = (hipStream_t)stream;
    if (s == c->stream) return DMANN_OK;
    FLUSH(c);
    // work queued on the old stream (map clears, renders, reads) comes
    // before anything enqueued on the new one
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->order_ev) HIPCHK(c, hipEventCreateWithFlags(&c->order_ev, hipEventDisableTiming));
    HIPCHK(c, hipEventRecord(c->order_ev, c->stream));
    HIPCHK(c, hipStreamWaitEvent(s, c->order_ev, 0));
    c->stream = s;
    return DMANN_OK;
}

int dmann_render_device(dmann_ctx *c, const dmann_particle *dev_parts, int64_t n)
{
    if (!c) return DMANN_ERR_ARG;
    if (n < 0) return fail(c, DMANN_ERR_ARG, "negative particle count");
    if (n == 0) return DMANN_OK;
    if (!dev_parts) return fail(c, DMANN_ERR_ARG, "null particle pointer");
    if (c->P.flux_kind == DMANN_FLUX_ANN_CORE_NOSOM && !c->halos_set)
        return fail(c, DMANN_ERR_ARG, "the halo-core flux plug-in needs dmann_set_halo_cores first");
    if (n > c->max_parts)
        return fail(c, DMANN_ERR_CAPACITY, "%lld particles exceed the context capacity %lld", (long long)n,
                    (long long)c->max_parts);
    HIPCHK(c, hipSetDevice(c->device));
    FLUSH(c);
    hipEvent_t *ev = nullptr;
    if (c->timing) {
        for (int i = 0; i < 5; i++) {
            hipEvent_t e;
            HIPCHK(c, hipEventCreate(&e));
            c->events.push_back(e);
        }
        ev = &c->events[c->events.size() - 5];
    }
    if (c->gc.k > 0 && !group_ready(c->gc)) return fail(c, DMANN_ERR_HIP, "grouping buffers are missing");
    return render_launch(c, reinterpret_cast<const Particle *>(dev_parts), n, ev);
}

int dmann_adjoint_device(dmann_ctx *c, const dmann_particle *dev_parts, int64_t n, const double *dev_grad_map,
                         double *dev_grad_parts)
{
    if (!c) return DMANN_ERR_ARG;
    if (n < 0) return fail(c, DMANN_ERR_ARG, "negative particle count");
    if (n == 0) return DMANN_OK;
    if (!dev_parts || !dev_grad_map || !dev_grad_parts) return fail(c, DMANN_ERR_ARG, "null pointer");
    if (c->P.flux_kind == DMANN_FLUX_ANN_ADIABATIC || c->P.flux_kind == DMANN_FLUX_ANN_CORE_NOSOM)
        return fail(c, DMANN_ERR_ARG, "no adjoint for the adiabatic or halo-core flux plug-ins (their factor "
                                      "depends on position)");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, launch_adjoint(reinterpret_cast<const Particle *>(dev_parts), n, c->P, dev_grad_map, dev_grad_parts,
                             c->cfg.medium_blocks, c->stream));
    return DMANN_OK;
}

int dmann_timing(dmann_ctx *c, int enable, double *ms, int *calls)
{
    if (!c) return DMANN_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    FLUSH(c);
    if (ms || calls) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        double acc[4] = {0, 0, 0, 0};
        for (size_t i = 0; i + 4 < c->events.size(); i += 5)
            for (int k = 0; k < 4; k++) {
                float t = 0;
                HIPCHK(c, hipEventElapsedTime(&t, c->events[i + k], c->events[i + k + 1]));
                acc[k] += t;
            }
        if (ms)
            for (int k = 0; k < 4; k++) ms[k] = acc[k];
        if (calls) *calls = (int)(c->events.size() / 5);
    }
    drop_events(c);
    c->timing = enable != 0;
    return DMANN_OK;
}

int dmann_render(dmann_ctx *c, const dmann_particle *host_parts, int64_t n)
{
    if (!c) return DMANN_ERR_ARG;
    if (n < 0) return fail(c, DMANN_ERR_ARG, "negative particle count");
    if (n == 0) return DMANN_OK;
    if (!host_parts) return fail(c, DMANN_ERR_ARG, "null particle pointer");
    if (c->P.flux_kind == DMANN_FLUX_ANN_CORE_NOSOM && !c->halos_set)
        return fail(c, DMANN_ERR_ARG, "the halo-core flux plug-in needs dmann_set_halo_cores first");
    if (n > c->max_parts)
        return fail(c, DMANN_ERR_CAPACITY, "%lld particles exceed the context capacity %lld", (long long)n,
                    (long long)c->max_parts);
    HIPCHK(c, hipSetDevice(c->device));
    if (c->gc.k > 0 && !group_ready(c->gc)) return fail(c, DMANN_ERR_HIP, "grouping buffers are missing");
    const int e = ensure_stage(c);
    if (e != DMANN_OK) return e;
    // the caller's buffer is copied into pinned staging before returning
    const dmann_particle *src = host_parts;
    int64_t left = n;
    while (left > 0) {
        if (c->fill == 0) HIPCHK(c, hipEventSynchronize(c->copied[c->cur]));
        const int64_t take = left < kStage - c->fill ? left : kStage - c->fill;
        memcpy(c->pin[c->cur] + c->fill, src, (size_t)take * sizeof(dmann_particle));
        c->fill += take;
        src += take;
        left -= take;
        if (c->fill == kStage) FLUSH(c);
    }
    return DMANN_OK;
}

int dmann_synchronize(dmann_ctx *c)
{
    if (!c) return DMANN_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    FLUSH(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    unsigned long long bad = 0;
    HIPCHK(c, hipMemcpy(&bad, c->d_stats + kStatBad, sizeof bad, hipMemcpyDeviceToHost));
    if (bad)
        return fail(c, DMANN_ERR_DOMAIN, "%llu particle(s) with undefined direction (d == 0 or |cos theta| > 1)",
                    bad);
    return DMANN_OK;
}

int dmann_get_map_device(dmann_ctx *c, double *dev_map)
{
    if (!c) return DMANN_ERR_ARG;
    if (!dev_map) return fail(c, DMANN_ERR_ARG, "null map pointer");
    HIPCHK(c, hipSetDevice(c->device));
    FLUSH(c);
    HIPCHK(c, launch_finalize(c->acc, c->map, c->npix, c->P.dOmega, dev_map, c->stream));
    return DMANN_OK;
}

int dmann_get_map(dmann_ctx *c, double *host_map)
{
    if (!c) return DMANN_ERR_ARG;
    if (!host_map) return fail(c, DMANN_ERR_ARG, "null map pointer");
    HIPCHK(c, hipSetDevice(c->device));
    FLUSH(c);
    if (!c->d_out) HIPCHK(c, hipMalloc(&c->d_out, c->npix * sizeof(double)));
    HIPCHK(c, launch_finalize(c->acc, c->map, c->npix, c->P.dOmega, c->d_out, c->stream));
    HIPCHK(c, hipMemcpyAsync(host_map, c->d_out, c->npix * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return dmann_synchronize(c);
}

int dmann_reset(dmann_ctx *c)
{
    if (!c) return DMANN_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    c->fill = 0; // gathered particles not yet rendered are dropped with the map
    c->fix_set = false;
    c->fix_total = 0.0;
    c->map.fix_scale = 1.0;
    HIPCHK(c, launch_clear(c->acc, c->map, c->npix, c->d_stats, c->stream));
    return DMANN_OK;
}

int dmann_get_stats(dmann_ctx *c, dmann_stats *st)
{
    if (!c || !st) return DMANN_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    FLUSH(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    unsigned long long h[kNumStats];
    HIPCHK(c, hipMemcpy(h, c->d_stats, sizeof h, hipMemcpyDeviceToHost));
    st->kept = (int64_t)h[kStatKept];
    st->subpixel = (int64_t)h[kStatSub];
    st->fallback = (int64_t)h[kStatFallback];
    st->extended = (int64_t)h[kStatExt];
    st->contributions = (int64_t)h[kStatContrib];
    st->large = (int64_t)h[kStatLarge];
    st->large_contributions = (int64_t)h[kStatLargeContrib];
    return DMANN_OK;
}

int64_t dmann_npix(const dmann_ctx *c) { return c ? c->npix : 0; }

void dmann_destroy(dmann_ctx *c)
{
    if (!c) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    drop_events(c);
    if (c->order_ev) (void)hipEventDestroy(c->order_ev);
    free_all(c);
    delete c;
}

const char *dmann_last_error(const dmann_ctx *c) { return c ? c->msg : g_create_msg; }

static void fill_devparams(const dmann_params *p, DevParams &P) { dmann::fill_devparams(p, P, 1 << 22); }

int64_t dmann_host_ring_check(const dmann_params *p, const dmann_particle *q)
{
    if (!p || !q) return -1;
    DevParams P;
    fill_devparams(p, P);
    DiscRec rec;
    bool bad = false;
    const int cls = prepass_one(P, *reinterpret_cast<const Particle *>(q), rec, bad);
    if (bad || cls != kClsMedium) return -2;
    const PGeo g = particle_geo(rec, P.nside, P.footprint_pad);
    const RingGeo G = ring_geo(P.nside);
    int64_t mism = 0;
    for (int gi = 0; gi < g.npart; ++gi) {
        const RingRec a = ring_setup(g, P.nside, gi + g.nnorth);
        const RingLite b = ring_lite(g, G, g.irmin + gi);
        // integer decisions exactly; FP32 terms to a few units of rounding
        // (the step is an FP32 quotient, the offset a fraction of a step)
        const bool fl = fabsf(a.A - b.A) <= 1e-7f * a.A && a.B == b.B &&
                        fabsf(a.step - b.step) <= 2.4e-7f * a.step &&
                        fabsf(a.offm - b.offm) <= 2.4e-7f * a.step;
        if (a.cnt != b.cnt || (a.cnt && a.col0 != b.col0) || !fl || (a.cnt && a.kc_rel != b.kc_rel) ||
            a.startpix != b.startpix || a.nr != b.nr)
            mism++;
    }
    return mism;
}

int64_t dmann_host_nest_cells(int nside, const double *theta, const double *phi, int64_t n, uint32_t *cells)
{
    if (nside < 1 || nside > kMaxNestNside || (nside & (nside - 1)) != 0 || n < 0 ||
        (n > 0 && (!theta || !phi || !cells)))
        return -1;
    for (int64_t i = 0; i < n; ++i) cells[i] = nest_cell(nside, theta[i], phi[i]);
    return n;
}

int64_t dmann_host_sky_check(const dmann_params *p, const dmann_particle *q)
{
    if (!p || !q || p->nside < 1 || p->nside > kMaxNside) return -1;
    DevParams P;
    fill_devparams(p, P);
    DiscRec rec;
    bool bad = false;
    const int cls = prepass_one(P, *reinterpret_cast<const Particle *>(q), rec, bad);
    if (bad || cls != kClsLarge) return -2;
    const PGeo g = particle_geo(rec, P.nside, P.footprint_pad);
    // the per-particle ring walk: weight of every footprint pixel
    std::vector<float> wmap((size_t)P.npix, -1.0f);
    int64_t mism = 0, k = 0;
    for (int gi = 0; gi < g.nrings; ++gi) {
        const RingRec rr = ring_setup(g, P.nside, gi);
        int abase, kw;
        ring_addr(rr, (int)k, abase, kw);
        for (int t = 0; t < rr.cnt; ++t, ++k) {
            const int px = ring_pixel(abase, kw, rr.nr, (int)k);
            if (wmap[px] >= 0.0f) mism++; // a pixel twice
            wmap[px] = ring_weight32(g, rr, t);
        }
    }
    // the gather: rings of the range, segments, slots
    int rlo, rhi;
    sky_ring_range(g, P.nside, rlo, rhi);
    int64_t seen = 0;
    for (int r = rlo; r <= rhi; ++r) {
        int startpix, nr, shifted;
        ring_info(P.nside, r, startpix, nr, shifted);
        const RingRec rr = ring_setup(g, P.nside, ring_index(g, r));
        if (rr.startpix != startpix) {
            mism++;
            continue;
        }
        int c0 = rr.col0 % nr;
        if (c0 < 0) c0 += nr;
        for (int s0 = 0; s0 < nr; s0 += kSkySeg) {
            const int ws = nr - s0 < kSkySeg ? nr - s0 : kSkySeg;
            const SegSpan sp = seg_span(c0, rr.cnt, nr, s0, ws);
            for (int i = 0; i < sp.na + sp.nb; ++i) {
                const int t = i < sp.na ? sp.u + i : i - sp.na;
                const int slot = seg_slot(c0, t, nr, s0);
                if (t < 0 || t >= rr.cnt || slot < 0 || slot >= ws) {
                    mism++;
                    continue;
                }
                const int px = startpix + s0 + slot;
                const float w = ring_weight32(g, rr, t);
                if (wmap[px] != w) mism++; // not in the footprint, or another weight
                wmap[px] = -2.0f;          // taken
                seen++;
            }
        }
    }
    return mism + (k - seen);
}

int64_t dmann_host_footprint(const dmann_params *p, const dmann_particle *q, int32_t *pix, float *w, int64_t cap,
                             double *flux, int *kind)
{
    // (the halo-core plug-in's table lives in a context on the device)
    if (!p || !q || p->nside < 1 || p->nside > kMaxNside || p->flux_kind < DMANN_FLUX_ANN_NOSOM ||
        p->flux_kind > DMANN_FLUX_ANN_ADIABATIC)
        return -1;
    DevParams P;
    fill_devparams(p, P);
    DiscRec rec;
    bool bad = false;
    const int cls = prepass_one(P, *reinterpret_cast<const Particle *>(q), rec, bad);
    if (kind) *kind = 0;
    if (bad) return -2;
    if (cls == kClsSkip) return 0;
    if (flux) *flux = rec.flux;
    if (cls == kClsPoint) {
        if (kind) *kind = 1;
        if (cap > 0 && pix) pix[0] = rec.p0;
        if (cap > 0 && w) w[0] = 1.0f;
        return 1;
    }
    const PGeo g = particle_geo(rec, P.nside, P.footprint_pad);
    int64_t k = 0;
    if (cls == kClsMedium) {
        // the grouped kernel's per-ring terms and weights
        const RingGeo G = ring_geo(P.nside);
        for (int iz = g.irmin; iz <= g.irmax; ++iz) {
            const RingLite rl = ring_lite(g, G, iz);
            for (int t = 0; t < rl.cnt; ++t, ++k) {
                if (k < cap) {
                    int col = rl.col0 + t;
                    if (col < 0) col += rl.nr;
                    else if (col >= rl.nr) col -= rl.nr;
                    if (pix) pix[k] = rl.startpix + col;
                    if (w) w[k] = lite_weight32(g, rl, t);
                }
            }
        }
        if (kind) *kind = k < 2 ? 2 : 3;
        if (k < 2) {
            if (cap > 0 && pix) pix[0] = rec.p0;
            if (cap > 0 && w) w[0] = 1.0f;
        }
        return k;
    }
    for (int gi = 0; gi < g.nrings; ++gi) {
        const RingRec rr = ring_setup(g, P.nside, gi);
        int abase, kw;
        ring_addr(rr, (int)k, abase, kw);
        for (int t = 0; t < rr.cnt; ++t, ++k) {
            if (k < cap) {
                if (pix) pix[k] = ring_pixel(abase, kw, rr.nr, (int)k);
                if (w) w[k] = ring_weight32(g, rr, t);
            }
        }
    }
    if (kind) *kind = k < 2 ? 2 : 3;
    if (k < 2) {
        if (cap > 0 && pix) pix[0] = rec.p0;
        if (cap > 0 && w) w[0] = 1.0f;
    }
    return k;
}

} // extern "C"
