// HIP kernels of the MI355X annihilation-map projector (gfx950, wave64).
//
// Pipeline for one render call (all on the context's stream):
//   k_prepass      one thread per particle, FP64: filter, observer-frame
//                  distance, flux, angles, angular radius, RING pixel; the
//                  sub-pixel rule deposits whole particles; extended
//                  particles get a 48-byte record in a medium or a
//                  sky-covering (large) class list.
//   k_disc_medium  one wavefront per particle: FP64 query_disc ring ranges
//                  (one ring per lane) staged in LDS, Gaussian weights in
//                  FP32 from a haversine form, normalisation, deposit.
//   k_disc_large   one 256-thread workgroup per sky-covering particle,
//                  ring-major, FP64 normalisation sum.
//   k_finalize     map / dOmega into the FP64 RING output.
//
// Reference semantics followed: src/genmap.cpp:22-175 (CPU generator).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dmann_device.h"
#include "dmann_footprint.h"
#include "dmann_geom.h"
#include "dmann_kernels.h"

namespace dmann {

// size classes of the sky-covering kernel's largest-first order
constexpr int kLargeClasses = 10;
// particles per thread per pre-pass iteration
constexpr int kPrepassItems = 2;

// ---------------------------------------------------------------- pre-pass

// One thread per particle (grid-stride).  Mirrors src/genmap.cpp:82-110 with
// the helpers of src/anglefuncs.cpp:15-71 and src/flux/flux_ann_nosom.cpp:12-23.
// In the reproducible form (kAccFix) each tile of 512 particles places its
// records at the offsets tile_med / tile_large hold (exclusive prefix sums
// of k_prepass_count's per-tile counts), so the record order follows the
// particle order whatever the blocks' timing.
template <int ACC>
__global__ __launch_bounds__(256) void k_prepass(const Particle *__restrict__ parts, int64_t n,
                                                 DevParams P, DiscRec *__restrict__ recs,
                                                 unsigned *__restrict__ counts, MapRef map,
                                                 unsigned long long *__restrict__ stats,
                                                 const unsigned *__restrict__ tile_med,
                                                 const unsigned *__restrict__ tile_large)
{
    // Each block takes kPrepassItems x 256 particles per iteration (item j
    // of thread t is particle base + j 256 + t, so loads stay coalesced)
    // and appends its medium and sky-covering records with one global
    // atomic each: the two append counters are single addresses, and one
    // atomic per wave and 64 particles made them the bottleneck.
    constexpr int J = kPrepassItems;
    constexpr bool DET = ACC == kAccFix;
    __shared__ unsigned s_cnt[2][4], s_base[2];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x * J;
    unsigned long long n_kept = 0, n_sub = 0, n_bad = 0;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x * J; base < n; base += stride) {
        DiscRec rec[J];
        int cls[J];
        uint64_t mm[J], ml[J];
        unsigned tm = 0, tl = 0;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int64_t i = base + (int64_t)j * blockDim.x + threadIdx.x;
            cls[j] = kClsSkip;
            rec[j].p0 = 0;
            rec[j].flux = 0.0;
            if (i < n) {
                const Particle q = parts[i];
                bool bad;
                const int c = prepass_one(P, q, rec[j], bad);
                if (bad) n_bad++;
                if (c != kClsSkip) n_kept++;
                if (c == kClsPoint) n_sub++;
                cls[j] = c;
            }
            // sub-pixel particles: the whole flux to their pixel, reduced
            // over the wavefront by pixel when two or more lanes have one
            const bool sub = cls[j] == kClsPoint;
            const uint64_t ms = __ballot(sub);
            if (P.subpixel_sort && __popcll(ms) >= 2) wave_subpixel_deposit<ACC>(map, sub, rec[j].p0, rec[j].flux);
            else if (sub) acc_point<ACC>(map, rec[j].p0, rec[j].flux);
            mm[j] = __ballot(cls[j] == kClsMedium);
            ml[j] = __ballot(cls[j] == kClsLarge);
            tm += (unsigned)__popcll(mm[j]);
            tl += (unsigned)__popcll(ml[j]);
        }
        if (lane == 0) {
            s_cnt[0][wid] = tm;
            s_cnt[1][wid] = tl;
        }
        __syncthreads();
        if (threadIdx.x < 2) {
            const unsigned t = s_cnt[threadIdx.x][0] + s_cnt[threadIdx.x][1] + s_cnt[threadIdx.x][2] +
                               s_cnt[threadIdx.x][3];
            if constexpr (DET) {
                // the tile's place in particle order; the counters keep the totals
                const int64_t tile = base / ((int64_t)blockDim.x * J);
                s_base[threadIdx.x] = (threadIdx.x == 0 ? tile_med : tile_large)[tile];
                if (t) atomicAdd(&counts[threadIdx.x], t);
            } else {
                s_base[threadIdx.x] = t ? atomicAdd(&counts[threadIdx.x], t) : 0u;
            }
        }
        __syncthreads();
        unsigned om = s_base[0], ol = s_base[1];
        for (int w = 0; w < wid; ++w) {
            om += s_cnt[0][w];
            ol += s_cnt[1][w];
        }
#pragma unroll
        for (int j = 0; j < J; ++j) {
            if (cls[j] == kClsMedium) recs[om + (unsigned)__popcll(mm[j] & below)] = rec[j];
            if (cls[j] == kClsLarge) recs[n - 1 - (int64_t)(ol + (unsigned)__popcll(ml[j] & below))] = rec[j];
            om += (unsigned)__popcll(mm[j]);
            ol += (unsigned)__popcll(ml[j]);
        }
        __syncthreads(); // s_cnt / s_base are rewritten by the next iteration
    }
    wave_count(&stats[kStatKept], n_kept);
    wave_count(&stats[kStatSub], n_sub);
    wave_count(&stats[kStatBad], n_bad);
}

// Records per tile of the pre-pass (the reproducible form's first pass):
// the same classification, no deposits; tile_med[t] / tile_large[t] = the
// medium / sky-covering records of particles t 512 .. t 512 + 511.
__global__ __launch_bounds__(256) void k_prepass_count(const Particle *__restrict__ parts, int64_t n, DevParams P,
                                                       unsigned *__restrict__ tile_med,
                                                       unsigned *__restrict__ tile_large)
{
    constexpr int J = kPrepassItems;
    __shared__ unsigned s_cnt[2][4];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x * J;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x * J; base < n; base += stride) {
        unsigned tm = 0, tl = 0;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int64_t i = base + (int64_t)j * blockDim.x + threadIdx.x;
            int c = kClsSkip;
            if (i < n) {
                DiscRec rec;
                bool bad;
                c = prepass_one(P, parts[i], rec, bad);
            }
            tm += (unsigned)__popcll(__ballot(c == kClsMedium));
            tl += (unsigned)__popcll(__ballot(c == kClsLarge));
        }
        if (lane == 0) {
            s_cnt[0][wid] = tm;
            s_cnt[1][wid] = tl;
        }
        __syncthreads();
        if (threadIdx.x < 2) {
            const unsigned t = s_cnt[threadIdx.x][0] + s_cnt[threadIdx.x][1] + s_cnt[threadIdx.x][2] +
                               s_cnt[threadIdx.x][3];
            (threadIdx.x == 0 ? tile_med : tile_large)[base / ((int64_t)blockDim.x * J)] = t;
        }
        __syncthreads();
    }
}

// Sum of the kept particles' fluxes: per block in a fixed order (each
// thread over its grid-stride particles, then a tree over the block), then
// over the blocks in index order, so the total is the same on every run.
__global__ __launch_bounds__(256) void k_flux_part(const Particle *__restrict__ parts, int64_t n, DevParams P,
                                                   double *__restrict__ part)
{
    __shared__ double s[256];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double dx, dy, dz, d, flux;
        if (kept_flux(P, parts[i], dx, dy, dz, d, flux) && flux > 0.0) acc += flux;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void k_flux_sum(const double *__restrict__ part, int nb, double *__restrict__ total)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double t = 0.0;
        for (int b = 0; b < nb; ++b) t += part[b];
        *total = t;
    }
}

__global__ void k_fix_shift(unsigned long long *fix, int64_t npix, int k)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += (int64_t)gridDim.x * blockDim.x)
        fix[i] = (unsigned long long)((long long)fix[i] >> k);
}

// ---------------------------------------------------------------- footprint

// FP64 weight exactly as the reference forms it (src/genmap.cpp:133-143):
// centre vector, acos of the dot product, exp(-0.5 d2 / 0.333).
__device__ __forceinline__ double weight64(double z, double s, double phic, const PGeo &g,
                                           double vx, double vy)
{
    const double ux = s * cos(phic), uy = s * sin(phic);
    double d2 = acos(ux * vx + uy * vy + z * g.z0) / g.r;
    d2 = d2 * d2;
    return exp(-0.5 * d2 / 0.333);
}

// LDS layout of one ring window (64 rings) for one wave.
struct WinLDS {
    int pre_end[64];
    int kref[64];     // pre_start + kc_rel
    float4 f[64];     // A, B, offm, step
    float4 gq[64];    // E, Ac, D (general weight form)
    int4 a[64];       // abase, k_wrap, nr, kcol (= pre_start - col0)
    double4 d[64];    // z, s, shift, 2pi/nr  (FP64 form only)
};

// Sets up rings [win*64, win*64+64) of the particle; returns the window's
// pixel count.  Lane j handles ring win*64+j.
template <bool F64>
__device__ __forceinline__ int setup_window(const PGeo &g, int nside, int win, WinLDS &L, int lane,
                                            int kbase)
{
    const int gi = win * 64 + lane;
    RingRec rr;
    int cnt = 0;
    if (gi < g.nrings) {
        rr = ring_setup(g, nside, gi);
        cnt = rr.cnt;
    }
    const int incl = wave_incl_scan(cnt, lane);
    const int total = __shfl(incl, 63, 64);
    const int pre_start = kbase + incl - cnt;
    L.pre_end[lane] = kbase + incl;
    if (gi < g.nrings) {
        L.kref[lane] = pre_start + rr.kc_rel;
        L.f[lane] = make_float4(rr.A, rr.B, rr.offm, rr.step);
        L.gq[lane] = make_float4(rr.E, rr.Ac, rr.D, 0.0f);
        int abase, kw;
        ring_addr(rr, pre_start, abase, kw);
        L.a[lane] = make_int4(abase, kw, rr.nr, pre_start - rr.col0);
        if constexpr (F64) L.d[lane] = make_double4(rr.z, rr.s, rr.shift, kTwoPi / (double)rr.nr);
    }
    wave_sync();
    return total;
}

__device__ __forceinline__ int pix_addr(const int4 a, int k) { return ring_pixel(a.x, a.y, a.z, k); }

// FP32 weight of flat pixel k in ring j of the window (see weight32).
template <int MODE>
__device__ __forceinline__ float wt32(float x, const float4 &f, const WinLDS &L, int j, float K4)
{
    if constexpr (MODE == 0) {
        return weight32<0>(x, f.x, f.y, 0.0f, 0.0f, K4);
    } else if constexpr (MODE == 2) {
        return weight32<2>(x, f.x, f.y, 0.0f, 0.0f, K4);
    } else {
        const float4 q = L.gq[j];
        return weight32<1>(x, f.x, q.x, q.y, q.z, K4);
    }
}

// Register capacity of the single-evaluation path: footprints of up to
// 64*kNW pixels keep their weights in VGPRs between the norm and the deposit.
constexpr int kNW = 24;

// Renders one particle with one wavefront.  Returns the footprint size.
template <int ACC, int MODE>
__device__ int render_particle_wave(const PGeo &g, int nside, WinLDS &L, int lane, const MapRef &map)
{
    const int nwin = (g.nrings + 63) / 64;
    if constexpr (ACC != kAccF64) {
        if (nwin == 1) {
            const int T = setup_window<false>(g, nside, 0, L, lane, 0);
            if (T < 2) {
                if (lane == 0) acc_point<ACC>(map, g.p0, g.flux);
                wave_sync();
                return T;
            }
            if (T <= 64 * kNW) {
                float wv[kNW];
                int av[kNW];
                float nsum = 0.0f;
                int j = 0;
#pragma unroll
                for (int it = 0; it < kNW; ++it) {
                    const int k = lane + 64 * it;
                    wv[it] = 0.0f;
                    av[it] = 0;
                    if (k < T) {
                        while (L.pre_end[j] <= k) ++j;
                        const float4 f = L.f[j];
                        const float kf = (float)(k - L.kref[j]);
                        const float w = wt32<MODE>(fmaf(kf, f.w, f.z), f, L, j, g.K4);
                        wv[it] = w;
                        av[it] = pix_addr(L.a[j], k);
                        nsum += w;
                    }
                }
                const double norm = wave_sum((double)nsum);
                const float scale = (float)(g.flux / norm);
#pragma unroll
                for (int it = 0; it < kNW; ++it) {
                    const int k = lane + 64 * it;
                    if (k < T) acc_add<ACC>(map, av[it], wv[it] * scale, 0.0);
                }
                wave_sync();
                return T;
            }
            wave_sync();
        }
    }
    // two passes over ring windows: normalisation, then deposit
    double nacc = 0.0;
    int T = 0;
    const double vx = sin(g.theta) * cos(g.phi), vy = sin(g.theta) * sin(g.phi);
    for (int win = 0; win < nwin; ++win) {
        const int Tw = setup_window<ACC == kAccF64>(g, nside, win, L, lane, 0);
        int j = 0;
        if constexpr (ACC == kAccF64) {
            for (int k = lane; k < Tw; k += 64) {
                while (L.pre_end[j] <= k) ++j;
                const double4 dd = L.d[j];
                const int col = k - L.a[j].w;
                nacc += weight64(dd.x, dd.y, ((double)col + dd.z) * dd.w, g, vx, vy);
            }
        } else {
            float part = 0.0f;
            int c = 0;
            for (int k = lane; k < Tw; k += 64) {
                while (L.pre_end[j] <= k) ++j;
                const float4 f = L.f[j];
                const float kf = (float)(k - L.kref[j]);
                part += wt32<MODE>(fmaf(kf, f.w, f.z), f, L, j, g.K4);
                if (++c == 16) {
                    nacc += (double)part;
                    part = 0.0f;
                    c = 0;
                }
            }
            nacc += (double)part;
        }
        T += Tw;
        wave_sync();
    }
    if (T < 2) {
        if (lane == 0) acc_point<ACC>(map, g.p0, g.flux);
        wave_sync();
        return T;
    }
    const double norm = wave_sum(nacc);
    const double scale64 = g.flux / norm;
    const float scale = (float)scale64;
    for (int win = 0; win < nwin; ++win) {
        const int Tw = setup_window<ACC == kAccF64>(g, nside, win, L, lane, 0);
        int j = 0;
        for (int k = lane; k < Tw; k += 64) {
            while (L.pre_end[j] <= k) ++j;
            const int addr = pix_addr(L.a[j], k);
            if constexpr (ACC == kAccF64) {
                const double4 dd = L.d[j];
                const int col = k - L.a[j].w;
                const double w = weight64(dd.x, dd.y, ((double)col + dd.z) * dd.w, g, vx, vy);
                acc_add<ACC>(map, addr, 0.0f, (w / norm) * g.flux);
            } else {
                const float4 f = L.f[j];
                const float kf = (float)(k - L.kref[j]);
                const float w = wt32<MODE>(fmaf(kf, f.w, f.z), f, L, j, g.K4);
                acc_add<ACC>(map, addr, w * scale, 0.0);
            }
        }
        wave_sync();
    }
    (void)scale64;
    return T;
}

// One wavefront per particle of the medium class (grid-stride over waves).
template <int ACC>
__global__ __launch_bounds__(256) void k_disc_medium(const DiscRec *__restrict__ recs,
                                                     const unsigned *__restrict__ counts, int nside, double pad,
                                                     MapRef map, unsigned long long *__restrict__ stats)
{
    __shared__ WinLDS lds[4];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    WinLDS &L = lds[wid];
    const unsigned count = counts[0];
    const unsigned nwaves = gridDim.x * 4;
    unsigned long long n_ext = 0, n_fb = 0, n_contrib = 0;
    for (unsigned it = blockIdx.x * 4 + wid; it < count; it += nwaves) {
        const PGeo g = particle_geo(recs[it], nside, pad);
        // the pre-pass only puts fast-form particles in this class
        const int T = render_particle_wave<ACC, 0>(g, nside, L, lane, map);
        if (T < 2) n_fb++;
        else {
            n_ext++;
            n_contrib += (unsigned long long)T;
        }
    }
    if (lane == 0) {
        if (n_ext) atomicAdd(&stats[kStatExt], n_ext);
        if (n_fb) atomicAdd(&stats[kStatFallback], n_fb);
        if (n_contrib) atomicAdd(&stats[kStatContrib], n_contrib);
    }
}

// Sky-covering particles: one 256-thread workgroup per particle; each wave
// walks whole rings (ring-major) from 64-ring windows, FP64 normalisation.
template <int ACC>
__global__ __launch_bounds__(256) void k_disc_large(const DiscRec *__restrict__ recs, int64_t n,
                                                    unsigned *__restrict__ counts, int nside, double pad,
                                                    MapRef map, unsigned long long *__restrict__ stats)
{
    __shared__ WinLDS lds[4];
    __shared__ double s_red[4];
    __shared__ int s_cnt[4];
    __shared__ unsigned s_g, s_list[256];
    __shared__ int s_wn[4];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    WinLDS &L = lds[wid];
    const unsigned count = counts[1];
    // Footprints here range from a few thousand pixels to the whole sky, so
    // blocks take work from a counter (counts[2], zeroed with the class
    // counters), largest footprints first: work item g is chunk g % nchunk
    // (256 particles) of size class g / nchunk (floor(log2 est) below the
    // map's, the last class taking the rest).  Each block compacts the
    // chunk's particles of that class and renders them in turn.
    const unsigned nchunk = (count + 255u) / 256u;
    const int lgmax = 63 - __clzll((unsigned long long)npix_of(nside));
    for (;;) {
        if (threadIdx.x == 0) s_g = atomicAdd(&counts[2], 1u);
        __syncthreads();
        const unsigned g = s_g;
        if (g >= nchunk * (unsigned)kLargeClasses) break;
        const int want = (int)(g / nchunk);
        const unsigned item = (g % nchunk) * 256u + threadIdx.x;
        bool mine = false;
        if (item < count) {
            const int est = recs[n - 1 - (int64_t)item].est;
            const int lg = est > 0 ? 31 - __clz(est) : 0;
            int c = lgmax - lg;
            c = c < 0 ? 0 : (c >= kLargeClasses ? kLargeClasses - 1 : c);
            mine = c == want;
        }
        const uint64_t m = __ballot(mine);
        if (lane == 0) s_wn[wid] = __popcll(m);
        __syncthreads();
        int base = 0;
        for (int w = 0; w < wid; ++w) base += s_wn[w];
        if (mine) s_list[base + __popcll(m & ((1ull << lane) - 1ull))] = item;
        const int nmine = s_wn[0] + s_wn[1] + s_wn[2] + s_wn[3];
        __syncthreads();
        for (int l = 0; l < nmine; ++l) {
            const unsigned it = s_list[l];
            const PGeo g = particle_geo(recs[n - 1 - (int64_t)it], nside, pad);
            const int nwin = (g.nrings + 63) / 64;
            const double vx = sin(g.theta) * cos(g.phi), vy = sin(g.theta) * sin(g.phi);
            // the particle's windows are walked from a window that depends
            // on the particle, so blocks rendering overlapping sky-covering
            // footprints do not sweep the same rings' pixels with atomics
            // at the same time
            const int rot = (int)((it * 2654435761u) >> 8) % nwin;
            double scale64 = 0.0;
            for (int pass = 0; pass < 2; ++pass) {
                double nacc = 0.0;
                int T = 0;
                for (int k = wid; k < nwin; k += 4) {
                    const int win = k + rot < nwin ? k + rot : k + rot - nwin;
                    const int Tw = setup_window<ACC == kAccF64>(g, nside, win, L, lane, 0);
                    T += Tw;
                    for (int j = 0; j < 64; ++j) {
                        const int pre_start = j ? L.pre_end[j - 1] : 0;
                        const int cnt = L.pre_end[j] - pre_start;
                        if (cnt <= 0) continue;
                        const float4 f = L.f[j];
                        const float4 gq = L.gq[j];
                        const int kref = L.kref[j];
                        const int4 a = L.a[j];
                        const double4 dd = L.d[j];
                        float part = 0.0f;
                        for (int t = lane; t < cnt; t += 64) {
                            const int k = pre_start + t;
                            if constexpr (ACC == kAccF64) {
                                const double w = weight64(dd.x, dd.y, ((double)(k - a.w) + dd.z) * dd.w, g, vx, vy);
                                if (pass == 0) nacc += w;
                                else acc_add<ACC>(map, pix_addr(a, k), 0.0f, (w * scale64));
                            } else {
                                const float kf = (float)(k - kref);
                                float w;
                                if (g.mode == 0) w = weight32<0>(fmaf(kf, f.w, f.z), f.x, f.y, 0.0f, 0.0f, g.K4);
                                else if (g.mode == 2) w = weight32<2>(fmaf(kf, f.w, f.z), f.x, f.y, 0.0f, 0.0f, g.K4);
                                else w = weight32<1>(fmaf(kf, f.w, f.z), f.x, gq.x, gq.y, gq.z, g.K4);
                                if (pass == 0) part += w;
                                else acc_add<ACC>(map, pix_addr(a, k), w * (float)scale64, 0.0);
                            }
                        }
                        if constexpr (ACC != kAccF64) nacc += (double)part;
                    }
                    wave_sync();
                }
                if (pass == 0) {
                    const double ws = wave_sum(nacc);
                    if (lane == 0) {
                        s_red[wid] = ws;
                        s_cnt[wid] = T;
                    }
                    __syncthreads();
                    const double norm = s_red[0] + s_red[1] + s_red[2] + s_red[3];
                    const int Ttot = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
                    __syncthreads();
                    if (Ttot < 2) {
                        if (threadIdx.x == 0) {
                            acc_point<ACC>(map, g.p0, g.flux);
                            atomicAdd(&stats[kStatFallback], 1ull);
                        }
                        break;
                    }
                    if (ACC == kAccF64) scale64 = g.flux / norm;
                    else scale64 = g.flux / norm;
                    if (threadIdx.x == 0) {
                        atomicAdd(&stats[kStatExt], 1ull);
                        atomicAdd(&stats[kStatLarge], 1ull);
                        atomicAdd(&stats[kStatContrib], (unsigned long long)Ttot);
                        atomicAdd(&stats[kStatLargeContrib], (unsigned long long)Ttot);
                    }
                }
            }
            __syncthreads();
        }
    }
}


// ---------------------------------------------------------------- adjoint

// Gradient of a loss L(map) with respect to each particle's mass, density,
// smoothing length and position, given g = dL/dmap (FP64, the finished map,
// i.e. after the division by dOmega).  The footprint is recomputed with the
// forward's ring walk (membership is piecewise constant, so only the smooth
// dependence through flux, r and the direction is differentiated), weights
// and their derivatives in FP64:
//   w_p = exp(-c psi_p^2 / r^2), c = 0.5/0.333,
//   dw_p/dr = w_p 2 c psi_p^2 / r^3,
//   dw_p/dv = w_p 2 c psi_p / (r^2 sin psi_p) (u_p - cos psi_p v)
// with v the particle direction and u_p the pixel centre.  Per particle the
// wave sums S = sum g w, N = sum w and the r- and v-derivatives of both;
// then dL/dflux = S/N, dL/dr = flux (S_r N - S N_r)/N^2,
// dL/dv = flux (S_v N - S N_v)/N^2, chained to (m, rho, h, pos) through
// flux = K rho m / (4 pi d^2), r = h/d, v = diag(-1,-1,1) R^T (pos-opos)/d.
// One wavefront per particle; out[6 i + {0..5}] = d/d(mass, dens, hsmooth,
// posx, posy, posz).
__global__ __launch_bounds__(256) void k_adjoint(const Particle *__restrict__ parts, int64_t n, DevParams P,
                                                 const double *__restrict__ gmap, double *__restrict__ out)
{
    __shared__ WinLDS lds[4];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    WinLDS &L = lds[wid];
    const double c = 0.5 / 0.333;
    const double inv_dO = 1.0 / P.dOmega;
    for (int64_t i = (int64_t)blockIdx.x * 4 + wid; i < n; i += (int64_t)gridDim.x * 4) {
        const Particle q = parts[i];
        DiscRec rec;
        bool bad;
        const int cls = prepass_one(P, q, rec, bad);
        double res[6] = {0, 0, 0, 0, 0, 0};
        if (cls != kClsSkip) {
            const double dx = (double)q.posx - P.opos[0];
            const double dy = (double)q.posy - P.opos[1];
            const double dz = (double)q.posz - P.opos[2];
            const double d = sqrt(dx * dx + dy * dy + dz * dz);
            const double flux = rec.flux;
            const double h = (double)q.hsmooth;
            double S = 0, N = 0, Sr = 0, Nr = 0, Sv[3] = {0, 0, 0}, Nv[3] = {0, 0, 0};
            int T = 0;
            double r = rec.r;
            if (cls != kClsPoint) {
                const PGeo g = particle_geo(rec, P.nside, P.footprint_pad);
                const double st = sin(g.theta);
                const double v[3] = {st * cos(g.phi), st * sin(g.phi), g.z0};
                const int nwin = (g.nrings + 63) / 64;
                for (int win = 0; win < nwin; ++win) {
                    const int Tw = setup_window<true>(g, P.nside, win, L, lane, 0);
                    // the gradient map value of the lane's next pixel is
                    // loaded while this one is computed (the loop is bound
                    // by that load's latency, not by its FP64 work)
                    int j = 0;
                    if (lane < Tw)
                        while (L.pre_end[j] <= lane) ++j;
                    double gp_next = lane < Tw ? gmap[pix_addr(L.a[j], lane)] : 0.0;
                    for (int k = lane; k < Tw; k += 64) {
                        const double gp_k = gp_next;
                        const int jk = j;
                        if (k + 64 < Tw) {
                            while (L.pre_end[j] <= k + 64) ++j;
                            gp_next = gmap[pix_addr(L.a[j], k + 64)];
                        }
                        const double4 dd = L.d[jk];
                        const int4 a = L.a[jk];
                        const double phic = ((double)(k - a.w) + dd.z) * dd.w;
                        double sp, cp;
                        sincos(phic, &sp, &cp);
                        const double u[3] = {dd.y * cp, dd.y * sp, dd.x};
                        const double cosp = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
                        const double cx = u[1] * v[2] - u[2] * v[1];
                        const double cy = u[2] * v[0] - u[0] * v[2];
                        const double cz = u[0] * v[1] - u[1] * v[0];
                        const double sinp = sqrt(cx * cx + cy * cy + cz * cz);
                        const double psi = atan2(sinp, cosp);
                        const double w = exp(-c * psi * psi / (r * r));
                        const double gp = gp_k * inv_dO;
                        const double dwdr = w * 2.0 * c * psi * psi / (r * r * r);
                        const double ratio = sinp > 1e-300 ? psi / sinp : 1.0;
                        const double fv = w * 2.0 * c * ratio / (r * r);
                        S += gp * w;
                        N += w;
                        Sr += gp * dwdr;
                        Nr += dwdr;
#pragma unroll
                        for (int e = 0; e < 3; e++) {
                            const double dv = fv * (u[e] - cosp * v[e]);
                            Sv[e] += gp * dv;
                            Nv[e] += dv;
                        }
                    }
                    T += Tw;
                    wave_sync();
                }
                S = wave_sum(S);
                N = wave_sum(N);
                Sr = wave_sum(Sr);
                Nr = wave_sum(Nr);
#pragma unroll
                for (int e = 0; e < 3; e++) {
                    Sv[e] = wave_sum(Sv[e]);
                    Nv[e] = wave_sum(Nv[e]);
                }
            }
            double dLdflux, dLdr = 0.0, dLdv[3] = {0, 0, 0};
            if (cls == kClsPoint || T < 2) {
                dLdflux = gmap[rec.p0] * inv_dO;
            } else {
                dLdflux = S / N;
                dLdr = flux * (Sr * N - S * Nr) / (N * N);
                for (int e = 0; e < 3; e++) dLdv[e] = flux * (Sv[e] * N - S * Nv[e]) / (N * N);
            }
            // every plug-in is linear in the mass, the annihilation ones in
            // the density too (the decay one does not depend on it)
            Particle q1 = q;
            q1.mass = 1.0f;
            res[0] = dLdflux * plugin_flux(P, q1, d);
            q1 = q;
            q1.dens = 1.0f;
            res[1] = P.flux_kind == DMANN_FLUX_DECAY ? 0.0 : dLdflux * plugin_flux(P, q1, d);
            res[2] = dLdr / d;
            const double dLdd = dLdflux * (-2.0 * flux / d) + dLdr * (-h / (d * d));
            // R D dL/dv, D = diag(-1, -1, 1); R row-major
            const double wv[3] = {-dLdv[0], -dLdv[1], dLdv[2]};
            double rw[3];
            for (int a = 0; a < 3; a++) rw[a] = P.rot[3 * a] * wv[0] + P.rot[3 * a + 1] * wv[1] + P.rot[3 * a + 2] * wv[2];
            const double dh[3] = {dx / d, dy / d, dz / d};
            const double proj = rw[0] * dh[0] + rw[1] * dh[1] + rw[2] * dh[2];
            for (int a = 0; a < 3; a++) res[3 + a] = (rw[a] - proj * dh[a]) / d + dLdd * dh[a];
        }
        if (lane < 6) out[6 * i + lane] = res[lane];
    }
}

hipError_t launch_adjoint(const Particle *d_parts, int64_t n, const DevParams &P, const double *gmap, double *out,
                          int blocks, hipStream_t st)
{
    const int64_t want = (n + 3) / 4;
    const int b = (int)(want < blocks ? want : blocks);
    if (b > 0) k_adjoint<<<b, 256, 0, st>>>(d_parts, n, P, gmap, out);
    return hipGetLastError();
}


// ---------------------------------------------------------------- batched

// Many maps (observers, each with its own nside) from one particle set in
// one persistent launch.  Work items are (map, chunk of `chunk` particles),
// taken from a device counter by whole wavefronts; for each particle the
// wave runs the pre-pass for that map's observer and renders the footprint
// (fast or general weight form).  Every wave leaves when the counter passes
// the last item.
template <int ACC>
__global__ __launch_bounds__(256) void k_batch(const Particle *__restrict__ parts, int64_t n,
                                               const DevParams *__restrict__ Ps, const MapRef *__restrict__ maps,
                                               int nmaps, int chunk, unsigned long long *__restrict__ work,
                                               unsigned long long *__restrict__ stats)
{
    __shared__ WinLDS lds[4];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    WinLDS &L = lds[wid];
    const unsigned long long nblk = (unsigned long long)((n + chunk - 1) / chunk);
    const unsigned long long total = nblk * (unsigned long long)nmaps;
    for (;;) {
        unsigned long long item = 0;
        if (lane == 0) item = atomicAdd(work, 1ull);
        item = __shfl(item, 0, 64);
        if (item >= total) break;
        const int m = (int)(item / nblk);
        const int64_t i0 = (int64_t)(item % nblk) * chunk;
        const int64_t i1 = i0 + chunk < n ? i0 + chunk : n;
        const DevParams P = Ps[m];
        const MapRef map = maps[m];
        unsigned long long c_kept = 0, c_sub = 0, c_fb = 0, c_ext = 0, c_contrib = 0, c_bad = 0;
        for (int64_t i = i0; i < i1; ++i) {
            const Particle q = parts[i];
            DiscRec rec;
            bool bad;
            const int cls = prepass_one(P, q, rec, bad);
            if (bad) c_bad++;
            if (cls == kClsSkip) continue;
            c_kept++;
            if (cls == kClsPoint) {
                if (lane == 0) acc_point<ACC>(map, rec.p0, rec.flux);
                c_sub++;
                continue;
            }
            const PGeo g = particle_geo(rec, P.nside, P.footprint_pad);
            const int T = g.mode == 0   ? render_particle_wave<ACC, 0>(g, P.nside, L, lane, map)
                          : g.mode == 2 ? render_particle_wave<ACC, 2>(g, P.nside, L, lane, map)
                                        : render_particle_wave<ACC, 1>(g, P.nside, L, lane, map);
            if (T < 2) c_fb++;
            else {
                c_ext++;
                c_contrib += (unsigned long long)T;
            }
        }
        if (lane == 0) {
            unsigned long long *st = stats + (size_t)m * kNumStats;
            if (c_kept) atomicAdd(&st[kStatKept], c_kept);
            if (c_sub) atomicAdd(&st[kStatSub], c_sub);
            if (c_fb) atomicAdd(&st[kStatFallback], c_fb);
            if (c_ext) atomicAdd(&st[kStatExt], c_ext);
            if (c_contrib) atomicAdd(&st[kStatContrib], c_contrib);
            if (c_bad) atomicAdd(&st[kStatBad], c_bad);
        }
    }
}

__global__ void k_fill_u64n(unsigned long long *p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

hipError_t launch_batch(int acc, const Particle *d_parts, int64_t n, const DevParams *Ps, const MapRef *maps,
                        int nmaps, int chunk, unsigned long long *work, unsigned long long *stats, int blocks,
                        hipStream_t st)
{
    k_fill_u64n<<<1, 64, 0, st>>>(work, 1);
    switch (acc) {
    case kAccF32: k_batch<kAccF32><<<blocks, 256, 0, st>>>(d_parts, n, Ps, maps, nmaps, chunk, work, stats); break;
    case kAccComp: k_batch<kAccComp><<<blocks, 256, 0, st>>>(d_parts, n, Ps, maps, nmaps, chunk, work, stats); break;
    default: k_batch<kAccF64><<<blocks, 256, 0, st>>>(d_parts, n, Ps, maps, nmaps, chunk, work, stats); break;
    }
    return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long *p, int64_t n, hipStream_t st)
{
    const int blocks = (int)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
    k_fill_u64n<<<blocks > 0 ? blocks : 1, 256, 0, st>>>(p, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------- map ops

__global__ void k_fill_f32(float *p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

__global__ void k_fill_f64(double *p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

__global__ void k_fill_u32(unsigned *p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

__global__ void k_fill_u64(unsigned long long *p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

// map / dOmega into FP64 RING output (src/genmap.cpp:164-169).
template <int ACC>
__global__ void k_finalize(MapRef map, int64_t npix, double dOmega, double *__restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npix;
         i += (int64_t)gridDim.x * blockDim.x) {
        double v;
        if constexpr (ACC == kAccF32) v = (double)map.f32[i];
        else if constexpr (ACC == kAccComp) v = (double)map.f32[i] + (double)map.lo[i];
        else if constexpr (ACC == kAccFix) v = (double)(long long)map.fix[i] / map.fix_scale;
        else v = map.f64[i]; // FP64 and W64 maps; in place when out == map.f64
        out[i] = v / dOmega;
    }
}

// ---------------------------------------------------------------- launchers

template <int ACC>
static hipError_t launch_render_t(const Particle *d_parts, int64_t n, const DevParams &P, DiscRec *recs,
                                  unsigned *counts, MapRef map, unsigned long long *stats,
                                  const LaunchCfg &cfg, hipStream_t st, hipEvent_t *ev, const GroupCfg *gc,
                                  const DetCfg *det)
{
    constexpr bool DET = ACC == kAccFix;
    constexpr bool SERIES = ACC == kAccF32 || ACC == kAccComp || ACC == kAccFix;
    if (DET && !det) return hipErrorInvalidValue;
    // groups of the reproducible form come from the stable radix sort
    if (DET && gc && gc->k > 0 && !gc->sort_tmp) return hipErrorInvalidValue;
    k_fill_u32<<<1, 64, 0, st>>>(counts, 4);
    const int64_t want = (n + 256 * kPrepassItems - 1) / (256 * kPrepassItems);
    const int pre_blocks = (int)(want < cfg.prepass_blocks ? want : cfg.prepass_blocks);
    if (ev) (void)hipEventRecord(ev[0], st);
    const unsigned *tmed = nullptr, *tlarge = nullptr;
    if (DET && pre_blocks > 0) {
        // records placed in particle order: per-tile counts, their
        // exclusive prefix sums, then the pre-pass proper
        const int ntiles = (int)want;
        k_prepass_count<<<pre_blocks, 256, 0, st>>>(d_parts, n, P, det->tile_med, det->tile_large);
        hipError_t e;
        if ((e = launch_scan_u32(det->tile_med, ntiles, det->tile_sums, st)) != hipSuccess) return e;
        if ((e = launch_scan_u32(det->tile_large, ntiles, det->tile_sums, st)) != hipSuccess) return e;
        tmed = det->tile_med;
        tlarge = det->tile_large;
    }
    if (pre_blocks > 0)
        k_prepass<ACC><<<pre_blocks, 256, 0, st>>>(d_parts, n, P, recs, counts, map, stats, tmed, tlarge);
    if (ev) (void)hipEventRecord(ev[1], st);
    if (SERIES && cfg.sky_blocks > 0 && gc && gc->keys) {
        const hipError_t e = launch_sky(ACC, recs, n, counts, P.nside, P.footprint_pad, map, stats, gc->keys,
                                        cfg.large_blocks, cfg.sky_blocks, st);
        if (e != hipSuccess) return e;
    } else {
        k_disc_large<ACC><<<cfg.large_blocks, 256, 0, st>>>(recs, n, counts, P.nside, P.footprint_pad, map, stats);
    }
    if (ev) (void)hipEventRecord(ev[2], st);
    if (SERIES && gc && gc->k > 0) {
        const hipError_t e = launch_group(ACC, recs, counts, P.nside, P.footprint_pad, *gc, map, stats, st,
                                          ev ? ev[3] : nullptr, n);
        if (e != hipSuccess) return e;
    } else {
        if (ev) (void)hipEventRecord(ev[3], st);
        k_disc_medium<ACC><<<cfg.medium_blocks, 256, 0, st>>>(recs, counts, P.nside, P.footprint_pad, map, stats);
    }
    if (ev) (void)hipEventRecord(ev[4], st);
    return hipGetLastError();
}

hipError_t launch_render(int acc, const Particle *d_parts, int64_t n, const DevParams &P, DiscRec *recs,
                         unsigned *counts, MapRef map, unsigned long long *stats, const LaunchCfg &cfg,
                         hipStream_t st, hipEvent_t *ev, const GroupCfg *gc, const DetCfg *det)
{
    switch (acc) {
    case kAccF32: return launch_render_t<kAccF32>(d_parts, n, P, recs, counts, map, stats, cfg, st, ev, gc, nullptr);
    case kAccComp: return launch_render_t<kAccComp>(d_parts, n, P, recs, counts, map, stats, cfg, st, ev, gc, nullptr);
    case kAccW64:
        return launch_render_t<kAccW64>(d_parts, n, P, recs, counts, map, stats, cfg, st, ev, nullptr, nullptr);
    case kAccFix: return launch_render_t<kAccFix>(d_parts, n, P, recs, counts, map, stats, cfg, st, ev, gc, det);
    default: return launch_render_t<kAccF64>(d_parts, n, P, recs, counts, map, stats, cfg, st, ev, gc, nullptr);
    }
}

hipError_t launch_flux_total(const Particle *d_parts, int64_t n, const DevParams &P, const DetCfg &det, int blocks,
                             hipStream_t st)
{
    k_flux_part<<<blocks, 256, 0, st>>>(d_parts, n, P, det.flux_part);
    k_flux_sum<<<1, 64, 0, st>>>(det.flux_part, blocks, det.flux_total);
    return hipGetLastError();
}

hipError_t launch_fix_shift(unsigned long long *fix, int64_t npix, int k, hipStream_t st)
{
    const int blocks = (int)((npix + 255) / 256 < 8192 ? (npix + 255) / 256 : 8192);
    k_fix_shift<<<blocks, 256, 0, st>>>(fix, npix, k);
    return hipGetLastError();
}

hipError_t launch_finalize(int acc, MapRef map, int64_t npix, double dOmega, double *out, hipStream_t st)
{
    const int blocks = (int)((npix + 255) / 256 < 8192 ? (npix + 255) / 256 : 8192);
    switch (acc) {
    case kAccF32: k_finalize<kAccF32><<<blocks, 256, 0, st>>>(map, npix, dOmega, out); break;
    case kAccComp: k_finalize<kAccComp><<<blocks, 256, 0, st>>>(map, npix, dOmega, out); break;
    case kAccFix: k_finalize<kAccFix><<<blocks, 256, 0, st>>>(map, npix, dOmega, out); break;
    default: k_finalize<kAccF64><<<blocks, 256, 0, st>>>(map, npix, dOmega, out); break;
    }
    return hipGetLastError();
}

hipError_t launch_clear(int acc, MapRef map, int64_t npix, unsigned long long *stats, hipStream_t st)
{
    const int blocks = (int)((npix + 255) / 256 < 8192 ? (npix + 255) / 256 : 8192);
    if (acc == kAccF64 || acc == kAccW64) k_fill_f64<<<blocks, 256, 0, st>>>(map.f64, npix);
    else if (acc == kAccFix) k_fill_u64n<<<blocks, 256, 0, st>>>(map.fix, npix);
    else k_fill_f32<<<blocks, 256, 0, st>>>(map.f32, npix);
    if (acc == kAccComp) k_fill_f32<<<blocks, 256, 0, st>>>(map.lo, npix);
    k_fill_u64<<<1, 64, 0, st>>>(stats, kNumStats);
    return hipGetLastError();
}

} // namespace dmann
