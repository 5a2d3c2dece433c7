// Sky-covering particles (near the observer, over a pole, or wider than the
// medium class), gathered over pixels.
//
// Their footprints run from thousands of pixels to the whole sky and
// overlap heavily; deposited pixel by pixel they cost one global atomic per
// footprint pixel (two in the two-float form), all of them forwarded to
// memory.  Here the map is instead built ring by ring:
//   k_sky_norm    one 256-thread workgroup per particle, largest footprints
//                 first: the normalisation sum over the footprint (FP32
//                 weights, FP64 sum), the statistics and the < 2-pixel
//                 fallback (src/genmap.cpp:125-128).  The record's flux
//                 becomes flux / norm, and the particle's ring range goes
//                 to a packed 4-byte entry of its own (`ranges`), so the
//                 gather's scans read 4 bytes per particle.
//   k_sky_gather  one wavefront per HEALPix ring, owning that ring's
//                 pixels.  It scans the particles whose ring range holds
//                 the ring and takes them in batches of up to 64: each lane
//                 sets up one particle's column range and weight terms on
//                 the ring (the FP64 query_disc construction of ring_setup);
//                 then, one 2048-column segment of the ring at a time, the
//                 wave walks the batch's particles over their columns in
//                 the segment, adding w * (flux / norm) into an FP32 copy of
//                 the segment in LDS, and adds the segment into the map
//                 (two-float with an exact TwoSum in the compensated form)
//                 with plain loads and stores: no other wave writes the
//                 ring's pixels during the launch, so no atomics are needed.
// Each LDS slot sums at most 64 terms (one batch) before it goes into the
// map, so the FP32 partial sums keep a relative error below 64 * 2^-24.
// Weights are those of the per-particle sky-covering kernel (ring_setup,
// weight32), the reference's definition (src/genmap.cpp:119-153).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dmann_device.h"
#include "dmann_footprint.h"
#include "dmann_geom.h"
#include "dmann_kernels.h"

namespace dmann {

namespace {

constexpr int kSkyClasses = 10; // size classes of the largest-first order

// Ring range of a rendered particle packed into 32 bits (rings are below
// 2^16 for nside <= 8192); the empty range (1, 0) marks a fallback.
__device__ __forceinline__ int pack_rings(int lo, int hi) { return (int)((unsigned)lo | ((unsigned)hi << 16)); }

struct NormLDS {
    float4 f[4][64]; // A, B, offm, step
    float4 q[4][64]; // E, Ac, D, -
    int2 c[4][64];   // cnt, kc_rel
    double red[4];
    int tot[4];
    unsigned g, list[256];
    int wn[4];
};

// This lane's share of the weights of one ring row (cnt = c.x pixels, the
// column nearest the particle at c.y): four columns per lane at a time in
// four FP32 partial sums while the wave has 256 columns left (at most 32
// terms each for rings of 8192 pixels), then one at a time.  B2 is B for
// the series forms and E for the general form.
template <int MODE>
__device__ __forceinline__ double row_weights(const int2 &c, const float4 &f, float B2, const float4 &q, float K4,
                                              int lane)
{
    float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    int t0 = 0;
    for (; t0 + 256 <= c.x; t0 += 256) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int t = t0 + 64 * k + lane;
            part[k] += weight32<MODE>(fmaf((float)(t - c.y), f.w, f.z), f.x, B2, q.y, q.z, K4);
        }
    }
    for (int t = t0 + lane; t < c.x; t += 64)
        part[0] += weight32<MODE>(fmaf((float)(t - c.y), f.w, f.z), f.x, B2, q.y, q.z, K4);
    return (double)((part[0] + part[1]) + (part[2] + part[3]));
}

template <int ACC>
__global__ __launch_bounds__(256) void k_sky_norm(DiscRec *__restrict__ recs, int64_t n,
                                                  unsigned *__restrict__ counts, int nside, double pad, MapRef map,
                                                  unsigned long long *__restrict__ stats,
                                                  unsigned *__restrict__ ranges)
{
    __shared__ NormLDS S;
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const unsigned count = counts[1];
    // work item gi: chunk gi % nchunk (256 particles) of size class
    // gi / nchunk (floor(log2 est) below the map's, the last class taking
    // the rest), from the counter counts[2]; each block compacts the
    // chunk's particles of that class and renders them in turn
    const unsigned nchunk = (count + 255u) / 256u;
    const int lgmax = 63 - __clzll((unsigned long long)npix_of(nside));
    for (;;) {
        if (threadIdx.x == 0) S.g = atomicAdd(&counts[2], 1u);
        __syncthreads();
        const unsigned gi = S.g;
        if (gi >= nchunk * (unsigned)kSkyClasses) break;
        const int want = (int)(gi / nchunk);
        const unsigned item = (gi % nchunk) * 256u + threadIdx.x;
        bool mine = false;
        if (item < count) {
            const int est = recs[n - 1 - (int64_t)item].est;
            const int lg = est > 0 ? 31 - __clz(est) : 0;
            int c = lgmax - lg;
            c = c < 0 ? 0 : (c >= kSkyClasses ? kSkyClasses - 1 : c);
            mine = c == want;
        }
        const uint64_t m = __ballot(mine);
        if (lane == 0) S.wn[wid] = __popcll(m);
        __syncthreads();
        int base = 0;
        for (int w = 0; w < wid; ++w) base += S.wn[w];
        if (mine) S.list[base + __popcll(m & ((1ull << lane) - 1ull))] = item;
        const int nmine = S.wn[0] + S.wn[1] + S.wn[2] + S.wn[3];
        __syncthreads();
        for (int l = 0; l < nmine; ++l) {
            const unsigned it = S.list[l];
            DiscRec &rec = recs[n - 1 - (int64_t)it];
            const PGeo g = particle_geo(rec, nside, pad);
            const int nwin = (g.nrings + 63) / 64;
            double nacc = 0.0;
            int T = 0;
            // each wave takes 64-ring windows: lane j sets up ring j of the
            // window, then the wave walks the window's rings one by one
            for (int k = wid; k < nwin; k += 4) {
                const int gr = k * 64 + lane;
                int cnt = 0;
                if (gr < g.nrings) {
                    const RingRec rr = ring_setup(g, nside, gr);
                    cnt = rr.cnt;
                    S.f[wid][lane] = make_float4(rr.A, rr.B, rr.offm, rr.step);
                    S.q[wid][lane] = make_float4(rr.E, rr.Ac, rr.D, 0.0f);
                    S.c[wid][lane] = make_int2(cnt, rr.kc_rel);
                } else {
                    S.c[wid][lane] = make_int2(0, 0);
                }
                T += cnt;
                wave_sync();
                for (int j = 0; j < 64; ++j) {
                    const int2 c = S.c[wid][j];
                    if (c.x <= 0) continue;
                    const float4 f = S.f[wid][j];
                    const float4 q = S.q[wid][j];
                    if (g.mode == 0) nacc += row_weights<0>(c, f, f.y, q, g.K4, lane);
                    else if (g.mode == 2) nacc += row_weights<2>(c, f, f.y, q, g.K4, lane);
                    else nacc += row_weights<1>(c, f, q.x, q, g.K4, lane);
                }
                wave_sync();
            }
            const double ws = wave_sum(nacc);
            const int Tw = wave_reduce_full<DppAdd>(T, 0);
            if (lane == 0) {
                S.red[wid] = ws;
                S.tot[wid] = Tw;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                const double norm = S.red[0] + S.red[1] + S.red[2] + S.red[3];
                const int Ttot = S.tot[0] + S.tot[1] + S.tot[2] + S.tot[3];
                if (Ttot < 2) {
                    acc_point<ACC>(map, g.p0, g.flux);
                    atomicAdd(&stats[kStatFallback], 1ull);
                    ranges[it] = pack_rings(1, 0);
                } else {
                    rec.flux = g.flux / norm;
                    atomicAdd(&stats[kStatExt], 1ull);
                    atomicAdd(&stats[kStatLarge], 1ull);
                    atomicAdd(&stats[kStatContrib], (unsigned long long)Ttot);
                    atomicAdd(&stats[kStatLargeContrib], (unsigned long long)Ttot);
                    int rlo, rhi;
                    sky_ring_range(g, nside, rlo, rhi);
                    ranges[it] = pack_rings(rlo, rhi);
                }
            }
            __syncthreads();
        }
    }
}

// The batch's particles on the current ring, and the LDS copy of one
// segment of the ring.
struct GatherLDS {
    float4 f[64]; // A, B (series forms) or E (general form), offm, step
    float4 q[64]; // Ac, D, K4, flux / norm
    int4 c[64];   // first column in [0, nr), count, kc_rel, weight form
    unsigned cand[64];
    float win[kSkySeg];
};

// The k-th ring to gather: the 2 nside + 1 rings of 4 nside pixels first,
// then the polar rings from the largest down, north and south alternately.
__device__ __forceinline__ int gather_ring(int k, int nside)
{
    if (k < 2 * nside + 1) return nside + k;
    const int j = k - (2 * nside + 1);
    return (j & 1) == 0 ? nside - 1 - (j >> 1) : 3 * nside + 1 + (j >> 1);
}

// Adds one particle's columns in a ring segment into the LDS copy: while
// the wave has 256 columns left, four per lane at a time -- their weights
// first, then the four slot reads, then the four writes (the columns are
// distinct, so the slots are), four independent read-modify-writes in
// flight -- then one at a time.
template <int MODE>
__device__ __forceinline__ void gather_row(float *win, const SegSpan &sp, const int4 &ci, const float4 &f,
                                           const float4 &q, int nr, int s0, int lane)
{
    const int total = sp.na + sp.nb;
    int i0 = 0;
    for (; i0 + 256 <= total; i0 += 256) {
        int sl[4];
        float v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = i0 + 64 * k + lane;
            const int t = i < sp.na ? sp.u + i : i - sp.na;
            v[k] = weight32<MODE>(fmaf((float)(t - ci.z), f.w, f.z), f.x, f.y, q.x, q.y, q.z) * q.w;
            sl[k] = seg_slot(ci.x, t, nr, s0);
        }
        float a[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) a[k] = win[sl[k]];
#pragma unroll
        for (int k = 0; k < 4; ++k) win[sl[k]] = a[k] + v[k];
    }
    for (int i = i0 + lane; i < total; i += 64) {
        const int t = i < sp.na ? sp.u + i : i - sp.na;
        const float v = weight32<MODE>(fmaf((float)(t - ci.z), f.w, f.z), f.x, f.y, q.x, q.y, q.z) * q.w;
        const int sl = seg_slot(ci.x, t, nr, s0);
        win[sl] = win[sl] + v;
    }
}

// Renders the batch S.cand[0..cn) on ring r (first pixel startpix, nr
// pixels).
template <int ACC>
__device__ void gather_batch(GatherLDS &S, int cn, const DiscRec *__restrict__ recs, int64_t n, int nside,
                             double pad, int r, int startpix, int nr, const MapRef &map, int lane)
{
    if (lane < cn) {
        const DiscRec rec = recs[n - 1 - (int64_t)S.cand[lane]];
        const PGeo g = particle_geo(rec, nside, pad);
        const RingRec rr = ring_setup(g, nside, ring_index(g, r));
        int c0 = rr.col0 % nr;
        if (c0 < 0) c0 += nr;
        S.f[lane] = make_float4(rr.A, g.mode == 1 ? rr.E : rr.B, rr.offm, rr.step);
        S.q[lane] = make_float4(rr.Ac, rr.D, g.K4, (float)rec.flux);
        S.c[lane] = make_int4(c0, rr.cnt, rr.kc_rel, g.mode);
    }
    wave_sync();
    for (int s0 = 0; s0 < nr; s0 += kSkySeg) {
        const int ws = nr - s0 < kSkySeg ? nr - s0 : kSkySeg;
        for (int c = 0; c < cn; ++c) {
            const int4 ci = S.c[c];
            if (ci.y <= 0) continue;
            const SegSpan sp = seg_span(ci.x, ci.y, nr, s0, ws);
            if (sp.na + sp.nb == 0) continue;
            const float4 f = S.f[c];
            const float4 q = S.q[c];
            if (ci.w == 0) gather_row<0>(S.win, sp, ci, f, q, nr, s0, lane);
            else if (ci.w == 2) gather_row<2>(S.win, sp, ci, f, q, nr, s0, lane);
            else gather_row<1>(S.win, sp, ci, f, q, nr, s0, lane);
            // the next particle's slot reads are ordered after these
            // writes by the wave's in-order LDS queue; only the compiler
            // must not move them
            asm volatile("" ::: "memory");
        }
        wave_sync();
        // the segment into the map; its slots go back to zero
        for (int s = lane; s < ws; s += 64) {
            const float v = S.win[s];
            if (v != 0.0f) {
                const int p = startpix + s0 + s;
                if constexpr (ACC == kAccComp) {
                    const float old = map.f32[p];
                    const float sum = old + v;
                    const float bb = sum - old;
                    const float err = (old - (sum - bb)) + (v - bb);
                    map.f32[p] = sum;
                    if (err != 0.0f) map.lo[p] += err;
                } else if constexpr (ACC == kAccFix) {
                    map.fix[p] += (unsigned long long)__double2ll_rn((double)v * map.fix_scale);
                } else {
                    map.f32[p] += v;
                }
                S.win[s] = 0.0f;
            }
        }
        wave_sync();
    }
}

template <int ACC>
__global__ __launch_bounds__(64) void k_sky_gather(const DiscRec *__restrict__ recs, int64_t n,
                                                   unsigned *__restrict__ counts, int nside, double pad, MapRef map,
                                                   const unsigned *__restrict__ ranges)
{
    __shared__ GatherLDS S;
    const int lane = threadIdx.x;
    const unsigned nsky = counts[1];
    const int nrings = 4 * nside - 1;
    if (nsky == 0) return;
    for (int s = lane; s < kSkySeg; s += 64) S.win[s] = 0.0f;
    wave_sync();
    for (;;) {
        unsigned k = 0;
        if (lane == 0) k = atomicAdd(&counts[3], 1u);
        k = (unsigned)wave_read((int)k, 0);
        if (k >= (unsigned)nrings) break;
        const int r = gather_ring((int)k, nside);
        int startpix, nr, shifted;
        ring_info(nside, r, startpix, nr, shifted);
        int cn = 0;
        // the scan reads four packed ranges per lane at a time (one 16-byte
        // load; the buffer and every 256-entry block start are 16-byte
        // aligned), entries past nsky read as the empty range
        for (unsigned base = 0; base < nsky; base += 256) {
            const unsigned i0 = base + 4u * (unsigned)lane;
            uint4 pr4 = make_uint4(1u, 1u, 1u, 1u);
            if (i0 + 3u < nsky) {
                pr4 = *reinterpret_cast<const uint4 *>(ranges + i0);
            } else {
                if (i0 < nsky) pr4.x = ranges[i0];
                if (i0 + 1u < nsky) pr4.y = ranges[i0 + 1u];
                if (i0 + 2u < nsky) pr4.z = ranges[i0 + 2u];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const unsigned pr = k == 0 ? pr4.x : k == 1 ? pr4.y : k == 2 ? pr4.z : pr4.w;
                const bool hit = (int)(pr & 0xffffu) <= r && r <= (int)(pr >> 16);
                const uint64_t hm = __ballot(hit);
                const int nh = __popcll(hm);
                if (cn + nh > 64) {
                    gather_batch<ACC>(S, cn, recs, n, nside, pad, r, startpix, nr, map, lane);
                    cn = 0;
                }
                if (hit) S.cand[cn + __popcll(hm & ((1ull << lane) - 1ull))] = i0 + (unsigned)k;
                cn += nh;
                wave_sync();
            }
        }
        if (cn > 0) gather_batch<ACC>(S, cn, recs, n, nside, pad, r, startpix, nr, map, lane);
    }
}

} // namespace

hipError_t launch_sky(int acc, DiscRec *recs, int64_t n, unsigned *counts, int nside, double pad, MapRef map,
                      unsigned long long *stats, unsigned *ranges, int norm_blocks, int gather_blocks, hipStream_t st)
{
    if (acc == kAccComp) {
        k_sky_norm<kAccComp><<<norm_blocks, 256, 0, st>>>(recs, n, counts, nside, pad, map, stats, ranges);
        k_sky_gather<kAccComp><<<gather_blocks, 64, 0, st>>>(recs, n, counts, nside, pad, map, ranges);
    } else if (acc == kAccF32) {
        k_sky_norm<kAccF32><<<norm_blocks, 256, 0, st>>>(recs, n, counts, nside, pad, map, stats, ranges);
        k_sky_gather<kAccF32><<<gather_blocks, 64, 0, st>>>(recs, n, counts, nside, pad, map, ranges);
    } else if (acc == kAccFix) {
        k_sky_norm<kAccFix><<<norm_blocks, 256, 0, st>>>(recs, n, counts, nside, pad, map, stats, ranges);
        k_sky_gather<kAccFix><<<gather_blocks, 64, 0, st>>>(recs, n, counts, nside, pad, map, ranges);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

} // namespace dmann
