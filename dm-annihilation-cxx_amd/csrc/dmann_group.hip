// Grouped footprint rendering with wave-private LDS accumulation.
//
// Global FP32 atomics on MI355X retire about one dword per clock per L2
// channel (~2.6e11/s chip-wide) even when an instruction's 64 lanes hit
// consecutive dwords (fewer when they spread over more cache lines), and
// LDS atomics are no faster; with ~1e12 footprint pixels per map, one atomic per pixel is
// the bottleneck.  Here the medium-class particles are first sorted by sky
// cell and footprint-size class (radix sort, or a counting sort in the
// graph-captured batch), then one wavefront takes K consecutive particles
// (a group of neighbours), in ascending flux order:
//   1. the normalisation pass (FP32 weights, FP64 sums): every (particle,
//      ring) row of the group on a lane of its own, 64 rows at a time;
//   2. the group's window -- the rings irmin..irmax of all members and the
//      union of their longitude ranges -- is walked in bands of rings that fit
//      the wave's LDS buffer; each member deposits its band pixels with a
//      plain LDS read-modify-write (no atomics: within one instruction every
//      lane holds a distinct pixel, and members are processed in turn);
//   3. the band is flushed with one global atomic (form) per non-zero pixel.
// A group's pixels are written ~K times less often than with per-pixel
// atomics.  The window holds FP32 partial sums of at most K deposits per
// pixel (one per member, the small ones first; error <= (K-1) u of the
// partial sum); the
// compensated form flushes them with the two-float atomic pair (exact
// TwoSum), so the sum over groups -- 1e4-1e5 terms per pixel -- stays
// compensated.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include "dmann_device.h"
#include "dmann_footprint.h"
#include "dmann_geom.h"
#include "dmann_kernels.h"

namespace dmann {

// LDS budget of one wave: 20 KB, so 8 single-wave blocks fit a 160 KB CU
// (the window takes what the tables leave, in whole kSeg blocks: per slot
// 4 bytes, and per kSeg slots one block-owner byte and one segment-ring
// byte).
#ifndef DMANN_GROUP_CAP
#define DMANN_GROUP_CAP ((15216 * kSeg) / (4 * kSeg + 2) / kSeg * kSeg)
#endif
constexpr int kGroupCapF32 = DMANN_GROUP_CAP; // FP32 window slots
// segments whose table ring is looked up directly (a window's segments all
// are: a band ring's row fills at most its columns plus a partial segment)
constexpr int kSegTab = kGroupCapF32 / kSeg + 64;

struct GroupLDS {
    // ring table of up to 64 rings of one member
    float4 rf[64];   // e0, e1, e2, step (RingLite weight terms)
    float rx[64];    // offm
    int4 ri[64];     // kc_rel, cnt, first segment, end segment
    int4 rs[64];     // window slot offset, window columns, column offset in the window ring, nr
    int2 rd[64];     // col0, first pixel of the ring (direct deposits)
    // band table: ring j owns slots bt[j].x .. bt[j+1].x - 1 (a multiple
    // of kSeg, at least window columns + kSeg)
    int4 bt[65];     // first slot, first window column (wrapped into [0, nr)), nr, first pixel
    int bnc[64];     // window columns
    unsigned char ros[kGroupCapF32 / kSeg]; // band ring owning each block of kSeg slots
    unsigned char q0[kSegTab];              // table ring of each segment
    float acc[kGroupCapF32];
};
static_assert(kGroupCapF32 % kSeg == 0, "window blocks");
#ifndef DMANN_GROUP_LDS_MAX
#define DMANN_GROUP_LDS_MAX 20480
#endif
static_assert(sizeof(GroupLDS) <= DMANN_GROUP_LDS_MAX, "eight waves per CU");

__device__ __forceinline__ double wave_min_d(double v) { return wave_reduce_full<DppMin>(v, 1e300); }
__device__ __forceinline__ double wave_max_d(double v) { return wave_reduce_full<DppMax>(v, -1e300); }
__device__ __forceinline__ int wave_min_i(int v) { return wave_reduce_full<DppMin>(v, 0x7fffffff); }
__device__ __forceinline__ int wave_max_i(int v) { return wave_reduce_full<DppMax>(v, -0x7fffffff - 1); }

// Ring terms packed for the per-wave cache: (e0, e1, e2, offm) and
// (cnt, col0, kc_rel in the low / nr in the high 16 bits, step).  The
// ring's first pixel is not kept: band tables hold it.
__device__ __forceinline__ void ring_pack(const RingLite &rl, float4 &f, int4 &i)
{
    f = make_float4(rl.e0, rl.e1, rl.e2, rl.offm);
    i = make_int4(rl.cnt, rl.col0, (int)(((unsigned)rl.nr << 16) | ((unsigned)rl.kc_rel & 0xffffu)),
                  __float_as_int(rl.step));
}
__device__ __forceinline__ RingLite ring_unpack(const float4 &f, const int4 &i)
{
    RingLite rl;
    rl.e0 = f.x;
    rl.e1 = f.y;
    rl.e2 = f.z;
    rl.offm = f.w;
    rl.cnt = i.x;
    rl.col0 = i.y;
    rl.nr = (int)((unsigned)i.z >> 16);
    rl.kc_rel = (int)(short)(i.z & 0xffff);
    rl.step = __int_as_float(i.w);
    rl.startpix = -1;
    return rl;
}

// Fills the LDS ring table from lane q's ring terms (lanes q < n hold ring
// iz0 + q): segment ranges, the table ring of every segment (the first
// kSegTab of them) and, with band_ring0 >= 0, the window slot terms of the
// band.  Returns the number of segments.  Every lane must call it.
__device__ __forceinline__ int ring_table_fill(const RingLite &rl, bool on, int iz0, GroupLDS &S, int lane,
                                               int band_ring0)
{
    const int cnt = on ? rl.cnt : 0;
    const int nseg = (cnt + kSeg - 1) / kSeg;
    const int incl = wave_scan_full<DppAdd>(nseg, 0);
    const int NS = wave_read(incl, 63);
    // this ring's segments in the segment-ring table
    const int qe = incl < kSegTab ? incl : kSegTab;
    for (int l = incl - nseg; l < qe; ++l) S.q0[l] = (unsigned char)lane;
    if (on) {
        S.rf[lane] = make_float4(rl.e0, rl.e1, rl.e2, rl.step);
        S.rx[lane] = rl.offm;
        S.ri[lane] = make_int4(rl.kc_rel, cnt, incl - nseg, incl);
        if (band_ring0 < 0) S.rd[lane] = make_int2(rl.col0, rl.startpix);
        if (band_ring0 >= 0) {
            // slot terms: the band ring owns ncols + kSeg slots from boff;
            // a segment may fill slots off .. off+len-1 below the limit.
            // A whole-ring window takes the ring's wrap into the kSeg spare
            // slots, which needs cnt + kSeg <= nr to keep the masked tail of
            // the last segment off the first one (else limit 0: direct).
            const int rel = iz0 + lane - band_ring0;
            const int4 b = S.bt[rel];
            S.rd[lane] = make_int2(rl.col0, b.w);
            // col0 in [-nr/2 - 1, nr), first window column in [0, nr)
            int off = rl.col0 - b.y;
            if (off < 0) off += rl.nr;
            if (off < 0) off += rl.nr;
            const int ncols = S.bnc[rel];
            const int lim = ncols < rl.nr ? ncols : (cnt + kSeg <= rl.nr ? rl.nr + kSeg : 0);
            S.rs[lane] = make_int4(b.x, lim, off, rl.nr);
        }
    } else {
        S.ri[lane] = make_int4(0, 0, incl, incl);
    }
    wave_sync();
    return NS;
}

// Deposits pixels t0 .. t0+len-1 of table ring q straight into the map
// (global atomics), for bands whose window does not fit the LDS buffer.
template <int ACC, int PV>
__device__ __forceinline__ void deposit_direct(const GroupLDS &S, int q, int t0, int len, float x0, const float4 &f,
                                               float K4, float sc, const MapRef &map)
{
    const int2 d = S.rd[q];
    const int nr = S.rs[q].w;
#pragma unroll
    for (int t = 0; t < kSeg; ++t) {
        if (t < len) {
            const float v = ring_term_weight<PV>(fmaf((float)t, f.w, x0), f.x, f.y, f.z, K4) * sc;
            int col = d.x + t0 + t;
            if (col < 0) col += nr;
            else if (col >= nr) col -= nr;
            acc_add<ACC>(map, d.y + col, v, 0.0);
        }
    }
}

// 1 for pixel t of a segment of len pixels (t < len), else 0: one VALU
// operation (a subtraction with the output clamp), since len and t are
// integers.
__device__ __forceinline__ float seg_mask(float lenf, int t)
{
    return __builtin_amdgcn_fmed3f(lenf - (float)t, 0.0f, 1.0f);
}

// Table ring of segment s: from the segment-ring table, or beyond it by
// stepping from the last tabled ring (at least q_from).
__device__ __forceinline__ int seg_ring(const GroupLDS &S, int s, int q_from)
{
    if (s < kSegTab) return S.q0[s];
    int q = q_from < S.q0[kSegTab - 1] ? S.q0[kSegTab - 1] : q_from;
    while (S.ri[q].w <= s) ++q;
    return q;
}

// Deposits weight * sc for the table's segments, walked round-robin: in
// iteration i lane L takes segment 64 i + L, so the 64 lanes hold
// consecutive segments -- within a ring their window slots are kSeg (odd)
// apart, in distinct LDS banks, and their ring-term reads mostly hit one
// table row (a broadcast).  kSeg pixels per segment with the ring terms
// loaded once per segment.  Every lane evaluates all kSeg pixels of its
// segment; those past the ring's end are masked to zero (a factor 0 in an
// fma: the weights there are finite), so the loop has no divergent branch.
// Pixel 0 of a segment always lies on the ring.
// The LDS reads are software-pipelined: the table ring of segment s + 128
// and the ring terms of segment s + 64 are read in the iteration of
// segment s, before its window slots, so each iteration waits for one LDS
// round trip instead of three (ring, terms, slots).  The table is read-only
// during the walk and the slots are a separate array, so the reordering
// changes no value.
// WINDOW: into the LDS window, where each band ring owns window-columns +
// kSeg slots, so the masked tail of a ring's last segment lands in slots no
// other lane touches; a segment that would leave its ring's slots (which
// the window construction rules out) goes to global atomics instead.
// Without WINDOW every deposit is a global atomic.
template <int ACC, bool WINDOW, int PV>
__device__ __forceinline__ void deposit_segments(GroupLDS &S, int NS, int lane, float K4, float sc,
                                                 const MapRef &map)
{
    const int nmax = (NS + 63) >> 6;
    int s = lane;
    int q = s < NS ? seg_ring(S, s, 0) : 0;
    int qn = s + 64 < NS ? seg_ring(S, s + 64, q) : q;
    float4 f = S.rf[q];
    int4 r = S.ri[q];
    float rx = S.rx[q];
    int4 w = S.rs[q];
    for (int i = 0; i < nmax; ++i) {
        // next segment's terms and the ring of the one after (read-only)
        const int qn2 = s + 128 < NS ? seg_ring(S, s + 128, qn) : qn;
        const float4 fn = S.rf[qn];
        const int4 rn = S.ri[qn];
        const float rxn = S.rx[qn];
        const int4 wn = S.rs[qn];
        if (s < NS) {
            const int t0 = (s - r.z) * kSeg;
            const int len = r.y - t0 < kSeg ? r.y - t0 : kSeg;
            const float lenf = (float)len;
            const float x0 = fmaf((float)(t0 - r.x), f.w, rx);
            if constexpr (WINDOW) {
                int off = w.z + t0;
                if (off >= w.w) off -= w.w;
                if (off + len <= w.y) {
                    float *a = &S.acc[w.x + off];
#pragma unroll
                    for (int t = 0; t < kSeg; ++t) {
                        const float x = t == 0 ? x0 : fmaf((float)t, f.w, x0);
                        const float v = ring_term_weight<PV>(x, f.x, f.y, f.z, K4);
                        a[t] = fmaf(v, t == 0 ? sc : sc * seg_mask(lenf, t), a[t]);
                    }
                } else {
                    deposit_direct<ACC, PV>(S, q, t0, len, x0, f, K4, sc, map);
                }
            } else {
                deposit_direct<ACC, PV>(S, q, t0, len, x0, f, K4, sc, map);
            }
        }
        s += 64;
        q = qn;
        qn = qn2;
        f = fn;
        r = rn;
        rx = rxn;
        w = wn;
    }
}

// Sum of the weights of one table row held in a lane's registers: pixels
// t = 0 .. cnt-1 at longitude offset x_t = offm + (t - kc_rel) step, four at
// a time from x of the first (two roundings, as in the segment walks), the
// last 0-3 masked.  FP32 partial sums of 16 weights (four groups of four)
// go to the FP64 total once per 16.  The trip count is the lane's own:
// rows of different lengths share the wavefront.
template <int PV>
__device__ __forceinline__ float four_weights(const RingLite &rl, int t0, float K4)
{
    const float xb = fmaf((float)(t0 - rl.kc_rel), rl.step, rl.offm);
    const float w0 = ring_term_weight<PV>(xb, rl.e0, rl.e1, rl.e2, K4);
    const float w1 = ring_term_weight<PV>(fmaf(1.0f, rl.step, xb), rl.e0, rl.e1, rl.e2, K4);
    const float w2 = ring_term_weight<PV>(fmaf(2.0f, rl.step, xb), rl.e0, rl.e1, rl.e2, K4);
    const float w3 = ring_term_weight<PV>(fmaf(3.0f, rl.step, xb), rl.e0, rl.e1, rl.e2, K4);
    return (w0 + w1) + (w2 + w3);
}

template <int PV>
__device__ __forceinline__ double row_sum(const RingLite &rl, float K4)
{
    const int n4 = rl.cnt >> 2;
    double acc = 0.0;
    int c = 0;
    for (; c + 4 <= n4; c += 4) {
        float part = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) part += four_weights<PV>(rl, 4 * (c + k), K4);
        acc += (double)part;
    }
    float part = 0.0f;
    for (; c < n4; ++c) part += four_weights<PV>(rl, 4 * c, K4);
    const int t0 = 4 * n4;
    const float lenf = (float)(rl.cnt - t0);
    const float xb = fmaf((float)(t0 - rl.kc_rel), rl.step, rl.offm);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float w = ring_term_weight<PV>(k == 0 ? xb : fmaf((float)k, rl.step, xb), rl.e0, rl.e1, rl.e2, K4);
        part = fmaf(w, seg_mask(lenf, k), part);
    }
    return acc + (double)part;
}

// deposit_segments with the particle's series lengths (fast_variant).
template <int ACC, bool WINDOW>
__device__ __forceinline__ void deposit_pv(int pv, GroupLDS &S, int NS, int lane, float K4, float sc,
                                           const MapRef &map)
{
    switch (pv) {
    case 0: deposit_segments<ACC, WINDOW, 0>(S, NS, lane, K4, sc, map); break;
    case 1: deposit_segments<ACC, WINDOW, 1>(S, NS, lane, K4, sc, map); break;
    default: deposit_segments<ACC, WINDOW, 2>(S, NS, lane, K4, sc, map); break;
    }
}

// The lo words of FB two-float additions whose hi atomics returned old:
// the exact rounding error of old + v (TwoSum).
template <int FB>
__device__ __forceinline__ void two_float_lo(const MapRef &map, const float *old, const float *v, const int *pix)
{
#pragma unroll
    for (int u = 0; u < FB; ++u) {
        const float sum = old[u] + v[u];
        const float bb = sum - old[u];
        const float err = (old[u] - (sum - bb)) + (v[u] - bb);
        if (err != 0.0f) unsafeAtomicAdd(map.lo + pix[u], err);
    }
}

// Flushes the band's window slots 0 .. Sb-1 (Sb >= 1): one global atomic
// (form) per non-zero slot.  Slots are taken FB x 64 at a time with
// unconditional (clamped) LDS reads, so the lookups of the FB chunks overlap
// and the returning atomics of the compensated form are all in flight
// before the first result is waited for.
template <int ACC>
__device__ __forceinline__ void flush_band(GroupLDS &S, int Sb, int lane, const MapRef &map, unsigned npix)
{
    constexpr int FB = ACC == kAccComp ? 8 : 4;
    float pold[FB], pv[FB];
    int ppix[FB];
    bool have = false;
    for (int t0 = 0; t0 < Sb; t0 += 64 * FB) {
        int tc[FB], j[FB], pix[FB];
        float v[FB];
#pragma unroll
        for (int u = 0; u < FB; ++u) {
            const int t = t0 + 64 * u + lane;
            tc[u] = t < Sb ? t : Sb - 1;
            j[u] = S.ros[(unsigned)tc[u] / kSeg];
        }
#pragma unroll
        for (int u = 0; u < FB; ++u) {
            const int4 b = S.bt[j[u]];
            const float a = S.acc[tc[u]];
            // slot offsets run below ncols + 2 kSeg (whole rings: nr + 2 kSeg)
            int col = b.y + (tc[u] - b.x);
            if (col >= b.z) col -= b.z;
            if (col >= b.z) col -= b.z;
            pix[u] = b.w + col;
            // (the pixel test only guards the map against a malformed table)
            v[u] = t0 + 64 * u + lane < Sb && (unsigned)pix[u] < npix ? a : 0.0f;
        }
        if constexpr (ACC == kAccComp) {
            // two-float accumulation: hi by a returning atomic, then the
            // exact rounding error of that addition (TwoSum) into lo.  The
            // lo atomics of a chunk are issued after the next chunk's hi
            // atomics, so two chunks of returning atomics are in flight
            float old[FB];
#pragma unroll
            for (int u = 0; u < FB; ++u) {
                old[u] = 0.0f;
                if (v[u] != 0.0f) old[u] = unsafeAtomicAdd(map.f32 + pix[u], v[u]);
            }
            if (have) two_float_lo<FB>(map, pold, pv, ppix);
#pragma unroll
            for (int u = 0; u < FB; ++u) {
                pold[u] = old[u];
                pv[u] = v[u];
                ppix[u] = pix[u];
            }
            have = true;
        } else {
#pragma unroll
            for (int u = 0; u < FB; ++u)
                if (v[u] != 0.0f) acc_add<ACC>(map, pix[u], v[u], 0.0);
        }
    }
    if constexpr (ACC == kAccComp)
        if (have) two_float_lo<FB>(map, pold, pv, ppix);
}

// A band of window rings b0 .. b0+BH-1 (BH <= 64, rings up to R1) whose
// slots fit the LDS buffer: band table, block owners and zeroed slots.
// The window spans longitudes phic + [wlo, whi] (whole rings when full).
// direct: ring b0 alone exceeds the buffer; its deposits go to global
// atomics and the table only holds the ring layout.
struct Band {
    int BH;      // rings in the band
    int Sb;      // window slots in use
    bool direct;
};
__device__ __forceinline__ Band band_setup(GroupLDS &S, int lane, int nside, int b0, int R1, double phic,
                                           double wlo, double whi, bool full)
{
    const int ring = b0 + lane;
    const bool vr = ring <= R1;
    int startpix = 0, nr = 1, shifted = 0, cl = 0, ncols = 0;
    if (vr) {
        ring_info(nside, ring, startpix, nr, shifted);
        const double shift = shifted ? 0.5 : 0.0;
        if (full) {
            cl = 0;
            ncols = nr;
        } else {
            cl = (int)floor((double)nr * kInvTwoPi * (phic + wlo) - shift) - 1;
            const int ch = (int)floor((double)nr * kInvTwoPi * (phic + whi) - shift) + 1;
            ncols = ch - cl + 1 < nr ? ch - cl + 1 : nr;
        }
    }
    const int alloc = vr ? (ncols + 2 * kSeg - 1) / kSeg * kSeg : 0;
    const int incl = wave_scan_full<DppAdd>(alloc, 0);
    const bool fits = vr && incl <= kGroupCapF32;
    Band bd;
    bd.BH = __popcll(__ballot(fits));
    bd.direct = bd.BH == 0;
    bd.Sb = 0;
    if (bd.direct) {
        bd.BH = (R1 - b0 + 1) < 64 ? (R1 - b0 + 1) : 64;
        // direct deposits still need the band table for slot terms
        if (lane < bd.BH) {
            S.bt[lane] = make_int4(0, 0, nr, startpix);
            S.bnc[lane] = 0;
        }
    } else {
        bd.Sb = wave_read(incl, bd.BH - 1);
        if (lane < bd.BH) {
            int clw = cl % nr;
            if (clw < 0) clw += nr;
            S.bt[lane] = make_int4(incl - alloc, clw, nr, startpix);
            S.bnc[lane] = ncols;
            for (int k = (incl - alloc) / kSeg; k < incl / kSeg; ++k) S.ros[k] = (unsigned char)lane;
        }
        if (lane == 0) S.bt[bd.BH] = make_int4(bd.Sb, 0, 1, 0);
        for (int t = lane; t < bd.Sb; t += 64) S.acc[t] = 0.0f;
    }
    wave_sync();
    return bd;
}

// Longitude window of the members with on set: the lead member's phi and
// the extent of the others' ranges relative to it.
struct Window {
    double phic, wlo, whi;
    bool full;
};
__device__ __forceinline__ Window group_window(bool on, double phi, double dphi)
{
    const uint64_t onmask = __ballot(on);
    const int lead = __ffsll((unsigned long long)onmask) - 1;
    Window w;
    w.phic = wave_read(phi, lead);
    double dl = phi - w.phic;
    if (dl > kPi) dl -= kTwoPi;
    if (dl < -kPi) dl += kTwoPi;
    w.wlo = wave_min_d(on ? dl - dphi : 1e30);
    w.whi = wave_max_d(on ? dl + dphi : -1e30);
    w.full = (w.whi - w.wlo) >= kTwoPi - 1e-9;
    return w;
}

#ifdef DMANN_GROUP_WPE
#define DMANN_GROUP_ATTR __attribute__((amdgpu_waves_per_eu(DMANN_GROUP_WPE)))
#else
#define DMANN_GROUP_ATTR
#endif

template <int ACC>
__global__ __launch_bounds__(64) DMANN_GROUP_ATTR void k_disc_group(const DiscRec *__restrict__ recs, const unsigned *__restrict__ idx,
                                                   const unsigned *__restrict__ counts, int nside, double pad, int K, MapRef map,
                                                   unsigned long long *__restrict__ stats, unsigned *__restrict__ work,
                                                   float4 *__restrict__ tab_f, int4 *__restrict__ tab_i,
                                                   unsigned spread)
{
    __shared__ GroupLDS S;
    const int lane = threadIdx.x;
    const RingGeo G = ring_geo(nside);
    const unsigned npix = (unsigned)npix_of(nside);
    const unsigned count = counts[0];
    const unsigned ngroups = (count + (unsigned)K - 1u) / (unsigned)K;
    const unsigned long long mul = ngroups < spread ? spread : 1u;
    unsigned long long n_ext = 0, n_fb = 0, n_contrib = 0;
    // groups are taken from a counter: footprint sizes vary a lot
    for (;;) {
        unsigned grp = 0;
        if (lane == 0) grp = atomicAdd(work, 1u);
        grp = (unsigned)wave_read((int)grp, 0);
        if (grp >= ngroups) break;
        // spread != 1: the waves in flight take groups spread over the
        // sorted order (g -> g spread mod ngroups, a bijection for a prime
        // spread above the group count), not consecutive ones, whose
        // windows would cover one patch of sky (NESTED order) and whose
        // flushes would queue on the same map addresses.
        grp = (unsigned)(((unsigned long long)grp * mul) % ngroups);
        const unsigned m0 = grp * (unsigned)K;
        const int nm = (int)((m0 + (unsigned)K < count ? m0 + (unsigned)K : count) - m0);
        // member geometry, one member per lane
        float my_scale = 0.0f, my_K4 = 0.0f;
        int my_pv = 2, my_p0 = 0;
        int my_lo = 0x7fffffff, my_hi = -0x7fffffff;
        double my_phi = 0.0, my_dphi = 0.0, my_z0 = 0.0, my_s0 = 0.0, my_xa = 0.0, my_cosr = 0.0, my_flux = 0.0;
        bool my_on = false;
        if (lane < nm) {
            const PGeo mg = particle_geo(recs[idx[m0 + lane]], nside, pad);
            my_lo = mg.irmin;
            my_hi = mg.irmax;
            my_phi = mg.phi;
            my_dphi = mg.dphi;
            my_pv = mg.pv;
            my_z0 = mg.z0;
            my_s0 = mg.s0;
            my_xa = mg.xa;
            my_cosr = mg.cosr;
            my_K4 = mg.K4;
            my_flux = mg.flux;
            my_p0 = mg.p0;
        }
        // Members in ascending flux order (lane order is deposit order): a
        // window slot's FP32 partial sum then takes the small deposits
        // before the large ones, so they are not rounded away against a
        // large partial sum.  A bitonic sort of (flux, lane) across the
        // wavefront, then every member field gathered from its new lane.
        {
            double key = lane < nm ? my_flux : 1e300;
            int src = lane;
#pragma unroll
            for (int kk = 2; kk <= 64; kk <<= 1) {
#pragma unroll
                for (int jj = kk >> 1; jj > 0; jj >>= 1) {
                    const double ok = __shfl_xor(key, jj, 64);
                    const int os = __shfl_xor(src, jj, 64);
                    const bool asc = (lane & kk) == 0, low = (lane & jj) == 0;
                    const bool other_less = ok < key || (ok == key && os < src);
                    if (low == asc ? other_less : !other_less) {
                        key = ok;
                        src = os;
                    }
                }
            }
            my_lo = __shfl(my_lo, src, 64);
            my_hi = __shfl(my_hi, src, 64);
            my_phi = __shfl(my_phi, src, 64);
            my_dphi = __shfl(my_dphi, src, 64);
            my_pv = __shfl(my_pv, src, 64);
            my_z0 = __shfl(my_z0, src, 64);
            my_s0 = __shfl(my_s0, src, 64);
            my_xa = __shfl(my_xa, src, 64);
            my_cosr = __shfl(my_cosr, src, 64);
            my_K4 = __shfl(my_K4, src, 64);
            my_flux = __shfl(my_flux, src, 64);
            my_p0 = __shfl(my_p0, src, 64);
        }
        // ---- phase 1: normalisation.  Every (member, ring) row of the group
        // gets a lane of its own, 64 rows per pass, in member order: the lane
        // computes the row's column range and weight terms (ring_lite),
        // keeps them in registers for the walk over the row's pixels, and
        // stores them to the wave's ring-term cache (rows below kCacheRows)
        // for the deposit bands.  Row sums are then added up per member.
        const int nrow = (lane < nm && my_hi >= my_lo) ? my_hi - my_lo + 1 : 0;
        const int rincl = wave_scan_full<DppAdd>(nrow, 0);
        const int rstart = rincl - nrow; // member's first row in the group
        const int NR = wave_read(rincl, 63);
        const size_t cache0 = (size_t)blockIdx.x * (size_t)K * 64;
        const int cache_rows = K * 64;
        double my_norm = 0.0;
        int my_T = 0;
        int carry = 0; // member of the last row of the previous pass
        // member marks of one pass (the ring table is not in use here)
        int *mark = reinterpret_cast<int *>(S.ri);
        for (int p0 = 0; p0 < NR; p0 += 64) {
            const int g = p0 + lane;
            const bool valid = g < NR;
            mark[lane] = -1;
            wave_sync();
            if (nrow > 0 && rstart >= p0 && rstart < p0 + 64) mark[rstart - p0] = lane;
            wave_sync();
            // member of row g: the last member starting at or before it
            int mj = wave_scan_full<DppMax>(mark[lane], -1);
            if (mj < 0) mj = carry;
            carry = wave_read(mj, 63);
            PGeo gm;
            gm.phi = __shfl(my_phi, mj, 64);
            gm.z0 = __shfl(my_z0, mj, 64);
            gm.s0 = __shfl(my_s0, mj, 64);
            gm.xa = __shfl(my_xa, mj, 64);
            gm.cosr = __shfl(my_cosr, mj, 64);
            gm.K4 = __shfl(my_K4, mj, 64);
            gm.pv = __shfl(my_pv, mj, 64);
            const int lo = __shfl(my_lo, mj, 64);
            const int rs0 = __shfl(rstart, mj, 64);
            RingLite rl;
            rl.cnt = 0;
            double rowsum = 0.0;
            if (valid) {
                rl = ring_lite(gm, G, lo + (g - rs0));
                if (g < cache_rows) {
                    float4 f;
                    int4 i;
                    ring_pack(rl, f, i);
                    tab_f[cache0 + (size_t)g] = f;
                    tab_i[cache0 + (size_t)g] = i;
                }
                if (rl.cnt > 0) {
                    switch (gm.pv) {
                    case 0: rowsum = row_sum<0>(rl, gm.K4); break;
                    case 1: rowsum = row_sum<1>(rl, gm.K4); break;
                    default: rowsum = row_sum<2>(rl, gm.K4); break;
                    }
                }
            }
            // per-member totals of this pass's rows
            const int mfirst = wave_read(mj, 0);
            const int mlast = wave_read(mj, NR - 1 - p0 < 63 ? NR - 1 - p0 : 63);
            for (int m = mfirst; m <= mlast; ++m) {
                const bool in = valid && mj == m;
                const double sm = wave_reduce_full<DppAdd>(in ? rowsum : 0.0, 0.0);
                const int tm = wave_reduce_full<DppAdd>(in ? rl.cnt : 0, 0);
                if (lane == m) {
                    my_norm += sm;
                    my_T += tm;
                }
            }
        }
        // footprints of fewer than 2 pixels: whole flux to the particle's
        // pixel (src/genmap.cpp:125-128)
        if (lane < nm) {
            if (my_T < 2) {
                acc_point<ACC>(map, my_p0, my_flux);
            } else {
                my_scale = (float)(my_flux / my_norm);
                my_on = true;
            }
        }
        {
            const uint64_t fb = __ballot(lane < nm && my_T < 2);
            const uint64_t ex = __ballot(my_on);
            n_fb += (unsigned long long)__popcll(fb);
            n_ext += (unsigned long long)__popcll(ex);
            n_contrib += (unsigned long long)wave_reduce_full<DppAdd>(my_on ? my_T : 0, 0);
        }
        // the cached ring terms are read back below
        __builtin_amdgcn_s_waitcnt(0);
        // ---- the group's window
        const int R0 = wave_min_i(my_on ? my_lo : 0x7fffffff);
        const int R1 = wave_max_i(my_on ? my_hi : -0x7fffffff);
        if (R0 > R1) continue;
        const Window win = group_window(my_on, my_phi, my_dphi);
        // ---- phase 2: bands of rings that fit the LDS buffer
        for (int b0 = R0; b0 <= R1;) {
            const Band bd = band_setup(S, lane, nside, b0, R1, win.phic, win.wlo, win.whi, win.full);
            const int BH = bd.BH, Sb = bd.Sb;
            const bool direct = bd.direct;
            const int b1 = b0 + BH - 1;
            // members meeting the band, in order (a ballot of them, taken
            // lowest first); the cached ring terms of the next two are
            // loaded while the current one is walked (the cache rows come
            // from L2 or memory)
            uint64_t meet = __ballot(my_on && my_hi >= b0 && my_lo <= b1);
            auto next_member = [&](int) {
                if (meet == 0) return nm;
                const int j = __ffsll((unsigned long long)meet) - 1;
                meet &= meet - 1;
                return j;
            };
            auto fetch = [&](int j, float4 &f, int4 &i) {
                const int lo = wave_read(my_lo, j), hi = wave_read(my_hi, j);
                const int rlo = b0 > lo ? b0 : lo, rhi = b1 < hi ? b1 : hi;
                const int row = wave_read(rstart, j) + (rlo - lo) + lane;
                if (lane <= rhi - rlo && row < cache_rows) {
                    f = tab_f[cache0 + (size_t)row];
                    i = tab_i[cache0 + (size_t)row];
                }
            };
            int j = next_member(-1);
            float4 nf = make_float4(0.0f, 0.0f, 0.0f, 0.0f), nf2 = nf;
            int4 ni = make_int4(0, 0, 0, 0), ni2 = ni;
            if (j < nm) fetch(j, nf, ni);
            int jn = next_member(j);
            if (jn < nm) fetch(jn, nf2, ni2);
            while (j < nm) {
                const float4 cf = nf;
                const int4 ci = ni;
                nf = nf2;
                ni = ni2;
                const int jn2 = next_member(jn);
                if (jn2 < nm) fetch(jn2, nf2, ni2);
                const int lo = wave_read(my_lo, j), hi = wave_read(my_hi, j);
                const int rlo = b0 > lo ? b0 : lo, rhi = b1 < hi ? b1 : hi;
                const bool on = lane <= rhi - rlo;
                RingLite rl;
                rl.cnt = 0;
                if (on) {
                    if (wave_read(rstart, j) + (rlo - lo) + lane < cache_rows) {
                        rl = ring_unpack(cf, ci);
                    } else {
                        // rows past the wave's cache
                        PGeo g;
                        g.phi = wave_read(my_phi, j);
                        g.z0 = wave_read(my_z0, j);
                        g.s0 = wave_read(my_s0, j);
                        g.xa = wave_read(my_xa, j);
                        g.cosr = wave_read(my_cosr, j);
                        g.K4 = wave_read(my_K4, j);
                        g.pv = wave_read(my_pv, j);
                        rl = ring_lite(g, G, rlo + lane);
                    }
                }
                const int NS = ring_table_fill(rl, on, rlo, S, lane, b0);
                const float sc = wave_read(my_scale, j);
                const float K4 = wave_read(my_K4, j);
                const int pv = wave_read(my_pv, j);
                if (direct) deposit_pv<ACC, false>(pv, S, NS, lane, K4, sc, map);
                else deposit_pv<ACC, true>(pv, S, NS, lane, K4, sc, map);
                wave_sync();
                j = jn;
                jn = jn2;
            }
            if (!direct) {
                flush_band<ACC>(S, Sb, lane, map, npix);
                wave_sync();
            }
            b0 += BH;
        }
    }
    if (lane == 0) {
        if (n_ext) atomicAdd(&stats[kStatExt], n_ext);
        if (n_fb) atomicAdd(&stats[kStatFallback], n_fb);
        if (n_contrib) atomicAdd(&stats[kStatContrib], n_contrib);
    }
}

// ---------------------------------------------------------------- sorting

// Key of the radix sort: the NESTED cell at the coarse nside nh, then the
// footprint-size class, then the NESTED sub-cell at the fine nside nf =
// nh 2^sub (NESTED indices nest: a cell at nh is the fine index >> 2 sub).
// Within a coarse cell a group holds footprints of one size class from a
// compact patch of sky (Z-order); where a cell holds fewer records of a
// class than a group, the group takes the next class of the same cell, so
// it stays within the cell.  Key bits: 2 log2 nf + 8.
__device__ __forceinline__ unsigned radix_key(const DiscRec &r, int nf, int sub)
{
    const int lg = r.est > 0 ? 31 - __clz(r.est) : 0;
    const unsigned bucket = (unsigned)(lg < 6 ? 0 : (lg - 6 > kSizeBuckets - 1 ? kSizeBuckets - 1 : lg - 6));
    const unsigned cell = nest_cell(nf, r.theta, r.phi);
    const unsigned lo = cell & ((1u << (2 * sub)) - 1u);
    return ((((cell >> (2 * sub)) << 4) | bucket) << (2 * sub)) | lo;
}

// Key of the counting sort (the graph-captured batch pipeline, whose
// histogram must stay small): the radix key's three parts with the fine
// cell at 4 nc, i.e. the 4 x 4 NESTED sub-cells of a coarse cell at nc --
// 12 nc^2 x 16 x 16 bins.
__device__ __forceinline__ unsigned count_key(const DiscRec &r, int nc)
{
    return radix_key(r, 4 * nc, 2);
}

// Counting-sort key of every medium-class record and the keys' histogram.
__global__ void k_sort_keys(const DiscRec *__restrict__ recs, const unsigned *__restrict__ counts, int nc,
                            unsigned *__restrict__ keys, unsigned *__restrict__ hist)
{
    const unsigned n = counts[0];
    // (keys are below 12 nc^2 x 256 by construction; the clamp keeps the
    // histogram update in bounds whatever the record holds)
    const unsigned top = 3072u * (unsigned)nc * (unsigned)nc - 1u;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned k = count_key(recs[i], nc);
        const unsigned key = k < top ? k : top;
        keys[i] = key;
        atomicAdd(&hist[key], 1u);
    }
}

// Keys of the stable radix sort: records below the medium count get their
// key, the rest of the nrec slots a key above every record's (they sort
// last), and each slot its own index as the value, so the sorted values
// list the medium records by key and, within a key, in record order.
__global__ void k_sort_keys_radix(const DiscRec *__restrict__ recs, const unsigned *__restrict__ counts, int nf,
                                  int sub, int64_t nrec, unsigned pad_key, unsigned *__restrict__ keys,
                                  unsigned *__restrict__ iota)
{
    const unsigned n = counts[0];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nrec; i += (int64_t)gridDim.x * blockDim.x) {
        // (records always sort before the pad slots: the kernel reads the
        // first n sorted indices as records)
        const unsigned k = i < (int64_t)n ? radix_key(recs[i], nf, sub) : pad_key;
        keys[i] = i < (int64_t)n && k >= pad_key ? pad_key - 1u : k;
        iota[i] = (unsigned)i;
    }
}

// Exclusive scan, three kernels: per 1024-element block, block sums, add.
__global__ __launch_bounds__(256) void k_scan_blocks(unsigned *__restrict__ data, int n, unsigned *__restrict__ sums)
{
    __shared__ unsigned s[256];
    const int base = blockIdx.x * 1024 + threadIdx.x * 4;
    unsigned v[4], t = 0;
    for (int q = 0; q < 4; q++) {
        v[q] = base + q < n ? data[base + q] : 0u;
        t += v[q];
    }
    s[threadIdx.x] = t;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const unsigned add = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0u;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    unsigned run = s[threadIdx.x] - t; // exclusive prefix of this thread
    for (int q = 0; q < 4; q++) {
        if (base + q < n) data[base + q] = run;
        run += v[q];
    }
    if (threadIdx.x == 255) sums[blockIdx.x] = s[255];
}

__global__ __launch_bounds__(256) void k_scan_sums(unsigned *__restrict__ sums, int nb)
{
    __shared__ unsigned carry;
    __shared__ unsigned s[256];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b = 0; b < nb; b += 256) {
        const int i = b + threadIdx.x;
        const unsigned v = i < nb ? sums[i] : 0u;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            const unsigned add = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0u;
            __syncthreads();
            s[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < nb) sums[i] = carry + s[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 255) carry += s[255];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_scan_add(unsigned *__restrict__ data, int n, const unsigned *__restrict__ sums)
{
    const int base = blockIdx.x * 1024;
    const unsigned add = sums[blockIdx.x];
    for (int q = threadIdx.x; q < 1024; q += 256)
        if (base + q < n) data[base + q] += add;
}

__global__ void k_sort_scatter(const unsigned *__restrict__ keys, const unsigned *__restrict__ counts,
                               unsigned *__restrict__ offs, unsigned *__restrict__ idx)
{
    const unsigned n = counts[0];
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        idx[atomicAdd(&offs[keys[i]], 1u)] = i;
}

__global__ void k_zero_u32(unsigned *p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// NESTED cell resolution of the radix sort: the largest power of two not
// above nside_c (and kMaxNestNside).
static int nest_nside(int nside_c)
{
    int nc = 1;
    while (2 * nc <= nside_c && 2 * nc <= kMaxNestNside) nc *= 2;
    return nc;
}
static int log2i(int v) { return 31 - __builtin_clz((unsigned)v); }

size_t group_sort_bytes(int64_t n)
{
    size_t bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, bytes, (const unsigned *)nullptr, (unsigned *)nullptr,
                                  (const unsigned *)nullptr, (unsigned *)nullptr, (size_t)(n > 0 ? n : 1), 0,
                                  2 * log2i(kMaxNestNside) + 8) != hipSuccess)
        return 0;
    return bytes;
}

hipError_t launch_scan_u32(unsigned *data, int n, unsigned *sums, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const int nb = (n + 1023) / 1024;
    k_scan_blocks<<<nb, 256, 0, st>>>(data, n, sums);
    k_scan_sums<<<1, 256, 0, st>>>(sums, nb);
    k_scan_add<<<nb, 256, 0, st>>>(data, n, sums);
    return hipGetLastError();
}

hipError_t launch_group(int acc, const DiscRec *recs, unsigned *counts, int nside, double pad, const GroupCfg &gc,
                        MapRef map,
                        unsigned long long *stats, hipStream_t st, hipEvent_t sorted, int64_t nrec)
{
    k_zero_u32<<<1, 64, 0, st>>>(gc.work, 1);
    if (gc.sort_tmp) {
        // stable radix sort of all nrec slots (the medium count is known on
        // the device only): medium records by key, then by record order.
        // At 1.1e9 records: 35 ms, against 100-150 ms for the counting sort,
        // whose two returning atomics per record are bound by the L2's
        // atomic rate.
        // Fine cells at the map's nside (at most kMaxNestNside), coarse
        // cells at nside_c; keys of bits = 2 log2 nf + 8 bits, below the
        // pad key (all ones): coarse cells < 12 nh^2 <= 2^(2 log2 nh + 4) -
        // 4 nh^2.
        const int nf = nest_nside(nside), nh = nest_nside(gc.nside_c < nf ? gc.nside_c : nf);
        const int sub = log2i(nf) - log2i(nh), bits = 2 * log2i(nf) + 8;
        const unsigned pad_key = bits >= 32 ? 0xffffffffu : (1u << bits) - 1u;
        k_sort_keys_radix<<<gc.sort_blocks, 256, 0, st>>>(recs, counts, nf, sub, nrec, pad_key, gc.keys, gc.iota);
        size_t bytes = gc.sort_tmp_bytes;
        const hipError_t e = rocprim::radix_sort_pairs(gc.sort_tmp, bytes, gc.keys, gc.keys_out, gc.iota, gc.idx,
                                                       (size_t)nrec, 0, bits, st);
        if (e != hipSuccess) return e;
    } else {
        // coarse cells at nc (a power of two, at most kMaxSortNside), fine
        // cells at 4 nc
        const int nc = nest_nside(gc.nside_c < kMaxSortNside ? gc.nside_c : kMaxSortNside);
        const int nbins = 12 * nc * nc * kSizeBuckets * 16;
        const int nb = (nbins + 1023) / 1024;
        k_zero_u32<<<nb < 4096 ? nb : 4096, 256, 0, st>>>(gc.hist, nbins);
        k_sort_keys<<<gc.sort_blocks, 256, 0, st>>>(recs, counts, nc, gc.keys, gc.hist);
        k_scan_blocks<<<nb, 256, 0, st>>>(gc.hist, nbins, gc.sums);
        k_scan_sums<<<1, 256, 0, st>>>(gc.sums, nb);
        k_scan_add<<<nb, 256, 0, st>>>(gc.hist, nbins, gc.sums);
        k_sort_scatter<<<gc.sort_blocks, 256, 0, st>>>(gc.keys, counts, gc.hist, gc.idx);
    }
    if (sorted) (void)hipEventRecord(sorted, st);
    // Groups in NESTED order are compact, and consecutive ones overlap:
    // spread them over the waves (at 1.1e9 particles, nside 2048, K = 32,
    // class-major key: 4868 ms spread against 5183 ms in order; the batched
    // 64-observer workload: 9305 ms spread against 9557 ms in order).
    const unsigned spread = 2654435761u;
    if (acc == kAccComp)
        k_disc_group<kAccComp><<<gc.group_blocks, 64, 0, st>>>(recs, gc.idx, counts, nside, pad, gc.k, map, stats, gc.work,
                                                               gc.tab_f, gc.tab_i, spread);
    else if (acc == kAccFix)
        k_disc_group<kAccFix><<<gc.group_blocks, 64, 0, st>>>(recs, gc.idx, counts, nside, pad, gc.k, map, stats, gc.work,
                                                              gc.tab_f, gc.tab_i, spread);
    else
        k_disc_group<kAccF32><<<gc.group_blocks, 64, 0, st>>>(recs, gc.idx, counts, nside, pad, gc.k, map, stats, gc.work,
                                                              gc.tab_f, gc.tab_i, spread);
    return hipGetLastError();
}

} // namespace dmann
