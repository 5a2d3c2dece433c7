// Grouped footprint rendering with wave-private LDS accumulation.
//
// Global FP32 atomics on MI355X retire about one dword per clock per L2
// channel (~2.6e11/s chip-wide) whatever the locality, and LDS atomics are
// no faster; with ~1e12 footprint pixels per map, one atomic per pixel is
// the bottleneck.  Here the medium-class particles are first sorted by a
// coarse sky cell (counting sort), then one wavefront takes K consecutive
// particles (a group of neighbours):
//   1. per particle, the normalisation pass (FP32 weights, FP64 sum);
//   2. the group's window -- the rings irmin..irmax of all members and the
//      union of their longitude ranges -- is walked in bands of rings that fit
//      the wave's LDS buffer; each member deposits its band pixels with a
//      plain LDS read-modify-write (no atomics: within one instruction every
//      lane holds a distinct pixel, and members are processed in turn);
//   3. the band is flushed with one global atomic (form) per non-zero pixel.
// A group's pixels are written ~K times less often than with per-pixel
// atomics.  The window holds FP32 partial sums of at most K deposits per
// pixel (one per member; error <= (K-1) u of the partial sum); the
// compensated form flushes them with the two-float atomic pair (exact
// TwoSum), so the sum over groups -- 1e4-1e5 terms per pixel -- stays
// compensated.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dmann_device.h"
#include "dmann_footprint.h"
#include "dmann_geom.h"
#include "dmann_kernels.h"

namespace dmann {

// LDS budget of one wave: 20 KB, so 7 single-wave blocks fit a 160 KB CU.
#ifndef DMANN_GROUP_CAP
#define DMANN_GROUP_CAP 3936
#endif
#ifndef DMANN_GROUP_SEG
#define DMANN_GROUP_SEG 8
#endif
constexpr int kGroupCapF32 = DMANN_GROUP_CAP; // FP32 window slots
constexpr int kSeg = DMANN_GROUP_SEG;         // pixels per segment

struct GroupLDS {
    // ring table of up to 64 rings of one member
    float4 rf[64];   // A, B, offm, step
    int4 ri[64];     // kc_rel, cnt, first segment, end segment
    int4 rs[64];     // window slot offset, window columns, column offset in the window ring, nr
    int2 rd[64];     // col0, first pixel of the ring (direct deposits)
    // band table
    int boff[68];    // band slot prefix (boff[j]..boff[j+1] for band ring j)
    int bcl[64];     // band ring: first window column, wrapped into [0, nr)
    int bnr[64];     // band ring: pixels in ring
    int bstart[64];  // band ring: first pixel index
    float acc[kGroupCapF32];
};

__device__ __forceinline__ double wave_min_d(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_max_d(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_min_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Sets up the ring table for rings iz0 .. iz0+n-1 (n <= 64) of a member:
// lane q handles ring iz0 + q.  With band_ring0 >= 0 the window slot terms
// of the band are filled too.  Returns the number of segments.
__device__ __forceinline__ int ring_table(const PGeo &g, int nside, int iz0, int n, GroupLDS &S, int lane,
                                          int band_ring0)
{
    const bool on = lane < n;
    RingLite rl;
    int cnt = 0;
    if (on) {
        rl = ring_lite(g, nside, iz0 + lane);
        cnt = rl.cnt;
    }
    const int nseg = (cnt + kSeg - 1) / kSeg;
    const int incl = wave_incl_scan(nseg, lane);
    const int NS = __shfl(incl, 63, 64);
    if (on) {
        S.rf[lane] = make_float4(rl.A, rl.B, rl.offm, rl.step);
        S.ri[lane] = make_int4(rl.kc_rel, cnt, incl - nseg, incl);
        S.rd[lane] = make_int2(rl.col0, rl.startpix);
        if (band_ring0 >= 0) {
            // slot terms: the band ring owns ncols + kSeg slots from boff;
            // a segment may fill slots off .. off+len-1 below the limit.
            // A whole-ring window takes the ring's wrap into the kSeg spare
            // slots, which needs cnt + kSeg <= nr to keep the masked tail of
            // the last segment off the first one (else limit 0: direct).
            const int rel = iz0 + lane - band_ring0;
            int off = (rl.col0 - S.bcl[rel]) % rl.nr;
            if (off < 0) off += rl.nr;
            const int ncols = S.boff[rel + 1] - S.boff[rel] - kSeg;
            const int lim = ncols < rl.nr ? ncols : (cnt + kSeg <= rl.nr ? rl.nr + kSeg : 0);
            S.rs[lane] = make_int4(S.boff[rel], lim, off, rl.nr);
        }
    } else {
        S.ri[lane] = make_int4(0, 0, incl, incl);
    }
    wave_sync();
    return NS;
}

// First ring q (of n) whose segment range contains segment s.
__device__ __forceinline__ int find_ring(const GroupLDS &S, int s)
{
    int q = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1)
        if (S.ri[q + step - 1].w <= s) q += step;
    return q;
}

// Deposits pixels t0 .. t0+len-1 of table ring q straight into the map
// (global atomics), for bands whose window does not fit the LDS buffer.
template <int ACC>
__device__ __forceinline__ void deposit_direct(const GroupLDS &S, int q, int t0, int len, float x0, float step,
                                               float A, float B, float K4, float sc, const MapRef &map)
{
    const int2 d = S.rd[q];
    const int nr = S.rs[q].w;
#pragma unroll
    for (int t = 0; t < kSeg; ++t) {
        if (t < len) {
            const float v = weight32<0>(fmaf((float)t, step, x0), A, B, 0.0f, 0.0f, K4) * sc;
            int col = d.x + t0 + t;
            if (col < 0) col += nr;
            else if (col >= nr) col -= nr;
            acc_add<ACC>(map, d.y + col, v, 0.0);
        }
    }
}

// Walks the table's segments: lane L takes a contiguous run of segments
// (NS/64 each), kSeg pixels per segment with the ring terms loaded once per
// segment.  Every lane evaluates all kSeg pixels of its segment; those past
// the ring's end are masked to zero, so the loop has no divergent branch.
// MODE 0 sums the weights (returns the lane's sum).  MODE 1 deposits
// weight * sc: into the LDS window (WINDOW), where each band ring owns
// window-columns + kSeg slots, so the masked tail of a ring's last segment
// lands in slots no other lane touches; a segment that would leave its
// ring's slots (which the window construction rules out) goes to global
// atomics instead.  Without WINDOW every deposit is a global atomic.
template <int MODE, int ACC, bool WINDOW>
__device__ __forceinline__ double walk_segments(GroupLDS &S, int NS, int lane, float K4, float sc, const MapRef &map)
{
    const int s0 = (lane * NS) >> 6, s1 = ((lane + 1) * NS) >> 6;
    const int nmax = (NS + 63) >> 6;
    double total = 0.0;
    float part = 0.0f;
    int q = s0 < s1 ? find_ring(S, s0) : 0;
    for (int i = 0; i < nmax; ++i) {
        const int s = s0 + i;
        if (s < s1) {
            while (S.ri[q].w <= s) ++q;
            const float4 f = S.rf[q];
            const int4 r = S.ri[q];
            const int t0 = (s - r.z) * kSeg;
            const int len = r.y - t0 < kSeg ? r.y - t0 : kSeg;
            const float x0 = fmaf((float)(t0 - r.x), f.w, f.z);
            if constexpr (MODE == 0) {
#pragma unroll
                for (int t = 0; t < kSeg; ++t) {
                    const float w = weight32<0>(fmaf((float)t, f.w, x0), f.x, f.y, 0.0f, 0.0f, K4);
                    part += t < len ? w : 0.0f;
                }
            } else if constexpr (WINDOW) {
                const int4 w = S.rs[q];
                int off = w.z + t0;
                if (off >= w.w) off -= w.w;
                if (off + len <= w.y) {
                    float *a = &S.acc[w.x + off];
#pragma unroll
                    for (int t = 0; t < kSeg; ++t) {
                        const float v = weight32<0>(fmaf((float)t, f.w, x0), f.x, f.y, 0.0f, 0.0f, K4) * sc;
                        a[t] += t < len ? v : 0.0f;
                    }
                } else {
                    deposit_direct<ACC>(S, q, t0, len, x0, f.w, f.x, f.y, K4, sc, map);
                }
            } else {
                deposit_direct<ACC>(S, q, t0, len, x0, f.w, f.x, f.y, K4, sc, map);
            }
        }
        if ((i & 3) == 3) {
            total += (double)part;
            part = 0.0f;
        }
    }
    return total + (double)part;
}

// Flushes the band's window slots 0 .. Sb-1: one global atomic (form) per
// non-zero slot.  Slots are taken FB x 64 at a time so that the returning
// atomics of the compensated form are all in flight before the first
// result is waited for.
template <int ACC>
__device__ __forceinline__ void flush_band(GroupLDS &S, int Sb, int lane, const MapRef &map)
{
    constexpr int FB = 4;
    int jb = 0;
    for (int t0 = 0; t0 < Sb; t0 += 64 * FB) {
        int pix[FB];
        float v[FB];
#pragma unroll
        for (int u = 0; u < FB; ++u) {
            const int t = t0 + 64 * u + lane;
            v[u] = 0.0f;
            pix[u] = 0;
            if (t < Sb) {
                while (S.boff[jb + 1] <= t) ++jb;
                const int nrb = S.bnr[jb];
                // slot offsets run to ncols + kSeg - 1 (whole rings: nr + kSeg - 1)
                int col = S.bcl[jb] + (t - S.boff[jb]);
                if (col >= nrb) col -= nrb;
                if (col >= nrb) col -= nrb;
                pix[u] = S.bstart[jb] + col;
                v[u] = S.acc[t];
            }
        }
        if constexpr (ACC == kAccComp) {
            // two-float accumulation: hi by a returning atomic, then the
            // exact rounding error of that addition (TwoSum) into lo
            float old[FB];
#pragma unroll
            for (int u = 0; u < FB; ++u) {
                old[u] = 0.0f;
                if (v[u] != 0.0f) old[u] = unsafeAtomicAdd(map.f32 + pix[u], v[u]);
            }
#pragma unroll
            for (int u = 0; u < FB; ++u) {
                const float sum = old[u] + v[u];
                const float bb = sum - old[u];
                const float err = (old[u] - (sum - bb)) + (v[u] - bb);
                if (err != 0.0f) unsafeAtomicAdd(map.lo + pix[u], err);
            }
        } else {
#pragma unroll
            for (int u = 0; u < FB; ++u)
                if (v[u] != 0.0f) acc_add<ACC>(map, pix[u], v[u], 0.0);
        }
    }
}

template <int ACC>
__global__ __launch_bounds__(64) void k_disc_group(const DiscRec *__restrict__ recs, const unsigned *__restrict__ idx,
                                                   const unsigned *__restrict__ counts, int nside, int K, MapRef map,
                                                   unsigned long long *__restrict__ stats, unsigned *__restrict__ work)
{
    __shared__ GroupLDS S;
    constexpr int CAP = kGroupCapF32;
    const int lane = threadIdx.x;
    const unsigned count = counts[0];
    const unsigned ngroups = (count + (unsigned)K - 1u) / (unsigned)K;
    unsigned long long n_ext = 0, n_fb = 0, n_contrib = 0;
    // groups are taken from a counter: footprint sizes vary a lot
    for (;;) {
        unsigned grp = 0;
        if (lane == 0) grp = atomicAdd(work, 1u);
        grp = __shfl(grp, 0, 64);
        if (grp >= ngroups) break;
        const unsigned m0 = grp * (unsigned)K;
        const int nm = (int)((m0 + (unsigned)K < count ? m0 + (unsigned)K : count) - m0);
        // ---- phase 1: normalisation of every member
        float my_scale = 0.0f, my_K4 = 0.0f;
        int my_lo = 0x7fffffff, my_hi = -0x7fffffff;
        double my_phi = 0.0, my_dphi = 0.0, my_z0 = 0.0, my_s0 = 0.0, my_xa = 0.0, my_cosr = 0.0;
        bool my_on = false;
        for (int j = 0; j < nm; ++j) {
            const PGeo g = particle_geo(recs[idx[m0 + j]], nside);
            double nacc = 0.0;
            int T = 0;
            for (int g0 = 0; g0 < g.nrings; g0 += 64) {
                const int n = g.nrings - g0 < 64 ? g.nrings - g0 : 64;
                const int NS = ring_table(g, nside, g.irmin + g0, n, S, lane, -1);
                T += lane < n ? S.ri[lane].y : 0;
                nacc += walk_segments<0, ACC, false>(S, NS, lane, g.K4, 0.0f, map);
                wave_sync();
            }
            const double norm = wave_sum(nacc);
            T = (int)wave_sum((double)T);
            if (T < 2) {
                if (lane == 0) acc_point<ACC>(map, g.p0, g.flux);
                n_fb++;
            } else {
                n_ext++;
                n_contrib += (unsigned long long)T;
                if (lane == j) {
                    my_scale = (float)(g.flux / norm);
                    my_lo = g.irmin;
                    my_hi = g.irmax;
                    my_phi = g.phi;
                    my_dphi = asin(fmin(1.0, sin(g.rad) / g.s0));
                    my_z0 = g.z0;
                    my_s0 = g.s0;
                    my_xa = g.xa;
                    my_cosr = g.cosr;
                    my_K4 = g.K4;
                    my_on = true;
                }
            }
        }
        // ---- the group's window
        const int R0 = wave_min_i(my_on ? my_lo : 0x7fffffff);
        const int R1 = wave_max_i(my_on ? my_hi : -0x7fffffff);
        if (R0 > R1) continue;
        const uint64_t onmask = __ballot(my_on);
        const int lead = __ffsll((unsigned long long)onmask) - 1;
        const double phic = __shfl(my_phi, lead, 64);
        double dl = my_phi - phic;
        if (dl > kPi) dl -= kTwoPi;
        if (dl < -kPi) dl += kTwoPi;
        const double wlo = wave_min_d(my_on ? dl - my_dphi : 1e30);
        const double whi = wave_max_d(my_on ? dl + my_dphi : -1e30);
        const bool full = (whi - wlo) >= kTwoPi - 1e-9;
        // ---- phase 2: bands of rings that fit the LDS buffer
        for (int b0 = R0; b0 <= R1;) {
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
            const int incl = wave_incl_scan(vr ? ncols + kSeg : 0, lane);
            const bool fits = vr && incl <= CAP;
            int BH = __popcll(__ballot(fits));
            const bool direct = BH == 0; // one window ring exceeds the buffer
            if (direct) BH = (R1 - b0 + 1) < 64 ? (R1 - b0 + 1) : 64;
            int Sb = 0;
            if (!direct) {
                if (lane < BH) {
                    int clw = cl % nr;
                    if (clw < 0) clw += nr;
                    S.boff[lane + 1] = incl;
                    S.bcl[lane] = clw;
                    S.bnr[lane] = nr;
                    S.bstart[lane] = startpix;
                }
                if (lane == 0) S.boff[0] = 0;
                Sb = __shfl(incl, BH - 1, 64);
                for (int t = lane; t < Sb; t += 64) S.acc[t] = 0.0f;
            } else {
                // direct deposits still need the band table for slot terms
                if (lane < BH) {
                    S.boff[lane + 1] = 0;
                    S.bcl[lane] = 0;
                }
                if (lane == 0) S.boff[0] = 0;
            }
            wave_sync();
            const int b1 = b0 + BH - 1;
            for (int j = 0; j < nm; ++j) {
                const int lo = __shfl(my_lo, j, 64), hi = __shfl(my_hi, j, 64);
                if (!__shfl((int)my_on, j, 64) || hi < b0 || lo > b1) continue;
                const float sc = __shfl(my_scale, j, 64);
                // the member's ring-walk terms, from its lane (computed in phase 1)
                PGeo g;
                g.phi = __shfl(my_phi, j, 64);
                g.z0 = __shfl(my_z0, j, 64);
                g.s0 = __shfl(my_s0, j, 64);
                g.xa = __shfl(my_xa, j, 64);
                g.cosr = __shfl(my_cosr, j, 64);
                g.K4 = __shfl(my_K4, j, 64);
                const int rlo = b0 > lo ? b0 : lo, rhi = b1 < hi ? b1 : hi;
                const int NS = ring_table(g, nside, rlo, rhi - rlo + 1, S, lane, b0);
                if (direct) walk_segments<1, ACC, false>(S, NS, lane, g.K4, sc, map);
                else walk_segments<1, ACC, true>(S, NS, lane, g.K4, sc, map);
                wave_sync();
            }
            if (!direct) {
                flush_band<ACC>(S, Sb, lane, map);
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

// Coarse sky cell (RING pixel at nside_c) of every medium-class record and
// its histogram.
__global__ void k_sort_keys(const DiscRec *__restrict__ recs, const unsigned *__restrict__ counts, int nside_c,
                            unsigned *__restrict__ keys, unsigned *__restrict__ hist)
{
    const unsigned n = counts[0];
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const DiscRec r = recs[i];
        const unsigned key = (unsigned)ang2pix(nside_c, r.theta, r.phi);
        keys[i] = key;
        atomicAdd(&hist[key], 1u);
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

hipError_t launch_group(int acc, const DiscRec *recs, unsigned *counts, int nside, const GroupCfg &gc, MapRef map,
                        unsigned long long *stats, hipStream_t st, hipEvent_t sorted)
{
    k_zero_u32<<<1, 64, 0, st>>>(gc.work, 1);
    const int nbins = 12 * gc.nside_c * gc.nside_c;
    const int nb = (nbins + 1023) / 1024;
    k_zero_u32<<<nb < 4096 ? nb : 4096, 256, 0, st>>>(gc.hist, nbins);
    k_sort_keys<<<gc.sort_blocks, 256, 0, st>>>(recs, counts, gc.nside_c, gc.keys, gc.hist);
    k_scan_blocks<<<nb, 256, 0, st>>>(gc.hist, nbins, gc.sums);
    k_scan_sums<<<1, 256, 0, st>>>(gc.sums, nb);
    k_scan_add<<<nb, 256, 0, st>>>(gc.hist, nbins, gc.sums);
    k_sort_scatter<<<gc.sort_blocks, 256, 0, st>>>(gc.keys, counts, gc.hist, gc.idx);
    if (sorted) (void)hipEventRecord(sorted, st);
    if (acc == kAccComp)
        k_disc_group<kAccComp><<<gc.group_blocks, 64, 0, st>>>(recs, gc.idx, counts, nside, gc.k, map, stats, gc.work);
    else
        k_disc_group<kAccF32><<<gc.group_blocks, 64, 0, st>>>(recs, gc.idx, counts, nside, gc.k, map, stats, gc.work);
    return hipGetLastError();
}

} // namespace dmann
