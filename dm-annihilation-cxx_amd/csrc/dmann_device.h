// Device helpers shared by the kernel files: wavefront reductions and
// scans, wave-local LDS ordering, and the map accumulation forms.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dmann_kernels.h"

namespace dmann {

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ int wave_incl_scan(int v, int lane)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// ---- full-wavefront scans and reductions with DPP moves (row shifts within
// 16-lane rows, then row broadcasts): no LDS round trips.  Every lane of the
// wavefront must be active.

struct DppAdd {
    template <class T> __device__ static T op(T a, T b) { return a + b; }
};
struct DppMin {
    template <class T> __device__ static T op(T a, T b) { return a < b ? a : b; }
};
struct DppMax {
    template <class T> __device__ static T op(T a, T b) { return a > b ? a : b; }
};

template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_move(int v, int ident)
{
    return __builtin_amdgcn_update_dpp(ident, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v, double ident)
{
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(ident), __double2loint(v), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(ident), __double2hiint(v), CTRL, ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Inclusive scan of OP over the 64 lanes; `ident` is OP's identity.
template <class OP, class T>
__device__ __forceinline__ T wave_scan_full(T v, T ident)
{
    v = OP::op(v, dpp_move<0x111, 0xf>(v, ident)); // row_shr:1
    v = OP::op(v, dpp_move<0x112, 0xf>(v, ident)); // row_shr:2
    v = OP::op(v, dpp_move<0x114, 0xf>(v, ident)); // row_shr:4
    v = OP::op(v, dpp_move<0x118, 0xf>(v, ident)); // row_shr:8
    v = OP::op(v, dpp_move<0x142, 0xa>(v, ident)); // row_bcast:15 into rows 1, 3
    v = OP::op(v, dpp_move<0x143, 0xc>(v, ident)); // row_bcast:31 into rows 2, 3
    return v;
}

__device__ __forceinline__ int wave_read(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ float wave_read(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double wave_read(double v, int lane)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// OP over all 64 lanes, the same value in every lane.
template <class OP, class T>
__device__ __forceinline__ T wave_reduce_full(T v, T ident)
{
    return wave_read(wave_scan_full<OP>(v, ident), 63);
}

// Adds one value to the map in the context's accumulation form.
template <int ACC>
__device__ __forceinline__ void acc_add(const MapRef &m, int addr, float v32, double v64)
{
    if constexpr (ACC == kAccF32) {
        unsafeAtomicAdd(m.f32 + addr, v32);
    } else if constexpr (ACC == kAccComp) {
        // two-float accumulation: hi by a returning atomic, then the exact
        // rounding error of that addition (TwoSum) into lo.
        float old = unsafeAtomicAdd(m.f32 + addr, v32);
        float s = old + v32;
        float bb = s - old;
        float err = (old - (s - bb)) + (v32 - bb);
        if (err != 0.0f) unsafeAtomicAdd(m.lo + addr, err);
    } else if constexpr (ACC == kAccW64) {
        unsafeAtomicAdd(m.f64 + addr, (double)v32);
    } else if constexpr (ACC == kAccFix) {
        atomicAdd(m.fix + addr, (unsigned long long)__double2ll_rn((double)v32 * m.fix_scale));
    } else {
        unsafeAtomicAdd(m.f64 + addr, v64);
    }
}

// Whole-particle deposit (sub-pixel rule and the npix_disc<2 fallback,
// src/genmap.cpp:107-110,125-128).
template <int ACC>
__device__ __forceinline__ void acc_point(const MapRef &m, int addr, double flux)
{
    if constexpr (ACC == kAccComp) {
        float f = (float)flux;
        float rem = (float)(flux - (double)f);
        float old = unsafeAtomicAdd(m.f32 + addr, f);
        float s = old + f;
        float bb = s - old;
        float err = (old - (s - bb)) + (f - bb) + rem;
        if (err != 0.0f) unsafeAtomicAdd(m.lo + addr, err);
    } else if constexpr (ACC == kAccW64) {
        unsafeAtomicAdd(m.f64 + addr, flux);
    } else if constexpr (ACC == kAccFix) {
        atomicAdd(m.fix + addr, (unsigned long long)__double2ll_rn(flux * m.fix_scale));
    } else {
        acc_add<ACC>(m, addr, (float)flux, flux);
    }
}

// Sub-pixel deposits of one wavefront (src/genmap.cpp:106-110: the whole
// flux to the RING pixel of the direction), reduced before they reach the
// map: the lanes' (pixel, flux) pairs are sorted by pixel across the
// wavefront (bitonic network on the unique keys pixel << 6 | lane, lanes
// without a deposit last), each run of equal pixels is summed in FP64 by a
// segmented scan (runs are contiguous once sorted), and the last lane of a
// run deposits its sum -- one atomic per distinct pixel instead of one per
// particle, which is what keeps many particles in one pixel (a distant
// halo at low nside) from serialising on that pixel's address.  Every lane
// of the wavefront must call it.
template <int ACC>
__device__ __forceinline__ void wave_subpixel_deposit(const MapRef &m, bool has, int pix, double flux)
{
    const int lane = __lane_id();
    uint64_t key = has ? (((uint64_t)(unsigned)pix << 6) | (uint64_t)lane) : ~0ull;
    double val = has ? flux : 0.0;
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            // compare-exchange with lane ^ j: the lower lane of the pair
            // keeps the smaller key in an ascending k-sequence, the larger
            // one in a descending one
            const uint64_t ok = (uint64_t)__shfl_xor((long long)key, j, 64);
            const double ov = __shfl_xor(val, j, 64);
            const bool lower = (lane & j) == 0;
            const bool up = (lane & k) == 0;
            if ((lower == up) ? (ok < key) : (ok > key)) {
                key = ok;
                val = ov;
            }
        }
    }
    const unsigned p = (unsigned)(key >> 6);
    double run = val;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double up = __shfl_up(run, o, 64);
        const unsigned upix = (unsigned)((uint64_t)__shfl_up((long long)key, o, 64) >> 6);
        if (lane >= o && upix == p) run += up;
    }
    const unsigned next = (unsigned)((uint64_t)__shfl_down((long long)key, 1, 64) >> 6);
    if (key != ~0ull && (lane == 63 || next != p)) acc_point<ACC>(m, (int)p, run);
}

// Wave-aggregated append: every lane of the wave must call it.
__device__ __forceinline__ unsigned wave_append(unsigned *counter, bool pred)
{
    uint64_t mask = __ballot(pred);
    if (mask == 0) return 0;
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)mask) - 1;
    const unsigned rank = (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(counter, (unsigned)__popcll(mask));
    base = __shfl(base, leader, 64);
    return base + rank;
}

__device__ __forceinline__ void wave_count(unsigned long long *ctr, unsigned long long v)
{
    // v is per lane; one atomic per wave
    unsigned long long s = v;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (__lane_id() == 0 && s) atomicAdd(ctr, s);
}

} // namespace dmann
