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
    } else {
        acc_add<ACC>(m, addr, (float)flux, flux);
    }
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
