// Internal interface between the C ABI (dmann_capi.hip) and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dmann.h"

namespace dmann {

// 48-byte particle record (src/datatypes.h:46-105).
struct Particle {
    float mass, dens, hsmooth;
    float posx, posy, posz;
    float velx, vely, velz;
    float sigmav, eps, phi;
};
static_assert(sizeof(Particle) == 48, "DMParticle is 12 floats");

// Accumulation forms: FP32 weights into an FP32 map, into a two-float map,
// FP64 weights into an FP64 map, FP32 weights into an FP64 map (FP64
// atomics; the allocation-free entry point accumulates into its output),
// and FP32 weights into a 64-bit fixed-point map (the reproducible form:
// integer additions commute, so the map does not depend on their order).
enum { kAccF32 = 0, kAccComp = 1, kAccF64 = 2, kAccW64 = 3, kAccFix = 4 };
enum { kClsSkip = 0, kClsPoint = 1, kClsMedium = 2, kClsLarge = 3 };
enum {
    kStatKept = 0,
    kStatSub = 1,
    kStatFallback = 2,
    kStatExt = 3,
    kStatContrib = 4,
    kStatLarge = 5,
    kStatBad = 6,
    kStatLargeContrib = 7,
    kNumStats = 8
};

// Record of one extended particle after the pre-pass (48 bytes).
struct DiscRec {
    double theta, phi, r, flux;
    int p0;  // RING pixel of the direction (fallback deposit)
    int est; // estimated footprint pixels
    int cls;
    int pad;
};
static_assert(sizeof(DiscRec) == 48, "DiscRec is 48 bytes");

// Per-map constants (host-computed).
struct DevParams {
    double opos[3];
    double rot[9];
    double flux_factor;
    double mass_cut;
    double theta0;  // acos(1 - dOmega / (2 * 3.141592)), src/genmap.cpp:38
    double dOmega;
    int nside;
    int npix;
    int large_pix;  // estimated footprint above which the sky-covering path is used
    int flux_kind;  // dmann_flux
    double decay_factor, vel_to_cgs, c_in_cgs;
    double cpos[3];
    double footprint_pad; // radians added to 2r for pixel membership
    int subpixel_sort;    // 1: sub-pixel deposits reduced per wavefront by pixel first
    // halo cores (ANN_CORE_NOSOM): rows (xc, yc, zc, core radius); a uniform
    // grid of cells, each listing the rows whose core meets it (ascending)
    const float4 *halos;
    const int *cell_start;
    const int *cell_list;
    double grid_lo[3];
    double grid_inv; // 1 / cell size
    int grid_dim[3];
    int nhalos;
};

// Device constants of one map from the C-ABI parameters.
void fill_devparams(const dmann_params *p, DevParams &P, int large_pix);

// Message of a failed call without a context (dmann_last_error(NULL)).
void set_global_error(const char *msg);

// Accumulation targets of one map.
struct MapRef {
    float *f32;  // FP32 map, or the hi word of the two-float map
    float *lo;   // lo word of the two-float map
    double *f64; // FP64 map
    unsigned long long *fix; // fixed-point map: value x 2^E as a two's-complement int64
    double fix_scale;        // 2^E
};

// Buffers of the reproducible form (kAccFix): the pre-pass places records
// by per-tile offsets (tiles of 512 particles in particle order) instead of
// atomic appends; the medium class is then sorted by the grouping's stable
// radix sort, so groups are the same in every run.
struct DetCfg {
    unsigned *tile_med;    // ceil(max_parts / 512): per-tile counts, then exclusive offsets
    unsigned *tile_large;
    unsigned *tile_sums;   // scan block sums, ceil(tiles / 1024)
    double *flux_part;     // per pre-pass block FP64 flux sums
    double *flux_total;    // their sum (one double)
};

// Bytes of radix-sort temporary storage for n keys (dmann_group.hip).
size_t group_sort_bytes(int64_t n);

// Sum of the kept particles' fluxes (FP64, fixed reduction order) into
// det.flux_total: a bound on every pixel of the map, which sets its
// fixed-point scale.
hipError_t launch_flux_total(const Particle *d_parts, int64_t n, const DevParams &P, const DetCfg &det, int blocks,
                             hipStream_t st);
// Fixed-point map values shifted right by k bits (a smaller scale).
hipError_t launch_fix_shift(unsigned long long *fix, int64_t npix, int k, hipStream_t st);
// Exclusive scan of n unsigned values in place (sums: ceil(n / 1024)).
hipError_t launch_scan_u32(unsigned *data, int n, unsigned *sums, hipStream_t st);

struct LaunchCfg {
    int prepass_blocks;
    int medium_blocks;
    int large_blocks;
    // single-wave workgroups of the ring-owned gather of the sky-covering
    // class (dmann_sky.hip; FP32 and two-float maps with the grouping's
    // buffers); 0 renders that class with the per-particle kernel
    // k_disc_large
    int sky_blocks;
};

// Grouped rendering of the medium class (dmann_group.hip): records sorted by
// coarse sky cell, then K neighbours per wavefront with LDS accumulation.
// Sort key of a medium record: coarse sky cell, then footprint-size class
// (floor(log2 est) - 6, clamped to 0..15), so that a group of neighbours
// also holds footprints of similar size and its window stays tight.
constexpr int kSizeBuckets = 16;
constexpr int kMaxSortNside = 256;  // counting sort (histogram of 12 nside_c^2 x 256 bins)
constexpr int kMaxNestNside = 4096; // radix sort (class and NESTED cell in 32 key bits)
// Default group size.  The window's FP32 partial sums hold up to K deposits
// per pixel; members deposit in ascending flux order, so small deposits are
// summed before large ones (in arbitrary order a deposit below half an ulp
// of the partial sum was lost).  On 2e8 particles of the 1.1e9 workload at
// nside 2048, compensated, worst pixel / total against the FP64 map:
// K = 32 1.0e-7 / 1.4e-10, K = 64 1.5e-7 / 1.1e-10 (in arbitrary order
// K = 32 gave 1.8e-7 / 5.6e-9); the grouped kernel took 4743 ms at K = 64
// against 4904 ms at K = 32 (1.1e9 particles).
constexpr int kDefaultGroupK = 64;
// Group size of a render of n particles into a map of npix pixels when the
// caller has not chosen one: K = 64 where the particles are dense (at
// least 8 per pixel), else 32.  On 1e7 particles (NFW + 100 subhalos) at
// nside 1024, 0.8 per pixel, K = 32 took 197 ms in the grouped kernel
// against 210 ms at K = 64 (nside 512: 58 against 59 ms); at 1.1e9
// particles, nside 2048, 22 per pixel, K = 64 was 2.5 % faster.
inline int auto_group_k(int64_t n, int64_t npix) { return n >= 8 * npix ? 64 : 32; }
// Default coarse cell resolution of the sort: nside_c = 32 (or the map's
// nside below 32), cells of 1.8 degrees whatever the map's nside.  Grouped
// kernel at 1.1e9 particles, nside 2048, K = 64, coarse nside 256 / 128 /
// 64 / 32 / 16 / 1: 4925 / 4752 / 4641 / 4601 / 4578 / 4854 ms; on 1e7
// particles at nside 256 / 512 / 1024 a coarse nside of 32 was as fast as
// or faster than nside / 32 (20.0 / 58.0 / 196 ms against 20.9 / 58.1 /
// 197 ms).
inline int default_cell_nside(int nside) { return nside < 32 ? nside : 32; }

struct GroupCfg {
    int k;            // particles per group (0: per-particle kernel instead)
    bool k_auto;      // each render takes min(k, auto_group_k(n, npix))
    int nside_c;      // sort cell resolution
    int sort_blocks;
    int group_blocks; // single-wave workgroups
    unsigned *keys;   // max_parts
    unsigned *idx;    // max_parts
    // radix sort (rocPRIM, stable): sorted keys and the record indices fed
    // to it, max_parts each, and its temporary storage; without sort_tmp
    // the counting sort below is used instead (its atomics are graph-safe,
    // and the batch pipeline is captured in a hipGraph)
    unsigned *keys_out;
    unsigned *iota;
    void *sort_tmp;
    size_t sort_tmp_bytes;
    unsigned *hist;   // counting sort: 12 nside_c^2 x kSizeBuckets x 16 bins
    unsigned *sums;   // ceil(bins / 1024)
    unsigned *work;   // group counter
    // per-wave cache of the members' ring terms (group_blocks x k x 64
    // rings), written by the normalisation pass, read back per band
    float4 *tab_f;
    int4 *tab_i;
};

hipError_t launch_group(int acc, const DiscRec *recs, unsigned *counts, int nside, double pad, const GroupCfg &gc,
                        MapRef map,
                        unsigned long long *stats, hipStream_t st,
                        hipEvent_t sorted = nullptr, int64_t nrec = 0);

// Sky-covering class by normalisation per particle, then a gather per
// ring (dmann_sky.hip); acc kAccF32, kAccComp or kAccFix; ranges: one word per
// sky-covering particle (the grouping's key buffer serves).
hipError_t launch_sky(int acc, DiscRec *recs, int64_t n, unsigned *counts, int nside, double pad, MapRef map,
                      unsigned long long *stats, unsigned *ranges, int norm_blocks, int gather_blocks,
                      hipStream_t st);

// ev (optional): 4 events recorded before the pre-pass, after it, after the
// sky-covering kernel and after the medium kernel.
// det: the reproducible form's buffers (required for kAccFix).
hipError_t launch_render(int acc, const Particle *d_parts, int64_t n, const DevParams &P, DiscRec *recs,
                         unsigned *counts, MapRef map, unsigned long long *stats, const LaunchCfg &cfg,
                         hipStream_t st, hipEvent_t *ev = nullptr, const GroupCfg *gc = nullptr,
                         const DetCfg *det = nullptr);
hipError_t launch_adjoint(const Particle *d_parts, int64_t n, const DevParams &P, const double *gmap, double *out,
                          int blocks, hipStream_t st);
hipError_t launch_batch(int acc, const Particle *d_parts, int64_t n, const DevParams *Ps, const MapRef *maps,
                        int nmaps, int chunk, unsigned long long *work, unsigned long long *stats, int blocks,
                        hipStream_t st);
hipError_t launch_fill_u64(unsigned long long *p, int64_t n, hipStream_t st);
hipError_t launch_finalize(int acc, MapRef map, int64_t npix, double dOmega, double *out, hipStream_t st);
hipError_t launch_clear(int acc, MapRef map, int64_t npix, unsigned long long *stats, hipStream_t st);

} // namespace dmann
