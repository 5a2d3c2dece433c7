/*
 * dmann.h -- C ABI of the MI355X annihilation-map projector (libdmann.so).
 *
 * Drop-in for the device API of the reference CUDA render path
 * (src/chealpixrender.h:60-64, implemented at src/chealpixrender.cu:33-125):
 *
 *   reference (C++ linkage, module-global state)      this library
 *   ---------------------------------------------      -----------------------------
 *   initializeCUDA(nside, numofparts)   cu:33-80   ->  dmann_create
 *   calculateMapByGPU(renderpart*, n)   cu:83-103  ->  dmann_render / dmann_render_device
 *   getCUDAMap(MAPTYPE* map)            cu:105-120 ->  dmann_get_map
 *   cudaCleaingUp()                     cu:122-125 ->  dmann_destroy
 *
 * Differences by design:
 *  - The caller passes raw 48-byte DMParticle records (src/datatypes.h:46-105),
 *    not host-preprocessed `renderpart`s: the per-particle pre-pass of
 *    src/genmap_cuda_render.cpp:102-167 (filter, distance, flux, angles,
 *    angular radius, sub-pixel rule) runs on the device.
 *  - The map follows the CPU generator's semantics (src/genmap.cpp:22-175):
 *    footprint = pixel centres within 2h/d (HEALPix query_disc), single-pixel
 *    deposit when 2h/d < theta0 or the footprint has < 2 pixels, Gaussian
 *    weight exp(-0.5 (psi/r)^2 / 0.333) normalised over the footprint.
 *  - State lives in an opaque context (re-entrant; one context per map).
 *  - Every entry point returns a status (0 = success) and never exits the
 *    process; dmann_last_error() gives the message.
 *  - Renders accumulate into the context's map across calls, as successive
 *    calculateMapByGPU calls accumulate into d_map.
 */
#ifndef DMANN_H
#define DMANN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 48-byte particle record; field order of DMParticle (src/datatypes.h:48-60). */
typedef struct dmann_particle {
    float mass, dens, hsmooth;
    float posx, posy, posz;
    float velx, vely, velz;
    float sigmav, eps, phi;
} dmann_particle;

/* Accumulation / arithmetic form of a context.
 *
 * DMANN_PREC_FIXED (no counterpart in the reference) gives the same map,
 * bit for bit, every time the same particles are rendered in the same calls:
 * the pre-pass places records in particle order (per-tile counts and their
 * prefix sums instead of atomic appends), the grouped kernel's neighbours
 * come from a stable radix sort, and every deposit -- FP32 weight times
 * flux / norm, or a window's FP32 partial sum -- is rounded to a 64-bit
 * fixed-point map whose integer additions commute.  The map's scale 2^E is
 * set at the first render after a reset so that the kept flux so far sits
 * below 2^58 units (every pixel is at most that flux), read back from the
 * device, so each render call synchronises once; a later render that would
 * pass 2^61 shifts the map down first.  Accuracy is that of the compensated
 * form (FP32 weights and window partial sums, exact accumulation).  Single
 * maps only (not dmann_batch_*, not dmann_render_map_noalloc). */
enum dmann_precision {
    DMANN_PREC_FP32 = 0,        /* FP32 weights, plain FP32 atomic accumulation          */
    DMANN_PREC_COMPENSATED = 1, /* FP32 weights, two-float (hi+lo) compensated accumulation */
    DMANN_PREC_FP64 = 2,        /* FP64 weights as the reference computes them; FP64 map  */
    DMANN_PREC_FIXED = 3        /* FP32 weights, 64-bit fixed-point map: bit-reproducible */
};

/* Status codes. */
enum dmann_status {
    DMANN_OK = 0,
    DMANN_ERR_ARG = 1,      /* bad argument (null pointer, nside, sizes)          */
    DMANN_ERR_CAPACITY = 2, /* more particles than the context was created for   */
    DMANN_ERR_HIP = 3,      /* a HIP runtime call failed                          */
    DMANN_ERR_DOMAIN = 4    /* a kept particle has an undefined direction (d == 0
                               or |cos theta| > 1); the reference exits here
                               (src/anglefuncs.cpp:60-63)                          */
};

/* Per-map constants, computed on the host from Parameters.
 *  opos       params.opos (src/parameters.cpp:121)
 *  rotmatrix  rotmatrix[i][j] row-major (src/parameters.cpp:183-289)
 *  flux_factor  unit_factor of src/flux/flux_ann_nosom.cpp:13-16
 *  mass_cut   (double)(float)min positive particle mass * 1.1; particles with
 *             mass >= mass_cut or dens < 0 are skipped (src/genmap.cpp:52-58,82-85).
 *             Use FLT_MAX*1.1 when no masses are configured.
 *  nside      HEALPix nside, 1..8192 (src/chealpixrender.cu:25 NS_MAX)
 *  precision  enum dmann_precision
 *  flux_kind  enum dmann_flux: the flux plug-in (src/flux/flux_*.cpp, one of
 *             which the reference links into each executable,
 *             src/Makefile:130-160); the terms below feed the plug-ins:
 *  decay_factor  unit factor of src/flux/flux_decay.cpp:11-14
 *  vel_to_cgs    codeunits.velocity_to_cgs (pwave, somv, somv2)
 *  c_in_cgs      natconst.c_in_cgs
 *  cpos          params.cpos, the halo centre (adiabatic)
 *  footprint_pad  radians added to the footprint radius 2r for pixel
 *             membership (weights unchanged): 0, the default, is the CPU
 *             generator's query_disc(2r) (src/genmap.cpp:119); pi/(4 nside)
 *             = theta_pix/2 reproduces the reference CUDA render's cutoff
 *             2r + theta_pix/2 (src/chealpixrender.cu:188).                */
typedef struct dmann_params {
    double opos[3];
    double rotmatrix[9];
    double flux_factor;
    double mass_cut;
    int32_t nside;
    int32_t precision;
    int32_t flux_kind;
    int32_t reserved;
    double decay_factor;
    double vel_to_cgs;
    double c_in_cgs;
    double cpos[3];
    double footprint_pad;
} dmann_params;

/* Flux plug-ins: K rho m / (4 pi d^2) times
 *   ANN_NOSOM      1                                  (flux_ann_nosom.cpp)
 *   ANN_PWAVE      (sigmav v_unit / c)^2              (flux_ann_pwave.cpp)
 *   ANN_SOMV       c / max(v, 1 km/s) / 1e4           (flux_ann_somv.cpp)
 *   ANN_SOMV2      (c / max(v, 5 km/s))^2 / 1e8       (flux_ann_somv2.cpp)
 *   ANN_ADIABATIC  the adiabatic-contraction profile factor squared, of the
 *                  distance to the halo centre        (flux_ann_adiabatic.cpp)
 *   ANN_CORE_NOSOM the halo-core correction squared: the product over the
 *                  halos whose core radius contains the particle of
 *                  0.55 (r / r_core)^-0.9            (flux_ann_core_nosom.cpp,
 *                  halocore/halocore.cpp:47-84); needs dmann_set_halo_cores
 * and DECAY = decay_factor m / (4 pi d^2)             (flux_decay.cpp). */
enum dmann_flux {
    DMANN_FLUX_ANN_NOSOM = 0,
    DMANN_FLUX_ANN_PWAVE = 1,
    DMANN_FLUX_ANN_SOMV = 2,
    DMANN_FLUX_ANN_SOMV2 = 3,
    DMANN_FLUX_DECAY = 4,
    DMANN_FLUX_ANN_ADIABATIC = 5,
    DMANN_FLUX_ANN_CORE_NOSOM = 6
};

/* Counters of the last dmann_get_stats call (cumulative since create/reset). */
typedef struct dmann_stats {
    int64_t kept;          /* particles passing the filter                        */
    int64_t subpixel;      /* deposited whole because 2r < theta0                 */
    int64_t fallback;      /* deposited whole because the footprint had < 2 pixels */
    int64_t extended;      /* spread over their footprint                         */
    int64_t contributions; /* total footprint pixels of extended particles        */
    int64_t large;         /* extended particles rendered by the sky-covering path */
    int64_t large_contributions; /* footprint pixels of those particles           */
} dmann_stats;

typedef struct dmann_ctx dmann_ctx;

/* Reads a reference ini file (key = value; src/parameters.cpp:310-470) and
 * fills *p with its map constants: observer, rotation matrix
 * (src/parameters.cpp:111-289), the nosom flux factor
 * (src/flux/flux_ann_nosom.cpp:13-16), the low-resolution mass cut
 * (src/genmap.cpp:52-58) and nside; `precision` is copied.  mem_buffer and
 * test_num (may be NULL) receive the ini's mem_buffer and test values.
 * Returns DMANN_ERR_ARG when the file cannot be read (where the reference's
 * Parameters constructor exits, src/parameters.cpp:471-474). */
int dmann_params_from_ini(const char *ini, int precision, dmann_params *p, int *mem_buffer, int *test_num);

/* Creates a context on HIP device `device` for maps of p->nside, accepting up
 * to max_parts particles per render call.  The map starts at zero.
 * Replaces initializeCUDA (src/chealpixrender.cu:33-80). */
int dmann_create(dmann_ctx **ctx, const dmann_params *p, int64_t max_parts, int device);

/* Adds n host-resident particles to the map.  Replaces calculateMapByGPU
 * (src/chealpixrender.cu:83-103).  The particles are copied into pinned
 * staging memory before the call returns (the caller may reuse its buffer
 * at once) and rendered in launches of 2^22 particles: small calls -- a
 * reader's mem_buffer chunks -- are gathered, larger ones split, and the
 * host-to-device copy of one launch overlaps the render of the previous
 * one.  A partly filled launch is rendered by the next call that needs
 * the map or orders work (dmann_synchronize, dmann_get_map,
 * dmann_get_map_device, dmann_get_stats, dmann_render_device,
 * dmann_timing, dmann_set_stream, dmann_set_grouping,
 * dmann_set_halo_cores); dmann_reset drops particles not yet rendered
 * together with the map.  Errors of the launches surface at those calls. */
int dmann_render(dmann_ctx *ctx, const dmann_particle *host_parts, int64_t n);

/* Same for particles already resident in device memory.  Asynchronous on the
 * context's stream; the pointer must stay valid until the next synchronising
 * call (dmann_get_map, dmann_synchronize). */
int dmann_render_device(dmann_ctx *ctx, const dmann_particle *dev_parts, int64_t n);

/* Launch shape and size-class threshold (no counterpart in the reference):
 * persistent workgroups per CU for the wavefront (medium) and workgroup
 * (large) footprint kernels, and the estimated footprint size above which a
 * particle goes to the workgroup kernel.  Defaults 8, 4, 4194304. */
int dmann_set_tuning(dmann_ctx *ctx, int medium_blocks_per_cu, int large_blocks_per_cu, int large_pix);

/* Sub-pixel particles (no counterpart in the reference; the rule is
 * src/genmap.cpp:106-110): on = 1 sorts each wavefront's sub-pixel
 * deposits by pixel and adds one FP64 sum per distinct pixel; 0 adds one
 * atomic per particle.  Default 1. */
int dmann_set_subpixel_sort(dmann_ctx *ctx, int on);

/* Sky-covering class (no counterpart in the reference): blocks_per_cu >= 1
 * renders it as a gather -- a normalisation pass per particle, then one
 * wavefront per HEALPix ring owning that ring's pixels, accumulating up to
 * 64 particles at a time in LDS and adding them into the map without
 * atomics -- with that many single-wavefront workgroups per CU; 0 renders
 * it with one workgroup per particle and per-pixel atomics.  Default 12.
 * The gather keeps each particle's ring range in the grouping's key
 * buffer: with grouping off (dmann_set_grouping K = 0), and in the FP64
 * form, the per-particle kernel is used. */
int dmann_set_sky_gather(dmann_ctx *ctx, int blocks_per_cu);

/* Grouped rendering of the medium class (no counterpart in the reference):
 * particles are sorted by coarse NESTED cell at the largest power of two
 * not above nside_c (1..min(nside, 4096)), then by footprint-size class,
 * then by NESTED cell at the map's nside, and K (1..64) neighbours share
 * one wavefront and an LDS window, cutting global atomics ~K-fold; k = 0
 * renders one particle per wavefront with per-pixel atomics.  Defaults
 * K = 64 (32 when a render has fewer than 8 particles per pixel) and
 * nside_c = min(nside, 32); a K set here holds for every render.  Ignored
 * by the FP64 form. */
int dmann_set_grouping(dmann_ctx *ctx, int k, int nside_c);

/* Uses `stream` (a hipStream_t, NULL = the null stream) for later calls. */
int dmann_set_stream(dmann_ctx *ctx, void *stream);

/* Waits for the context's queued work.  Reports DMANN_ERR_DOMAIN when a kept
 * particle had an undefined direction. */
int dmann_synchronize(dmann_ctx *ctx);

/* Adjoint of the map with respect to the particles (no counterpart in the
 * reference, which has no gradients).  Given dev_grad_map = dL/dmap for the
 * finished map (npix doubles on the device, the map after division by
 * dOmega), writes dev_grad_parts[6*i + {0..5}] = dL/d(mass, dens, hsmooth,
 * posx, posy, posz) of particle i (FP64, device).  The footprint is recomputed
 * with the forward's ring walk; weights and derivatives are FP64.  Pixel
 * membership of the footprint is piecewise constant: its jumps are not part of
 * the gradient.  Every flux plug-in except ANN_ADIABATIC (DMANN_ERR_ARG).
 * Asynchronous on the context's stream. */
int dmann_adjoint_device(dmann_ctx *ctx, const dmann_particle *dev_parts, int64_t n, const double *dev_grad_map,
                         double *dev_grad_parts);

/* Per-kernel timing with HIP events on the context's stream.  Reads the
 * summed durations (ms) of the kernels of the render calls made since the
 * last call -- ms[0] pre-pass, ms[1] sky-covering kernel, ms[2] sort of the
 * medium records by coarse cell, ms[3] medium footprint kernel -- and their
 * number (both may be NULL), then turns event recording on (enable != 0) or
 * off for the following render calls.  A measurement aid with no
 * counterpart in the reference. */
int dmann_timing(dmann_ctx *ctx, int enable, double *ms, int *calls);

/* Writes the finished map, RING order, npix doubles, already divided by
 * dOmega = 4 pi / npix (src/genmap.cpp:164-169), to host memory.
 * Replaces getCUDAMap (src/chealpixrender.cu:105-120) plus the host combine
 * of src/genmap_cuda_render.cpp:191-196. */
int dmann_get_map(dmann_ctx *ctx, double *host_map);

/* Same into device memory (npix doubles); asynchronous on the stream. */
int dmann_get_map_device(dmann_ctx *ctx, double *dev_map);

/* Sets the map and the counters back to zero. */
int dmann_reset(dmann_ctx *ctx);

int dmann_get_stats(dmann_ctx *ctx, dmann_stats *st);

/* Number of pixels of the context's map (12 nside^2). */
int64_t dmann_npix(const dmann_ctx *ctx);

/* Frees everything.  Replaces cudaCleaingUp (src/chealpixrender.cu:122-125). */
void dmann_destroy(dmann_ctx *ctx);

/* ---- batched maps (no counterpart in the reference, which renders one
 * observer per run; for callers that need many observers / resolutions of
 * one particle set) ----
 * All maps of a batch share the particles and the precision; each has its
 * own observer, rotation and nside (params[i]).  A render computes every map
 * afresh -- clears, the footprint work, FP64 finalisation -- and with
 * use_graph != 0 it is captured once in a hipGraph per (particle pointer, n)
 * and replayed.  The footprint work is, by default, the single-map pipeline
 * (pre-pass, sky-covering kernel, sort, grouped kernel) map after map on
 * buffers the maps share (records for n particles, allocated by the first
 * render with that n); or one persistent launch over (map, particle-chunk)
 * work items with per-pixel atomics (dmann_batch_set_tuning, chunk >= 1). */
typedef struct dmann_batch dmann_batch;
int dmann_batch_create(dmann_batch **b, const dmann_params *params, int nmaps, int device);
/* chunk 0: the per-map pipeline (default); chunk 1..2^20: the persistent
 * launch with that many particles per work item and blocks_per_cu (1..64)
 * persistent workgroups per CU */
int dmann_batch_set_tuning(dmann_batch *b, int chunk, int blocks_per_cu);
int dmann_batch_render_device(dmann_batch *b, const dmann_particle *dev_parts, int64_t n, int use_graph);
/* The caller's stream (a hipStream_t; NULL = the null stream, the default):
 * each render waits for the work queued on it before reading the particles,
 * and work queued on it after the render waits for the maps.  The render
 * itself (and its graph) runs on the batch's own stream. */
int dmann_batch_set_stream(dmann_batch *b, void *stream);
/* Waits for the batch's work.  Reports DMANN_ERR_DOMAIN when a kept particle
 * had an undefined direction for some map's observer. */
int dmann_batch_synchronize(dmann_batch *b);
int64_t dmann_batch_npix(const dmann_batch *b, int i);
/* device pointer of finished map i (npix doubles, / dOmega), valid until destroy */
int dmann_batch_map_device(dmann_batch *b, int i, double **dev_map);
int dmann_batch_get_map(dmann_batch *b, int i, double *host_map);
int dmann_batch_get_stats(dmann_batch *b, int i, dmann_stats *st);
long long dmann_batch_graph_replays(const dmann_batch *b);
void dmann_batch_destroy(dmann_batch *b);
const char *dmann_batch_last_error(const dmann_batch *b);

/* HEALPix FITS map file, as MapGenerator::saveToFits writes it
 * (src/mapgenerator.cpp:30-62): RING order, one float64 column.
 * Writes npix doubles of `map` for `nside`. */
int dmann_write_healpix_fits(const char *path, const double *map, int64_t npix, int nside);
/* Reads such a file; with map == NULL only *nside / *npix are filled. */
int dmann_read_healpix_fits(const char *path, double *map, int64_t cap, int *nside, int64_t *npix);

/* Halo cores for DMANN_FLUX_ANN_CORE_NOSOM (replaces the reference's
 * HaloCore, src/halocore/halocore.cpp:17-41): n rows of (xc, yc, zc,
 * radius) as in its halo file; the core radius is radius / (0.73 * 40000)
 * in code length units.  Copied to the device with a uniform grid over the
 * cores, so each particle visits only the cores of its grid cell; the
 * correction multiplies the containing cores' factors in row order, as the
 * reference's loop over all rows does.  Replaces any earlier table. */
int dmann_set_halo_cores(dmann_ctx *ctx, const float *halos, int64_t n);

/* Parses a halo file (whitespace-separated xc yc zc radius rows, read as
 * floats in order, as src/halocore/halocore.cpp:17-41 reads it) into out
 * (4 floats per row, up to cap rows); returns the row count (-1 if the file
 * cannot be read). */
int64_t dmann_read_halo_file(const char *path, float *out, int64_t cap);

/* Reads the particles the ini file names -- the native file (`datafile`) or
 * a tipsy file with its smooth side files (`partfile`, `densfile`,
 * `hsmfile`, `vdspfile`) -- with the ini's `test` truncation and mask, in
 * chunks of `mem_buffer`, exactly as the map generator reads them
 * (src/datareader.cpp:23-298; the reference's pywrapper, src/pywrapper/,
 * exposed this reader).  Copies up to cap records into out (may be NULL)
 * and returns the number of particles, or -DMANN_ERR_ARG when a file is
 * missing or malformed, -DMANN_ERR_DOMAIN when it ends early. */
int64_t dmann_read_particles(const char *ini, dmann_particle *out, int64_t cap);

/* Host-side footprint of one particle, computed with the same code the
 * kernels run (no GPU involved; medium-class particles with the grouped
 * kernel's ring terms and weight series, the others with the sky-covering
 * kernel's): pixels (RING) and unnormalised FP32 weights of the particle's
 * footprint in the kernels' ring order, up to cap entries.
 * *kind: 0 filtered, 1 sub-pixel (whole flux to pix[0]), 2 footprint of < 2
 * pixels (whole flux to pix[0]), 3 spread.  Returns the footprint size (1
 * for kinds 1 and 2), 0 when filtered, < 0 on bad input or an undefined
 * direction.  A debugging and test aid with no counterpart in the reference. */
int64_t dmann_host_footprint(const dmann_params *p, const dmann_particle *q, int32_t *pix, float *w,
                             int64_t cap, double *flux, int *kind);

/* Host-side check of the medium-class ring walk: the number of rings of a
 * medium-class particle where the fast ring setup (FP32 half-width with an
 * FP64 recomputation near column boundaries, FP32 column step) differs from
 * the exact FP64 setup -- in a column decision, or in an FP32 weight term by
 * more than a few units of rounding (0 expected); -2 when the particle is
 * not in the medium class.  A test aid with no counterpart in the
 * reference. */
int64_t dmann_host_ring_check(const dmann_params *p, const dmann_particle *q);

/* Host-side check of the sky-covering gather's indexing: for a particle of
 * the sky-covering class, the pixels and FP32 weights that the per-ring
 * gather visits (ring range, ring position, 2048-column segments, slots)
 * against those of the per-particle ring walk.  Returns the number of
 * differences (0 expected), -2 when the particle is not of that class, -1
 * for bad arguments. */
int64_t dmann_host_sky_check(const dmann_params *p, const dmann_particle *q);

/* Test hook (no counterpart in the reference): the HEALPix NESTED index at
 * nside (a power of two, 1..4096) of n directions (theta, phi), as the
 * grouping's sort computes it on the device, into cells.  Returns n, or -1
 * for bad arguments. */
int64_t dmann_host_nest_cells(int nside, const double *theta, const double *phi, int64_t n, uint32_t *cells);

/* Allocation-free render.  No counterpart in the reference: for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  Renders n particles
 * (device memory) for the map constants *p into dev_map (npix = 12 nside^2
 * doubles of device memory, the caller's): the map is cleared first, and on
 * return holds the finished FP64 RING map divided by dOmega, as
 * dmann_get_map_device gives it.  Accumulation is FP64, in dev_map itself
 * (FP32 weights for the FP32 and compensated precisions, FP64 weights for
 * FP64); the sort-free per-particle kernels are used.  No device memory is
 * allocated or freed: the per-chunk particle records and the counters live
 * in a fixed 48 MiB + 256 B device array that each GPU holds from its first
 * call until the process ends; particles pass through it in chunks of 2^20.
 * Runs on the current device, asynchronously on `stream` (a hipStream_t,
 * NULL = the null stream).  Calls that overlap on one GPU race on that
 * array: serialise them (one stream, or events).  With host_stats != NULL
 * the call waits for the render, fills *host_stats and returns
 * DMANN_ERR_DOMAIN when a kept particle had an undefined direction; with
 * NULL it only enqueues (such particles are skipped).  The halo-core flux
 * plug-in is not available here (DMANN_ERR_ARG).  Errors: dmann_last_error(NULL). */
int dmann_render_map_noalloc(const dmann_params *p, const dmann_particle *dev_parts, int64_t n, double *dev_map,
                             void *stream, dmann_stats *host_stats);

/* Message of the last failed call on ctx (or, when ctx is NULL, of the last
 * failed dmann_create or dmann_render_map_noalloc). */
const char *dmann_last_error(const dmann_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* DMANN_H */
