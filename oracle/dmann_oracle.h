/*
 * CPU oracle for the annihilation-map projector.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product path (the HIP library and the C++ drop-in)
 * never links, loads or calls it.
 *
 * It is a plain-C, FP64, single-threaded restatement of the reference CPU
 * generator (src/genmap.cpp:22-175) and of the pieces it calls:
 *   - Parameters: ini parser, unit chain and rotation matrix
 *       (src/parameters.cpp:42-82, 111-152, 183-289, 291-483)
 *   - getflux for the "nosom" plug-in (src/flux/flux_ann_nosom.cpp:12-23)
 *   - calc_angles / ang2vec (src/anglefuncs.cpp:15-71)
 *   - the native particle file and mask (src/datareader.cpp:101-272)
 *   - HEALPix C++ 3.00 RING primitives used by genmap.cpp (Healpix_Base::
 *     ang2pix, pix2vec, query_disc; third-party, not in the reference tree),
 *     restated from the published HEALPix algorithm (Gorski et al. 2005).
 * An OpenMP variant with per-thread maps is provided for the N-core CPU
 * baseline; it computes the same per-particle contributions.
 *
 * Parity pinning (each checked in tests/test_oracle_*.py):
 *   - Parameters, getflux, DataReader: known-answer vectors from the
 *     reference's own parameters.cpp / flux_ann_nosom.cpp / datareader.cpp
 *     compiled in this container (oracle/_ref), plus closed forms.
 *   - HEALPix primitives: brute force over all pixels (query_disc), round
 *     trips (ang2pix o pix2vec), ring pixel counts and equal areas.
 *   - genmap: brute-force all-pixel restatement on tiny inputs, the
 *     invariant sum(map)*dOmega == sum(flux), ring symmetry of a polar
 *     particle, and the npix_disc<2 / sub-pixel rules on hand-built cases.
 *   - calc_angles: closed forms on the axes of a known rotation (the
 *     reference anglefuncs.cpp includes HEALPix headers and cannot be
 *     compiled here).
 */
#ifndef DMANN_ORACLE_H
#define DMANN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 48-byte particle record, field order of DMParticle (src/datatypes.h:46-105). */
typedef struct {
    float mass, dens, hsmooth;
    float posx, posy, posz;
    float velx, vely, velz;
    float sigmav, eps, phi;
} orc_particle;

/* Everything the generator reads from Parameters, plus the unit chain. */
typedef struct {
    /* ini values */
    double Omega_M, h100, n_s, sigma_8, Lbox_in_Mpc, mass_to_Msun_in;
    double particle_masses_in[10];
    long particle_numbers_in[10];
    double centerpos[3], centervel[3];
    double align_vec[3], obs_pos[3];
    int nside, test, mem_buffer, is_native, is_mask;
    char datafile[512], outputmap[512], maskfile[512];
    /* derived (parameters.cpp:42-82, 111-152, 183-289) */
    double rho_crit_in_cgs, Delta_vir;
    double mass_to_cgs, mass_to_Msun, length_to_Mpc, length_to_cgs, time_to_cgs;
    double velocity_to_cgs, density_to_cgs, annihilation_flux_to_cgs;
    double opos[3], cpos[3], otheta, ophi;
    double particle_masses[10];
    double rotmatrix[9];           /* row-major rotmatrix[i][j] -> [3*i+j] */
    int npix;
    double dOmega;
    /* flux_ann_nosom.cpp:13-16 unit_factor */
    double flux_factor;
} orc_params;

/* Per-run constants the generator needs (subset of orc_params). */
/* Flux plug-ins (the src/flux/ files, one linked per executable in the
 * reference, src/Makefile:130-160). */
enum {
    ORC_FLUX_ANN_NOSOM = 0, /* flux_ann_nosom.cpp */
    ORC_FLUX_ANN_PWAVE = 1, /* flux_ann_pwave.cpp */
    ORC_FLUX_ANN_SOMV = 2,  /* flux_ann_somv.cpp */
    ORC_FLUX_ANN_SOMV2 = 3, /* flux_ann_somv2.cpp */
    ORC_FLUX_DECAY = 4,     /* flux_decay.cpp */
    ORC_FLUX_ANN_ADIABATIC = 5, /* flux_ann_adiabatic.cpp */
    ORC_FLUX_ANN_CORE_NOSOM = 6 /* flux_ann_core_nosom.cpp + halocore/halocore.cpp */
};

typedef struct {
    double opos[3];
    double rotmatrix[9];
    double flux_factor;
    double mass_cut;     /* hires_particle_mass(float) * 1.1 (genmap.cpp:52-58,82) */
    int nside;
    int flux_kind;       /* ORC_FLUX_* (default: nosom) */
    double decay_factor; /* flux_decay.cpp:11-14 unit factor */
    double vel_to_cgs;   /* codeunits.velocity_to_cgs */
    double c_in_cgs;     /* natconst.c_in_cgs */
    double cpos[3];      /* params.cpos (flux_ann_adiabatic.cpp:16-18) */
    double footprint_pad; /* added to the query_disc radius 2r (0 = genmap.cpp:119;
                             theta_pix/2 = the reference CUDA cutoff,
                             chealpixrender.cu:188) */
    const float *halos;   /* halo-core rows (xc, yc, zc, radius) for ANN_CORE_NOSOM */
    int64_t nhalos;
} orc_config;

typedef struct {
    int64_t kept;          /* particles passing the filter */
    int64_t subpixel;      /* deposited whole via 2r < theta0 */
    int64_t fallback;      /* deposited whole via npix_disc < 2 */
    int64_t extended;      /* spread over >= 2 pixels */
    int64_t contributions; /* total footprint pixels of extended particles */
    double total_flux;     /* sum of flux of kept particles */
} orc_stats;

/* ---- HEALPix RING primitives (HEALPix C++ 3.00 semantics) ---- */
int64_t orc_npix(int nside);
int orc_ring_above(int nside, double z);
double orc_ring2z(int nside, int ring);
void orc_ring_info(int nside, int ring, int64_t *startpix, int64_t *ringpix, int *shifted);
int64_t orc_ang2pix_ring(int nside, double theta, double phi);
void orc_pix2vec_ring(int nside, int64_t pix, double vec[3]);
void orc_pix2zphi_ring(int nside, int64_t pix, double *z, double *phi);
/* Pixels whose centres lie within `radius` of (theta, phi), ascending.
 * Writes at most `cap` entries; returns the total count. */
int64_t orc_query_disc(int nside, double theta, double phi, double radius,
                       int64_t *out, int64_t cap);

/* ---- Parameters ---- */
/* returns 0 on success, -1 if the file cannot be opened */
int orc_params_from_ini(const char *path, orc_params *p);
/* defaults of Parameters::initialize + constructor (parameters.cpp:291-320) */
void orc_params_defaults(orc_params *p);
/* unit chain, params, rotation (parameters.cpp:42-289) */
void orc_params_setup(orc_params *p);
void orc_config_from_params(const orc_params *p, orc_config *c);

/* ---- angles / flux ---- */
double orc_getflux(double flux_factor, const orc_particle *q, double d);
/* HaloCore::getCorrection (halocore.cpp:47-68) over rows (xc, yc, zc,
 * radius), every row in order. */
double orc_core_correction(const float *halos, int64_t n, float x, float y, float z);
/* getflux of the configured plug-in (c->flux_kind). */
double orc_getflux_cfg(const orc_config *c, const orc_particle *q, double d);
void orc_calc_angles(double x, double y, double z, double d, const double rot[9],
                     double *costheta, double *phi);

/* ---- particle file ---- */
/* Reads the native file (int32 N, N x 48 B).  Applies `test` truncation
 * (test != -1) and the mask file (0 byte -> dens = -1).  *out is malloc'd.
 * Returns the particle count, or -1 on error. */
int64_t orc_read_native(const char *path, int test, const char *maskfile,
                        orc_particle **out);
void orc_free(void *p);

/* ---- the generator ---- */
/* Adds the raw (not yet divided by dOmega) contributions of n particles to
 * map[0..npix).  Mirrors genmap.cpp:67-156 for one chunk. */
void orc_genmap(const orc_particle *parts, int64_t n, const orc_config *c,
                double *map, orc_stats *st);
/* Same contributions with nthreads OpenMP threads and per-thread maps. */
void orc_genmap_omp(const orc_particle *parts, int64_t n, const orc_config *c,
                    double *map, orc_stats *st, int nthreads);
/* genmap.cpp:161-169: returns sum(map) and divides map by dOmega. */
double orc_finalize(double *map, int64_t npix, double dOmega);

/* Single-particle footprint with weights, for the gradient checks.
 * Returns the footprint size (pixels and normalised weights w/norm). */
int64_t orc_particle_footprint(const orc_particle *q, const orc_config *c,
                               int64_t *pix, double *wnorm, int64_t cap,
                               double *flux_out);

#ifdef __cplusplus
}
#endif
#endif
