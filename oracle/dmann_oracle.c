/*
 * CPU oracle for the annihilation-map projector -- TEST INFRASTRUCTURE ONLY.
 * See dmann_oracle.h for what it restates and how each part is pinned.
 *
 * Build flags matter: the reference is plain C++ compiled for x86-64 without
 * FMA contraction, so this file is compiled with -ffp-contract=off and no
 * -ffast-math (oracle/Makefile) to reproduce its FP64 operation order.
 */
#include "dmann_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846 /* M_PI, src/datatypes.h:22,34 */

/* unit constants, src/datatypes.cpp:5-10 */
static const double kMsun_in_cgs = 1.9889212e+33;
static const double kMpc_in_cgs = 3.08568025e+24;
static const double keV_in_cgs = 1.602e-12;
static const double kG_in_cgs = 6.67259e-8;
static const double kc_in_cgs = 2.99792458e10;

/* HEALPix constants */
static const double kTwoPi = 6.283185307179586476925286766559005768394;
static const double kHalfPi = 1.570796326794896619231321691639751442099;
static const double kTwoThird = 2.0 / 3.0;

/* ===================================================================== */
/* HEALPix RING primitives.  Restated from the HEALPix C++ 3.00 algorithm */
/* (Healpix_Base: ring_above, ring2z, get_ring_info_small, loc2pix,       */
/* pix2loc, query_disc non-inclusive) as called by genmap.cpp:40,106,     */
/* 119,135.                                                               */
/* ===================================================================== */

int64_t orc_npix(int nside) { return 12LL * nside * nside; }

int orc_ring_above(int nside, double z)
{
    double az = fabs(z);
    if (az <= kTwoThird) return (int)(nside * (2 - 1.5 * z));
    int iring = (int)(nside * sqrt(3 * (1 - az)));
    return (z > 0) ? iring : 4 * nside - iring - 1;
}

double orc_ring2z(int nside, int ring)
{
    double fact2 = 4. / (double)orc_npix(nside);
    double fact1 = (double)(2 * nside) * fact2;
    int64_t r = ring;
    if (r < nside) return 1 - (double)(r * r) * fact2;
    if (r <= 3 * nside) return (double)(2 * nside - r) * fact1;
    r = 4 * nside - r;
    return (double)(r * r) * fact2 - 1;
}

void orc_ring_info(int nside, int ring, int64_t *startpix, int64_t *ringpix, int *shifted)
{
    int64_t r = ring, ns = nside;
    int64_t ncap = 2 * ns * (ns - 1);
    if (r < ns) {
        *shifted = 1;
        *ringpix = 4 * r;
        *startpix = 2 * r * (r - 1);
    } else if (r < 3 * ns) {
        *shifted = ((r - ns) & 1) == 0;
        *ringpix = 4 * ns;
        *startpix = ncap + (r - ns) * 4 * ns;
    } else {
        int64_t nr = 4 * ns - r;
        *shifted = 1;
        *ringpix = 4 * nr;
        *startpix = orc_npix(nside) - 2 * nr * (nr + 1);
    }
}

static double fmodulo(double v1, double v2)
{
    if (v1 >= 0) return (v1 < v2) ? v1 : fmod(v1, v2);
    double tmp = fmod(v1, v2) + v2;
    return (tmp == v2) ? 0. : tmp;
}

static int64_t loc2pix(int nside, double z, double phi, double sth, int have_sth)
{
    int64_t ns = nside;
    double za = fabs(z);
    double tt = fmodulo(phi * (1.0 / kHalfPi), 4.0);
    if (za <= kTwoThird) {
        int64_t nl4 = 4 * ns;
        double temp1 = ns * (0.5 + tt);
        double temp2 = ns * z * 0.75;
        int64_t jp = (int64_t)(temp1 - temp2);
        int64_t jm = (int64_t)(temp1 + temp2);
        int64_t ir = ns + 1 + jp - jm;
        int64_t kshift = 1 - (ir & 1);
        int64_t t1 = jp + jm - ns + kshift + 1 + nl4 + nl4;
        int64_t ip = (t1 >> 1) % nl4;
        return 2 * ns * (ns - 1) + (ir - 1) * nl4 + ip;
    } else {
        double tp = tt - (double)(int64_t)tt;
        double tmp = ((za < 0.99) || (!have_sth)) ? ns * sqrt(3 * (1 - za))
                                                  : ns * sth / sqrt((1. + za) / 3.);
        int64_t jp = (int64_t)(tp * tmp);
        int64_t jm = (int64_t)((1.0 - tp) * tmp);
        int64_t ir = jp + jm + 1;
        int64_t ip = (int64_t)(tt * ir);
        return (z > 0) ? 2 * ir * (ir - 1) + ip : orc_npix(nside) - 2 * ir * (ir + 1) + ip;
    }
}

int64_t orc_ang2pix_ring(int nside, double theta, double phi)
{
    if ((theta < 0.01) || (theta > 3.14159 - 0.01))
        return loc2pix(nside, cos(theta), phi, sin(theta), 1);
    return loc2pix(nside, cos(theta), phi, 0., 0);
}

static int64_t isqrt64(int64_t v) { return (int64_t)sqrt((double)v + 0.5); }

static void pix2loc(int nside, int64_t pix, double *z, double *phi, double *sth, int *have_sth)
{
    int64_t ns = nside;
    int64_t ncap = 2 * ns * (ns - 1);
    int64_t npix = orc_npix(nside);
    double fact2 = 4. / (double)npix;
    double fact1 = (double)(2 * ns) * fact2;
    *have_sth = 0;
    if (pix < ncap) {
        int64_t iring = (1 + isqrt64(1 + 2 * pix)) >> 1;
        int64_t iphi = (pix + 1) - 2 * iring * (iring - 1);
        double tmp = (double)(iring * iring) * fact2;
        *z = 1.0 - tmp;
        if (*z > 0.99) { *sth = sqrt(tmp * (2.0 - tmp)); *have_sth = 1; }
        *phi = (iphi - 0.5) * kHalfPi / iring;
    } else if (pix < (npix - ncap)) {
        int64_t nl4 = 4 * ns;
        int64_t ip = pix - ncap;
        int64_t tmp = ip / nl4;
        int64_t iring = tmp + ns;
        int64_t iphi = ip - nl4 * tmp + 1;
        double fodd = ((iring + ns) & 1) ? 1 : 0.5;
        *z = (double)(2 * ns - iring) * fact1;
        *phi = (iphi - fodd) * ORC_PI * 0.75 * fact1;
    } else {
        int64_t ip = npix - pix;
        int64_t iring = (1 + isqrt64(2 * ip - 1)) >> 1;
        int64_t iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
        double tmp = (double)(iring * iring) * fact2;
        *z = tmp - 1.0;
        if (*z < -0.99) { *sth = sqrt(tmp * (2.0 - tmp)); *have_sth = 1; }
        *phi = (iphi - 0.5) * kHalfPi / iring;
    }
}

void orc_pix2zphi_ring(int nside, int64_t pix, double *z, double *phi)
{
    double sth;
    int have;
    pix2loc(nside, pix, z, phi, &sth, &have);
}

void orc_pix2vec_ring(int nside, int64_t pix, double vec[3])
{
    double z, phi, sth = 0;
    int have;
    pix2loc(nside, pix, &z, &phi, &sth, &have);
    if (have) {
        vec[0] = sth * cos(phi);
        vec[1] = sth * sin(phi);
        vec[2] = z;
    } else {
        double st = sqrt((1. - z) * (1. + z));
        vec[0] = st * cos(phi);
        vec[1] = st * sin(phi);
        vec[2] = z;
    }
}

/* range append that keeps the list ascending (rangeset<I>::append) */
static inline void put_range(int64_t a, int64_t b, int64_t *out, int64_t cap, int64_t *n)
{
    for (int64_t p = a; p < b; ++p) {
        if (*n < cap) out[*n] = p;
        ++*n;
    }
}

int64_t orc_query_disc(int nside, double theta, double phi, double radius,
                       int64_t *out, int64_t cap)
{
    int64_t n = 0;
    int64_t npix = orc_npix(nside);
    /* pointing::normalize: theta already in [0,pi] here; phi into [0,2pi) */
    phi = fmodulo(phi, kTwoPi);

    if (radius >= ORC_PI) {
        put_range(0, npix, out, cap, &n);
        return n;
    }
    double cosrbig = cos(radius);
    double z0 = cos(theta);
    double xa = 1. / sqrt((1 - z0) * (1 + z0));

    double rlat1 = theta - radius;
    double zmax = cos(rlat1);
    int irmin = orc_ring_above(nside, zmax) + 1;
    if ((rlat1 <= 0) && (irmin > 1)) { /* north pole in the disc */
        int64_t sp, rp;
        int sh;
        orc_ring_info(nside, irmin - 1, &sp, &rp, &sh);
        put_range(0, sp + rp, out, cap, &n);
    }
    double rlat2 = theta + radius;
    double zmin = cos(rlat2);
    int irmax = orc_ring_above(nside, zmin);

    for (int iz = irmin; iz <= irmax; ++iz) {
        double z = orc_ring2z(nside, iz);
        double x = (cosrbig - z * z0) * xa;
        double ysq = 1 - z * z - x * x;
        double dphi = (ysq <= 0) ? 0 : atan2(sqrt(ysq), x);
        if (dphi > 0) {
            int64_t ipix1, nr;
            int sh;
            orc_ring_info(nside, iz, &ipix1, &nr, &sh);
            double shift = sh ? 0.5 : 0.;
            int64_t ipix2 = ipix1 + nr - 1;
            int64_t ip_lo = (int64_t)floor((double)nr * (1.0 / kTwoPi) * (phi - dphi) - shift) + 1;
            int64_t ip_hi = (int64_t)floor((double)nr * (1.0 / kTwoPi) * (phi + dphi) - shift);
            if (ip_lo <= ip_hi) {
                if (ip_hi >= nr) { ip_lo -= nr; ip_hi -= nr; }
                if (ip_lo < 0) {
                    put_range(ipix1, ipix1 + ip_hi + 1, out, cap, &n);
                    put_range(ipix1 + ip_lo + nr, ipix2 + 1, out, cap, &n);
                } else {
                    put_range(ipix1 + ip_lo, ipix1 + ip_hi + 1, out, cap, &n);
                }
            }
        }
    }
    if ((rlat2 >= ORC_PI) && (irmax + 1 < 4 * nside)) { /* south pole in the disc */
        int64_t sp, rp;
        int sh;
        orc_ring_info(nside, irmax + 1, &sp, &rp, &sh);
        put_range(sp, npix, out, cap, &n);
    }
    return n;
}

/* ===================================================================== */
/* Parameters (src/parameters.cpp)                                        */
/* ===================================================================== */

void orc_params_defaults(orc_params *p)
{
    memset(p, 0, sizeof(*p));
    /* Parameters::initialize (parameters.cpp:291-308) */
    p->nside = 512;
    p->mem_buffer = 10000000;
    p->test = -1;
    p->is_native = 1;
    p->is_mask = 0;
    /* constructor (parameters.cpp:311-319) */
    p->mass_to_Msun_in = -1.0;
}

static void setup_units(orc_params *p, double redshift)
{
    double Om = p->Omega_M;
    double rho = 1.8788309e-29 * pow(p->h100, 2);
    (void)rho;
    double Omega_m_at_z = Om * pow((1.0 + redshift), 3) /
                          (Om * pow((1.0 + redshift), 3) + (1.0 - Om));
    double x = Omega_m_at_z - 1.0;
    p->Delta_vir = (18.0 * ORC_PI * ORC_PI + 82.0 * x - 39.0 * x * x) / (1.0 + x);
    p->rho_crit_in_cgs = 3.0 * pow((100.0 * p->h100 / (kMpc_in_cgs * 1e-5)), 2) /
                         (8 * ORC_PI * kG_in_cgs);
    p->length_to_Mpc = p->Lbox_in_Mpc;
    p->length_to_cgs = p->Lbox_in_Mpc * kMpc_in_cgs;
    p->time_to_cgs = pow((p->rho_crit_in_cgs * kG_in_cgs), (-0.5));
    p->density_to_cgs = p->rho_crit_in_cgs;
    if (p->mass_to_Msun_in == -1) {
        p->mass_to_cgs = p->density_to_cgs * pow(p->length_to_cgs, 3);
        p->mass_to_Msun = p->mass_to_cgs / kMsun_in_cgs;
    } else {
        p->mass_to_Msun = p->mass_to_Msun_in;
        p->mass_to_cgs = p->mass_to_Msun * kMsun_in_cgs;
    }
    p->velocity_to_cgs = p->length_to_cgs / p->time_to_cgs;
    p->annihilation_flux_to_cgs =
        (p->density_to_cgs * p->mass_to_cgs / pow(p->length_to_cgs, 2));
}

static void setup_params(orc_params *p)
{
    for (int i = 0; i < 3; i++) {
        p->cpos[i] = p->centerpos[i];
        p->opos[i] = p->obs_pos[i];
    }
    for (int i = 0, j = 0; i < 10; i++) {
        p->particle_masses[i] = 0;
        if (p->particle_masses_in[i] > 0) {
            p->particle_masses[j] = p->particle_masses_in[i];
            j++;
        }
    }
    double radius = sqrt(pow(p->opos[0] - p->cpos[0], 2.0) + pow(p->opos[1] - p->cpos[1], 2.0) +
                         pow(p->opos[2] - p->cpos[2], 2.0));
    double otheta = acos((p->opos[2] - p->cpos[2]) / radius);
    double ophi = atan2((p->opos[1] - p->cpos[1]), (p->opos[0] - p->cpos[0]));
    if (ophi < 0.0) ophi += 2.0 * ORC_PI;
    p->otheta = otheta;
    p->ophi = ophi;
}

static void setup_rotation(orc_params *p)
{
    double otheta = p->otheta, ophi = p->ophi;
    double u = 0, v = -cos(otheta), w = sin(otheta) * sin(ophi);
    double rr = sqrt(u * u + v * v + w * w);
    u = u / rr;
    v = v / rr;
    w = w / rr;
    double c = -sin(otheta) * cos(ophi);
    double t = 1.0 - c;
    double s = rr;
    double rotm[3][3];
    if (rr > 0.0) {
        rotm[0][0] = u * u * t + c;
        rotm[0][1] = u * v * t + w * s;
        rotm[0][2] = u * w * t - v * s;
        rotm[1][0] = v * u * t - w * s;
        rotm[1][1] = v * v * t + c;
        rotm[1][2] = v * w * t + u * s;
        rotm[2][0] = w * u * t + v * s;
        rotm[2][1] = w * v * t - u * s;
        rotm[2][2] = w * w * t + c;
    } else {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) rotm[i][j] = (i == j) ? 1.0 : 0.0;
        if (cos(ophi) > 0.0) {
            rotm[0][0] = -1.0;
            rotm[1][1] = -1.0;
        }
    }
    double distance = sqrt(pow(p->opos[0] - p->cpos[0], 2.0) + pow(p->opos[1] - p->cpos[1], 2.0) +
                           pow(p->opos[2] - p->cpos[2], 2.0));
    double tv[3] = {p->cpos[0] + distance * p->align_vec[0] - p->opos[0],
                    p->cpos[1] + distance * p->align_vec[1] - p->opos[1],
                    p->cpos[2] + distance * p->align_vec[2] - p->opos[2]};
    double xt = tv[0] * rotm[0][0] + tv[1] * rotm[1][0] + tv[2] * rotm[2][0];
    double yt = tv[0] * rotm[0][1] + tv[1] * rotm[1][1] + tv[2] * rotm[2][1];
    double zt = tv[0] * rotm[0][2] + tv[1] * rotm[1][2] + tv[2] * rotm[2][2];
    (void)xt;
    double gamma_t = fabs(atan2(zt, yt));
    gamma_t *= -1.0;
    double r2[3][3] = {{1.0, 0.0, 0.0},
                       {0.0, cos(gamma_t), sin(gamma_t)},
                       {0.0, -sin(gamma_t), cos(gamma_t)}};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double val = 0;
            for (int k = 0; k < 3; k++) val += rotm[i][k] * r2[k][j];
            p->rotmatrix[3 * i + j] = val;
        }
}

void orc_params_setup(orc_params *p)
{
    /* parameters.cpp:476-480: setupUnits, setupOthers, setupHalo,
     * setupParams, setupRotation (redshift is always 0 there) */
    setup_units(p, 0.0);
    setup_params(p);
    setup_rotation(p);
    p->npix = 12 * p->nside * p->nside;
    p->dOmega = 4 * ORC_PI / p->npix; /* Map::setNside, datatypes.h:248-253 */
    /* flux_ann_nosom.cpp:13-16 */
    p->flux_factor = pow(pow((kc_in_cgs), 2) / (keV_in_cgs * 1.0e9), 2) /
                     (kMpc_in_cgs * 1.0e-3) * p->annihilation_flux_to_cgs;
}

static void read3(double *dst, char *val)
{
    for (int k = 0; k < 3; k++) {
        if (k) val = strtok(NULL, " ");
        if (!val) return;
        dst[k] = atof(val);
    }
}

int orc_params_from_ini(const char *path, orc_params *p)
{
    orc_params_defaults(p);
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char line[4096];
    while (fgets(line, sizeof line, f)) {
        size_t L = strlen(line);
        if (L && line[L - 1] == '\n') line[--L] = 0; /* getline drops '\n' */
        if (L == 0) continue;
        char *toks = strtok(line, " ");
        if (!toks) continue;
        if (toks[0] == '#') continue;
        char key[256];
        snprintf(key, sizeof key, "%s", toks);
        char *val = strtok(NULL, " ");
        val = strtok(NULL, " ");
        if (!val) continue;
        if (!strcmp(key, "Omega_M")) p->Omega_M = atof(val);
        else if (!strcmp(key, "h100")) p->h100 = atof(val);
        else if (!strcmp(key, "n_s")) p->n_s = atof(val);
        else if (!strcmp(key, "sigma_8")) p->sigma_8 = atof(val);
        else if (!strcmp(key, "Lbox_in_Mpc")) p->Lbox_in_Mpc = atof(val);
        else if (!strcmp(key, "mass_to_Msun")) p->mass_to_Msun_in = atof(val);
        else if (!strcmp(key, "particle_masses")) {
            int i = 0;
            do {
                p->particle_masses_in[i] = atof(val);
                i++;
                val = strtok(NULL, " ");
            } while (val != NULL && i < 10);
        } else if (!strcmp(key, "particle_numbers")) {
            int i = 0;
            do {
                p->particle_numbers_in[i] = atol(val);
                i++;
                val = strtok(NULL, " ");
            } while (val != NULL && i < 10);
        } else if (!strcmp(key, "centerpos")) read3(p->centerpos, val);
        else if (!strcmp(key, "centervel")) read3(p->centervel, val);
        else if (!strcmp(key, "align_vec")) read3(p->align_vec, val);
        else if (!strcmp(key, "obs_pos")) read3(p->obs_pos, val);
        else if (!strcmp(key, "nside")) p->nside = atoi(val);
        else if (!strcmp(key, "test")) p->test = atoi(val);
        else if (!strcmp(key, "datafile")) snprintf(p->datafile, sizeof p->datafile, "%s", val);
        else if (!strcmp(key, "partfile") || !strcmp(key, "densfile") ||
                 !strcmp(key, "hsmfile") || !strcmp(key, "vdspfile"))
            p->is_native = 0;
        else if (!strcmp(key, "outputmap")) snprintf(p->outputmap, sizeof p->outputmap, "%s", val);
        else if (!strcmp(key, "mem_buffer")) p->mem_buffer = atoi(val);
        else if (!strcmp(key, "mask")) {
            snprintf(p->maskfile, sizeof p->maskfile, "%s", val);
            p->is_mask = 1;
        }
    }
    fclose(f);
    orc_params_setup(p);
    return 0;
}

void orc_config_from_params(const orc_params *p, orc_config *c)
{
    memset(c, 0, sizeof(*c));
    for (int i = 0; i < 3; i++) c->opos[i] = p->opos[i];
    for (int i = 0; i < 9; i++) c->rotmatrix[i] = p->rotmatrix[i];
    c->flux_factor = p->flux_factor;
    c->nside = p->nside;
    /* genmap.cpp:52-58: float hires = FLT_MAX; min over positive masses */
    float hires = FLT_MAX;
    for (int i = 0; i < 10; i++)
        if (p->particle_masses[i] > 0 && p->particle_masses[i] < hires)
            hires = (float)p->particle_masses[i];
    c->mass_cut = (double)hires * 1.1;
    c->flux_kind = ORC_FLUX_ANN_NOSOM;
    /* flux_decay.cpp:11-14 */
    double decay_flux_to_cgs = (p->mass_to_cgs / pow(p->length_to_cgs, 2));
    c->decay_factor = pow(kc_in_cgs, 2) / (keV_in_cgs * 1.0e9) / (kMpc_in_cgs * 1.0e-3) * decay_flux_to_cgs;
    c->vel_to_cgs = p->velocity_to_cgs;
    c->c_in_cgs = kc_in_cgs;
    for (int i = 0; i < 3; i++) c->cpos[i] = p->cpos[i];
}

/* ===================================================================== */
/* flux / angles                                                          */
/* ===================================================================== */

double orc_getflux(double flux_factor, const orc_particle *q, double d)
{
    /* flux_ann_nosom.cpp:18 */
    return flux_factor * (double)q->dens * (double)q->mass / (4.0 * ORC_PI * d * d);
}

/* halocore.cpp:70-84: float differences, double distance */
static double orc_core_func(float x, float y, float z, float hxc, float hyc, float hzc, float radius)
{
    double xx = (x - hxc);
    double yy = (y - hyc);
    double zz = (z - hzc);
    double r = sqrt(xx * xx + yy * yy + zz * zz);
    double ratio = 1.0;
    if (r < radius) {
        if (r > 0) ratio = 0.55 * pow(r / radius, -0.9);
    }
    return ratio;
}

double orc_core_correction(const float *halos, int64_t n, float x, float y, float z)
{
    double xp = x, yp = y, zp = z;
    double r = 1.0;
    for (int64_t i = 0; i < n; i++) {
        const float *h = halos + 4 * i;
        r *= orc_core_func((float)xp, (float)yp, (float)zp, h[0], h[1], h[2], (float)(h[3] / (0.73 * 40000.0)));
    }
    return r;
}

double orc_getflux_cfg(const orc_config *c, const orc_particle *q, double d)
{
    const double K = c->flux_factor;
    switch (c->flux_kind) {
    case ORC_FLUX_ANN_PWAVE: {
        /* flux_ann_pwave.cpp:18-27: (v/c)^2 */
        double v = q->sigmav;
        double somer_v = (v * c->vel_to_cgs) / c->c_in_cgs;
        somer_v = somer_v * somer_v;
        return K * somer_v * q->dens * q->mass / (4.0 * ORC_PI * d * d);
    }
    case ORC_FLUX_ANN_SOMV: {
        /* flux_ann_somv.cpp:21-40: c / v / 1e4, v in km/s saturated at 1 */
        double v = q->sigmav * c->vel_to_cgs;
        v /= 1.0e5;
        if (v < 1) v = 1;
        double somer_v = c->c_in_cgs / v / 1.0e4;
        return K * somer_v * q->dens * q->mass / (4.0 * ORC_PI * d * d);
    }
    case ORC_FLUX_ANN_SOMV2: {
        /* flux_ann_somv2.cpp:21-41: (c / v)^2 / 1e8, v saturated at 5 km/s */
        double v = q->sigmav * c->vel_to_cgs;
        v /= 1.0e5;
        if (v < 5) v = 5;
        double somer_v = c->c_in_cgs / v;
        somer_v *= (somer_v / 1.0e8);
        return K * somer_v * q->dens * q->mass / (4.0 * ORC_PI * d * d);
    }
    case ORC_FLUX_DECAY:
        /* flux_decay.cpp:10-16 */
        return c->decay_factor * q->mass / (4.0 * ORC_PI * d * d);
    case ORC_FLUX_ANN_CORE_NOSOM: {
        /* flux_ann_core_nosom.cpp:10-18: correction squared times nosom */
        double r = orc_core_correction(c->halos, c->nhalos, q->posx, q->posy, q->posz);
        return r * r * K * q->dens * q->mass / (4.0 * ORC_PI * d * d);
    }
    case ORC_FLUX_ANN_ADIABATIC: {
        /* flux_ann_adiabatic.cpp:11-30: profile factor of the distance to
         * the halo centre in kpc/h */
        double dx = q->posx - c->cpos[0];
        double dy = q->posy - c->cpos[1];
        double dz = q->posz - c->cpos[2];
        double r = sqrt(dx * dx + dy * dy + dz * dz) * 40000.0;
        double factor = (14.6534 * pow(1 + 0.0355872 * r, 1.76)) /
                        (pow(1 + 0.10888 * pow(r, 0.76), 3.34211) * pow(r, 0.13));
        double f = K * q->dens * q->mass / (4.0 * ORC_PI * d * d);
        f *= (factor * factor);
        return f;
    }
    default:
        return orc_getflux(K, q, d);
    }
}

void orc_calc_angles(double x, double y, double z, double d, const double rot[9],
                     double *costheta, double *phi)
{
    /* anglefuncs.cpp:15-52: temp_i = sum_j R[j][i] v_j */
    double v[3] = {x, y, z}, t[3];
    for (int i = 0; i < 3; i++) {
        t[i] = 0;
        for (int j = 0; j < 3; j++) t[i] += rot[3 * j + i] * v[j];
    }
    *costheta = t[2] / d;
    double ph = atan2(t[1], t[0]);
    if (ph < 0) ph += 2.0 * ORC_PI;
    ph -= ORC_PI;
    if (ph < 0) ph = 2.0 * ORC_PI + ph;
    if (ph > 2 * ORC_PI) ph = ph - 2.0 * ORC_PI;
    *phi = ph;
}

/* ===================================================================== */
/* particle file (datareader.cpp:101-272, native path)                    */
/* ===================================================================== */

int64_t orc_read_native(const char *path, int test, const char *maskfile, orc_particle **out)
{
    *out = NULL;
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    int32_t n32 = 0;
    if (fread(&n32, 4, 1, f) != 1) { fclose(f); return -1; }
    int64_t n = (uint32_t)n32;
    FILE *mf = NULL;
    if (maskfile && maskfile[0]) {
        mf = fopen(maskfile, "rb");
        int32_t np = 0;
        if (!mf || fread(&np, 4, 1, mf) != 1 || np != (int32_t)n) {
            if (mf) fclose(mf);
            fclose(f);
            return -1;
        }
    }
    if (test != -1 && n > test) n = test;
    orc_particle *buf = (orc_particle *)malloc((size_t)(n ? n : 1) * sizeof(orc_particle));
    if (!buf) { fclose(f); if (mf) fclose(mf); return -1; }
    if ((int64_t)fread(buf, sizeof(orc_particle), (size_t)n, f) != n) {
        free(buf); fclose(f); if (mf) fclose(mf);
        return -1;
    }
    if (mf) {
        unsigned char *m = (unsigned char *)malloc((size_t)(n ? n : 1));
        if ((int64_t)fread(m, 1, (size_t)n, mf) != n) {
            free(m); free(buf); fclose(f); fclose(mf);
            return -1;
        }
        for (int64_t i = 0; i < n; i++)
            if (m[i] == 0) buf[i].dens = -1;
        free(m);
        fclose(mf);
    }
    fclose(f);
    *out = buf;
    return n;
}

void orc_free(void *p) { free(p); }

/* ===================================================================== */
/* the generator (genmap.cpp:22-175)                                      */
/* ===================================================================== */

typedef struct {
    int64_t *pix;
    double *w;
    int64_t cap;
} scratch_t;

static int scratch_reserve(scratch_t *s, int64_t n)
{
    if (n <= s->cap) return 0;
    int64_t c = s->cap ? s->cap : 1024;
    while (c < n) c *= 2;
    int64_t *np = (int64_t *)realloc(s->pix, (size_t)c * sizeof(int64_t));
    double *nw = (double *)realloc(s->w, (size_t)c * sizeof(double));
    if (!np || !nw) return -1;
    s->pix = np;
    s->w = nw;
    s->cap = c;
    return 0;
}

/* One particle.  Returns the number of footprint pixels written to sc
 * (>= 2 when spread), 1 for a single-pixel deposit, 0 when filtered.
 * *pix0, *flux filled for deposits. */
static int64_t one_particle(const orc_particle *q, const orc_config *c, double theta0,
                            scratch_t *sc, int64_t *pix0, double *flux, int *kind)
{
    *kind = 0;
    if ((double)q->mass >= c->mass_cut || q->dens < 0) return 0;
    double ox = c->opos[0], oy = c->opos[1], oz = c->opos[2];
    double d = sqrt(((double)q->posx - ox) * ((double)q->posx - ox) +
                    ((double)q->posy - oy) * ((double)q->posy - oy) +
                    ((double)q->posz - oz) * ((double)q->posz - oz));
    double f = orc_getflux_cfg(c, q, d);
    double costheta, phi;
    orc_calc_angles((double)q->posx - ox, (double)q->posy - oy, (double)q->posz - oz, d,
                    c->rotmatrix, &costheta, &phi);
    double theta = acos(costheta);
    double r = (double)q->hsmooth / d;
    /* ang2vec (anglefuncs.cpp:55-71) */
    double vec[3];
    double sz = sin(theta);
    vec[0] = sz * cos(phi);
    vec[1] = sz * sin(phi);
    vec[2] = cos(theta);
    int64_t pix = orc_ang2pix_ring(c->nside, theta, phi);
    *pix0 = pix;
    *flux = f;
    if (2.0 * r < theta0) { *kind = 1; return 1; }
    int64_t cnt = orc_query_disc(c->nside, theta, phi, 2.0 * r + c->footprint_pad, sc->pix, sc->cap);
    if (cnt > sc->cap) {
        if (scratch_reserve(sc, cnt)) return -1;
        cnt = orc_query_disc(c->nside, theta, phi, 2.0 * r + c->footprint_pad, sc->pix, sc->cap);
    }
    if (cnt < 2) { *kind = 2; return 1; }
    double norm = 0.0;
    for (int64_t j = 0; j < cnt; j++) {
        double u[3];
        orc_pix2vec_ring(c->nside, sc->pix[j], u);
        double d2 = acos(u[0] * vec[0] + u[1] * vec[1] + u[2] * vec[2]) / r;
        d2 = d2 * d2;
        sc->w[j] = exp(-0.5 * d2 / 0.333);
        norm += sc->w[j];
    }
    for (int64_t j = 0; j < cnt; j++) sc->w[j] = sc->w[j] / norm;
    *kind = 3;
    return cnt;
}

static double theta0_of(int nside)
{
    double dOmega = 4 * ORC_PI / (12.0 * nside * nside);
    return acos(1.0 - dOmega / (2.0 * 3.141592)); /* genmap.cpp:38 */
}

static void run_range(const orc_particle *parts, int64_t lo, int64_t hi, const orc_config *c,
                      double *map, orc_stats *st)
{
    double theta0 = theta0_of(c->nside);
    scratch_t sc = {0, 0, 0};
    scratch_reserve(&sc, 1 << 16);
    for (int64_t i = lo; i < hi; i++) {
        int64_t pix0;
        double f;
        int kind;
        int64_t cnt = one_particle(&parts[i], c, theta0, &sc, &pix0, &f, &kind);
        if (kind == 0) continue;
        st->kept++;
        st->total_flux += f;
        if (kind == 1 || kind == 2) {
            map[pix0] += f;
            if (kind == 1) st->subpixel++;
            else st->fallback++;
            continue;
        }
        st->extended++;
        st->contributions += cnt;
        for (int64_t j = 0; j < cnt; j++) map[sc.pix[j]] += sc.w[j] * f;
    }
    free(sc.pix);
    free(sc.w);
}

void orc_genmap(const orc_particle *parts, int64_t n, const orc_config *c, double *map,
                orc_stats *st)
{
    orc_stats local;
    memset(&local, 0, sizeof local);
    run_range(parts, 0, n, c, map, &local);
    if (st) *st = local;
}

void orc_genmap_omp(const orc_particle *parts, int64_t n, const orc_config *c, double *map,
                    orc_stats *st, int nthreads)
{
#ifdef _OPENMP
    int64_t npix = orc_npix(c->nside);
    if (nthreads < 1) nthreads = omp_get_max_threads();
    orc_stats tot;
    memset(&tot, 0, sizeof tot);
    double **maps = (double **)calloc((size_t)nthreads, sizeof(double *));
    orc_stats *sts = (orc_stats *)calloc((size_t)nthreads, sizeof(orc_stats));
#pragma omp parallel num_threads(nthreads)
    {
        int t = omp_get_thread_num();
        maps[t] = (double *)calloc((size_t)npix, sizeof(double));
        // small tasks keep the threads balanced on heavy-tailed footprints
        int64_t chunk = n / ((int64_t)nthreads * 64);
        if (chunk < 16) chunk = 16;
        if (chunk > 4096) chunk = 4096;
#pragma omp for schedule(dynamic, 1)
        for (int64_t b = 0; b < (n + chunk - 1) / chunk; b++) {
            int64_t lo = b * chunk, hi = lo + chunk < n ? lo + chunk : n;
            run_range(parts, lo, hi, c, maps[t], &sts[t]);
        }
    }
    for (int t = 0; t < nthreads; t++) {
        for (int64_t p = 0; p < npix; p++) map[p] += maps[t][p];
        free(maps[t]);
        tot.kept += sts[t].kept;
        tot.subpixel += sts[t].subpixel;
        tot.fallback += sts[t].fallback;
        tot.extended += sts[t].extended;
        tot.contributions += sts[t].contributions;
        tot.total_flux += sts[t].total_flux;
    }
    free(maps);
    free(sts);
    if (st) *st = tot;
#else
    (void)nthreads;
    orc_genmap(parts, n, c, map, st);
#endif
}

double orc_finalize(double *map, int64_t npix, double dOmega)
{
    double total = 0;
    for (int64_t i = 0; i < npix; i++) {
        total += map[i];
        map[i] /= dOmega;
    }
    return total;
}

int64_t orc_particle_footprint(const orc_particle *q, const orc_config *c, int64_t *pix,
                               double *wnorm, int64_t cap, double *flux_out)
{
    double theta0 = theta0_of(c->nside);
    scratch_t sc = {0, 0, 0};
    scratch_reserve(&sc, cap > 1024 ? cap : 1024);
    int64_t pix0;
    double f = 0;
    int kind;
    int64_t cnt = one_particle(q, c, theta0, &sc, &pix0, &f, &kind);
    *flux_out = f;
    int64_t out = 0;
    if (kind == 1 || kind == 2) {
        if (cap > 0) { pix[0] = pix0; wnorm[0] = 1.0; }
        out = 1;
    } else if (kind == 3) {
        for (int64_t j = 0; j < cnt && j < cap; j++) {
            pix[j] = sc.pix[j];
            wnorm[j] = sc.w[j];
        }
        out = cnt;
    }
    free(sc.pix);
    free(sc.w);
    return out;
}
