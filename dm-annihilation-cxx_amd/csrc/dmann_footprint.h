// Per-particle and per-ring math of the projector, shared by the kernels and
// by the host-side footprint enumerator (dmann_host_footprint), so the CPU
// tests exercise exactly the code the GPU runs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dmann_geom.h"
#include "dmann_kernels.h"

namespace dmann {

// The fast weight form (MODE 0 of weight32) holds when the footprint radius
// rad = 2r is at most 0.5 and its longitude half-width
// asin(sin(rad)/sin(theta)) stays below 0.445; this also keeps the disc off
// both poles.  s0 = sin(theta) as sqrt((1 - cos theta)(1 + cos theta)).
DMANN_HD bool fast_geometry(double rad, double s0) { return rad <= 0.5 && sin(rad) < s0 * 0.43; }

// coreFunc of src/halocore/halocore.cpp:70-84: float differences, FP64
// distance; 0.55 (r / radius)^-0.9 inside the core, 1 outside.
DMANN_HD double core_factor(float x, float y, float z, float hxc, float hyc, float hzc, float radius)
{
    const double xx = (x - hxc);
    const double yy = (y - hyc);
    const double zz = (z - hzc);
    const double r = sqrt(xx * xx + yy * yy + zz * zz);
    double ratio = 1.0;
    if (r < radius) {
        if (r > 0) ratio = 0.55 * pow(r / radius, -0.9);
    }
    return ratio;
}

// HaloCore::getCorrection (src/halocore/halocore.cpp:47-68): the product of
// the factors of all cores in row order; cores that do not contain the
// particle contribute exactly 1, so only the cores listed in the particle's
// grid cell are visited (device memory: device code only).
DMANN_HD double core_correction(const DevParams &P, float x, float y, float z)
{
    double c = 1.0;
    if (P.nhalos <= 0) return c;
    const double f[3] = {((double)x - P.grid_lo[0]) * P.grid_inv, ((double)y - P.grid_lo[1]) * P.grid_inv,
                         ((double)z - P.grid_lo[2]) * P.grid_inv};
    int ic[3];
    for (int a = 0; a < 3; a++) {
        if (!(f[a] >= 0.0)) return c;
        ic[a] = (int)f[a];
        if (ic[a] >= P.grid_dim[a]) return c;
    }
    const int cell = (ic[2] * P.grid_dim[1] + ic[1]) * P.grid_dim[0] + ic[0];
    for (int k = P.cell_start[cell]; k < P.cell_start[cell + 1]; ++k) {
        const float4 h = P.halos[P.cell_list[k]];
        c *= core_factor(x, y, z, h.x, h.y, h.z, h.w);
    }
    return c;
}

// getflux of the map's flux plug-in for particle q at distance d, with the
// reference's operation order (src/flux/flux_ann_nosom.cpp:12-23,
// flux_ann_pwave.cpp:13-29, flux_ann_somv.cpp:16-43, flux_ann_somv2.cpp:16-44,
// flux_decay.cpp:10-16, flux_ann_adiabatic.cpp:10-31).
DMANN_HD double plugin_flux(const DevParams &P, const Particle &q, double d)
{
    const double K = P.flux_factor;
    switch (P.flux_kind) {
    case DMANN_FLUX_ANN_PWAVE: {
        const double v = q.sigmav;
        double somer = (v * P.vel_to_cgs) / P.c_in_cgs;
        somer = somer * somer;
        return K * somer * (double)q.dens * (double)q.mass / (4.0 * kPi * d * d);
    }
    case DMANN_FLUX_ANN_SOMV: {
        double v = (double)q.sigmav * P.vel_to_cgs;
        v /= 1.0e5; // km/s, saturated at 1 km/s
        if (v < 1) v = 1;
        const double somer = P.c_in_cgs / v / 1.0e4;
        return K * somer * (double)q.dens * (double)q.mass / (4.0 * kPi * d * d);
    }
    case DMANN_FLUX_ANN_SOMV2: {
        double v = (double)q.sigmav * P.vel_to_cgs;
        v /= 1.0e5; // km/s, saturated at 5 km/s
        if (v < 5) v = 5;
        double somer = P.c_in_cgs / v;
        somer *= (somer / 1.0e8);
        return K * somer * (double)q.dens * (double)q.mass / (4.0 * kPi * d * d);
    }
    case DMANN_FLUX_DECAY:
        return P.decay_factor * (double)q.mass / (4.0 * kPi * d * d);
    case DMANN_FLUX_ANN_CORE_NOSOM: {
        const double r = core_correction(P, q.posx, q.posy, q.posz);
        return r * r * K * (double)q.dens * (double)q.mass / (4.0 * kPi * d * d);
    }
    case DMANN_FLUX_ANN_ADIABATIC: {
        // profile factor of the distance to the halo centre, in kpc/h
        const double dx = (double)q.posx - P.cpos[0];
        const double dy = (double)q.posy - P.cpos[1];
        const double dz = (double)q.posz - P.cpos[2];
        const double r = sqrt(dx * dx + dy * dy + dz * dz) * 40000.0;
        const double factor =
            (14.6534 * pow(1 + 0.0355872 * r, 1.76)) / (pow(1 + 0.10888 * pow(r, 0.76), 3.34211) * pow(r, 0.13));
        double f = K * (double)q.dens * (double)q.mass / (4.0 * kPi * d * d);
        f *= (factor * factor);
        return f;
    }
    default:
        return K * (double)q.dens * (double)q.mass / (4.0 * kPi * d * d);
    }
}

// The filter, the observer-frame offset and distance, and the flux of one
// particle (src/genmap.cpp:82-98); false when the particle is skipped.
DMANN_HD bool kept_flux(const DevParams &P, const Particle &q, double &dx, double &dy, double &dz, double &d,
                        double &flux)
{
    if (((double)q.mass >= P.mass_cut) || (q.dens < 0)) return false;
    dx = (double)q.posx - P.opos[0];
    dy = (double)q.posy - P.opos[1];
    dz = (double)q.posz - P.opos[2];
    d = sqrt(dx * dx + dy * dy + dz * dz);
    flux = plugin_flux(P, q, d);
    return true;
}

// Pre-pass of one particle (src/genmap.cpp:82-110; src/anglefuncs.cpp:15-71;
// the flux plug-in, plugin_flux).  Returns kClsSkip (filtered),
// kClsPoint (whole flux to rec.p0), kClsMedium or kClsLarge (footprint).
// *bad is set when the direction is undefined (the reference exits there).
DMANN_HD int prepass_one(const DevParams &P, const Particle &q, DiscRec &rec, bool &bad)
{
    bad = false;
    double dx, dy, dz, d, flux;
    if (!kept_flux(P, q, dx, dy, dz, d, flux)) return kClsSkip;
    const double v[3] = {dx, dy, dz};
    double t[3];
    for (int a = 0; a < 3; a++) {
        t[a] = 0;
        for (int b = 0; b < 3; b++) t[a] += P.rot[3 * b + a] * v[b];
    }
    const double costheta = t[2] / d;
    double ph = atan2(t[1], t[0]);
    if (ph < 0) ph += 2.0 * kPi;
    ph -= kPi;
    if (ph < 0) ph = 2.0 * kPi + ph;
    if (ph > 2 * kPi) ph = ph - 2.0 * kPi;
    const double theta = acos(costheta);
    const double r = (double)q.hsmooth / d;
    if (!(theta >= 0.0 && theta <= kPi)) {
        bad = true;
        return kClsSkip;
    }
    rec.theta = theta;
    rec.phi = ph;
    rec.r = r;
    rec.flux = flux;
    rec.p0 = ang2pix(P.nside, theta, ph);
    rec.pad = 0;
    if (2.0 * r < P.theta0) {
        rec.est = 1;
        rec.cls = kClsPoint;
        return kClsPoint;
    }
    const double rad = 2.0 * r + P.footprint_pad;
    const double est = (rad >= kPi) ? (double)P.npix : kTwoPi * (1.0 - cos(rad)) / P.dOmega;
    rec.est = est > 2.0e9 ? 2000000000 : (int)est;
    const double z0 = cos(theta);
    const double s0 = sqrt((1 - z0) * (1 + z0));
    // medium: one wavefront, fast weight form; large: everything else
    // (sky-covering, polar, wide) on the workgroup kernel
    rec.cls = (fast_geometry(rad, s0) && est <= (double)P.large_pix) ? kClsMedium : kClsLarge;
    return rec.cls;
}

// Per-particle quantities of the query_disc ring walk (FP64).
struct PGeo {
    double theta, phi, rad, r, z0, s0, xa, cosr, flux;
    double dphi; // widest longitude half-width of the disc, asin(sin(rad)/s0) (pi/2 when it reaches a pole)
    int irmin, irmax, nnorth, npart, nsouth, nrings, p0;
    bool all;
    int mode;  // 0: fast polynomials, 1: general (sin / asin), 2: sin series + fast asin^2 series
    int pv;    // series lengths of the fast form (fast_variant)
    float K4;  // -(0.5/0.333) * 4 * log2(e) / r^2
};

// Series lengths of the fast weight form for one particle, chosen so that
// each truncated series stays within 2e-8 relative over its footprint
// (|x| <= xmax, h <= sin^2(rad/2)):
//   0: sin^2 2 terms, asin^2 2 terms  (xmax <= 0.0518, rad <= 0.0366)
//   1: 3 and 3 terms                  (xmax <= 0.2717, rad <= 0.1498)
//   2: 4 and 6 terms                  (the whole fast geometry)
DMANN_HD int fast_variant(double rad, double xmax)
{
    if (rad <= 0.0366 && xmax <= 0.0518) return 0;
    if (rad <= 0.1498 && xmax <= 0.2717) return 1;
    return 2;
}

// pad: radians added to the footprint radius 2r (DevParams::footprint_pad).
DMANN_HD PGeo particle_geo(const DiscRec &rec, int nside, double pad)
{
    PGeo g;
    g.theta = rec.theta;
    g.phi = fmodulo(rec.phi, kTwoPi); // pointing::normalize
    g.r = rec.r;
    g.rad = 2.0 * rec.r + pad;
    g.flux = rec.flux;
    g.p0 = rec.p0;
    g.all = g.rad >= kPi;
    g.cosr = cos(g.rad);
    g.z0 = cos(g.theta);
    g.s0 = sqrt((1 - g.z0) * (1 + g.z0));
    g.xa = 1. / g.s0;
    const double rlat1 = g.theta - g.rad;
    const double zmax = cos(rlat1);
    g.irmin = ring_above(nside, zmax) + 1;
    const bool north = (rlat1 <= 0) && (g.irmin > 1);
    const double rlat2 = g.theta + g.rad;
    const double zmin = cos(rlat2);
    g.irmax = ring_above(nside, zmin);
    const bool south = (rlat2 >= kPi) && (g.irmax + 1 < 4 * nside);
    g.nnorth = north ? g.irmin - 1 : 0;
    g.npart = g.irmax >= g.irmin ? g.irmax - g.irmin + 1 : 0;
    g.nsouth = south ? 4 * nside - 1 - g.irmax : 0;
    g.nrings = g.all ? 4 * nside - 1 : g.nnorth + g.npart + g.nsouth;
    const double kc = 0.5 / 0.333; // src/genmap.cpp:143
    g.K4 = (float)(-kc * 4.0 * 1.4426950408889634 / (g.r * g.r));
    // the asin^2 series of the fast form holds for psi <= 0.5 whatever the
    // longitude offsets: footprints of radius <= 0.5 that reach far in
    // longitude (near a pole) take it with sin(x/2) from its own series
    g.mode = fast_geometry(g.rad, g.s0) ? 0 : (g.rad <= 0.5 ? 2 : 1);
    g.dphi = asin(fmin(1.0, sin(g.rad) / g.s0));
    g.pv = fast_variant(g.rad, g.dphi);
    return g;
}

// One ring of a footprint.
struct RingRec {
    int cnt;
    float A, B, offm, step; // haversine terms (FP32)
    float Ac;               // cos^2((theta_ring - theta)/2)
    float E, D;             // sin^2, cos^2 of (theta_ring + theta)/2
    int kc_rel;             // column nearest phi, relative to col0
    int col0, startpix, nr;
    double z, s, shift;     // FP64 terms (FP64 weight form)
};

// Ring `gi` (0-based position in the particle's ring list: north polar cap
// rings, then the partial rings irmin..irmax, then south polar cap rings).
DMANN_HD RingRec ring_setup(const PGeo &g, int nside, int gi)
{
    RingRec o;
    int iz;
    bool full;
    if (g.all) {
        iz = gi + 1;
        full = true;
    } else if (gi < g.nnorth) {
        iz = gi + 1;
        full = true;
    } else if (gi < g.nnorth + g.npart) {
        iz = g.irmin + (gi - g.nnorth);
        full = false;
    } else {
        iz = g.irmax + 1 + (gi - g.nnorth - g.npart);
        full = true;
    }
    int startpix, nr, shifted;
    ring_info(nside, iz, startpix, nr, shifted);
    const double shift = shifted ? 0.5 : 0.0;
    const double z = ring2z(nside, iz);
    const double fnr = (double)nr;
    const int kc = (int)floor(fnr * kInvTwoPi * g.phi - shift + 0.5);
    int col0 = 0, cnt = 0;
    if (full) {
        col0 = kc - nr / 2;
        cnt = nr;
    } else {
        // Healpix_Base::query_disc (non-inclusive), called at src/genmap.cpp:119
        const double x = (g.cosr - z * g.z0) * g.xa;
        const double ysq = 1 - z * z - x * x;
        const double dphi = (ysq <= 0) ? 0 : atan2(sqrt(ysq), x);
        if (dphi > 0) {
            const int ip_lo = (int)floor(fnr * kInvTwoPi * (g.phi - dphi) - shift) + 1;
            const int ip_hi = (int)floor(fnr * kInvTwoPi * (g.phi + dphi) - shift);
            if (ip_lo <= ip_hi) {
                col0 = ip_lo;
                cnt = ip_hi - ip_lo + 1;
            }
        }
    }
    const double s = sqrt((1 - z) * (1 + z));
    const double cdt = z * g.z0 + s * g.s0;         // cos(theta_ring - theta)
    double A = (1.0 - cdt) * 0.5;                   // hav(theta_ring - theta)
    double Ac = (1.0 + cdt) * 0.5;                  // 1 - hav(theta_ring - theta)
    if (A < 0) A = 0;
    if (Ac < 0) Ac = 0;
    const double cst = z * g.z0 - s * g.s0;         // cos(theta_ring + theta)
    double E = (1.0 - cst) * 0.5, D = (1.0 + cst) * 0.5;
    if (E < 0) E = 0;
    if (D < 0) D = 0;
    const double step = kTwoPi / fnr;
    o.cnt = cnt;
    o.A = (float)A;
    o.Ac = (float)Ac;
    o.E = (float)E;
    o.D = (float)D;
    o.B = (float)(s * g.s0);
    o.offm = (float)(((double)kc + shift) * step - g.phi);
    o.step = (float)step;
    o.kc_rel = kc - col0;
    o.col0 = col0;
    o.startpix = startpix;
    o.nr = nr;
    o.z = z;
    o.s = s;
    o.shift = shift;
    return o;
}

// Position of ring r in a particle's ring list (north polar cap, partial
// rings irmin..irmax, south polar cap), as ring_setup counts them; r must
// lie in the particle's ring range (sky_ring_range).
DMANN_HD int ring_index(const PGeo &g, int r)
{
    if (g.all || r < g.irmin) return r - 1;
    if (r <= g.irmax) return g.nnorth + (r - g.irmin);
    return g.nnorth + g.npart + (r - g.irmax - 1);
}

// First and last ring of a footprint: its ring list is contiguous.
DMANN_HD void sky_ring_range(const PGeo &g, int nside, int &rlo, int &rhi)
{
    rlo = (g.all || g.nnorth > 0) ? 1 : g.irmin;
    rhi = (g.all || g.nsouth > 0) ? 4 * nside - 1 : g.irmax;
}

// Ring columns per LDS segment of the sky-covering gather.
constexpr int kSkySeg = 2048;

// The columns of segment [s0, s0 + ws) of a ring of nr pixels that a
// footprint with columns c0, c0 + 1, ..., c0 + cnt - 1 (mod nr; c0 in
// [0, nr), cnt <= nr) covers, as positions t in the footprint's column
// range: [u, u + na) and, past the wrap, [0, nb).  The segment's columns
// are positions [u, u + ws) mod nr.
struct SegSpan {
    int u, na, nb;
};
DMANN_HD SegSpan seg_span(int c0, int cnt, int nr, int s0, int ws)
{
    SegSpan o;
    o.u = s0 - c0;
    if (o.u < 0) o.u += nr;
    const int a1 = cnt < o.u + ws ? cnt : o.u + ws;
    o.na = a1 > o.u ? a1 - o.u : 0;
    const int b1 = cnt < o.u + ws - nr ? cnt : o.u + ws - nr;
    o.nb = b1 > 0 ? b1 : 0;
    return o;
}

// Slot in segment s0 of footprint position t (t from seg_span).
DMANN_HD int seg_slot(int c0, int t, int nr, int s0)
{
    int col = c0 + t;
    if (col >= nr) col -= nr;
    int slot = col - s0;
    if (slot < 0) slot += nr;
    return slot;
}

// Partial ring `iz` of a disc that reaches no pole (the medium class), with
// the same column range as ring_setup but cheaper: the half-width dphi is
// first computed in FP32 from the FP64 terms x and 1 - z^2 - x^2 (a
// half-angle asin polynomial, within 3e-7 of their FP64 atan2 over 2e7
// random rings, host build); the floor() decisions for ip_lo / ip_hi are
// kept only when the boundary
// lies farther than a bound on the FP32 error (relative 1e-6 of dphi) from
// an integer, otherwise dphi is recomputed in FP64 exactly as ring_setup
// does.  Only the FP32 weight terms are filled (not z, s, shift); the
// column step is an FP32 quotient and the offset of column kc is its
// fraction of a column (|.| <= 1/2, from the same FP64 product as kc) times
// the step, so no FP64 division is made.
//
// e0, e1, e2 are the weight terms the ring walks use.  For the 2+2 series
// form (pv 0) they are the exponent as a quadratic in u = x^2,
//   log2 w = K4 G(A + B sin^2(x/2)) = e0 + e1 u + e2 u^2 + O(u^3),
//   G(h) = asin^2(sqrt h) = h + h^2/3 + 8h^3/45 + 4h^4/35 + ...,
//   e0 = K4 G(A), e1 = K4 G'(A) B/4, e2 = K4 (G''(A) B^2/32 - G'(A) B/48).  The u^3 remainder is at most (1/360) u^2 of the e1 u
// term (B in [0, 1], A <= 3.4e-4), below 2e-8 for the pv 0 range
// u <= 0.0518^2 -- the truncation error of the series form it replaces
// (ring_quad).
// For the other forms they are (A, B, 0).
struct RingLite {
    int cnt, col0, kc_rel, startpix, nr;
    float A, B, offm, step;
    float e0, e1, e2;
};

// Exponent polynomial of a pv 0 ring (RingLite), in FP32 from the FP32
// terms A, B: A <= 3.4e-4, so the series of G, G', G'' are exact to FP32
// rounding with three terms, and the rounding of A and B (as in the series
// form) dominates the error.
DMANN_HD void ring_quad(float A, float B, float K4, float &e0, float &e1, float &e2)
{
    const float G0 = A * fmaf(A, fmaf(A, 8.0f / 45.0f, 1.0f / 3.0f), 1.0f);
    const float G1 = fmaf(A, fmaf(A, 8.0f / 15.0f, 2.0f / 3.0f), 1.0f);
    const float G2 = fmaf(A, 16.0f / 15.0f, 2.0f / 3.0f);
    e0 = K4 * G0;
    e1 = (K4 * G1) * (B * 0.25f);
    e2 = K4 * (B * fmaf(G2 * (1.0f / 32.0f), B, G1 * (-1.0f / 48.0f)));
}

// Square root, the hardware instruction on the device (1 ulp).
DMANN_HD float fast_sqrt(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}

// Reciprocal, the hardware instruction on the device (1 ulp).
DMANN_HD float fast_rcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// asin(sqrt m) for m in [0, 1/2] (FP32): sqrt m times a degree-8
// polynomial fitted to asin(sqrt m) / sqrt m on [0, 1/2] (Chebyshev least
// squares, error 6e-9 relative; with the roundings about 2e-7).
DMANN_HD float asin_sqrt(float m)
{
    float g = fmaf(m, 1.2517987633e-01f, -1.3961639202e-01f);
    g = fmaf(m, g, 1.1148400957e-01f);
    g = fmaf(m, g, -9.2857844148e-03f);
    g = fmaf(m, g, 3.6501743054e-02f);
    g = fmaf(m, g, 4.3977226295e-02f);
    g = fmaf(m, g, 7.5037316312e-02f);
    g = fmaf(m, g, 1.6666578722e-01f);
    g = fmaf(m, g, 1.0f);
    return fast_sqrt(m) * g;
}

DMANN_HD RingLite ring_lite(const PGeo &g, const RingGeo &G, int iz)
{
    RingLite o;
    int shifted;
    ring_info(G.nside, iz, o.startpix, o.nr, shifted);
    const double shift = shifted ? 0.5 : 0.0;
    const double z = ring2z(G, iz);
    const double fnr = (double)o.nr;
    const double a = fnr * kInvTwoPi * g.phi;
    const int kc = (int)floor(a - shift + 0.5);
    o.cnt = 0;
    o.col0 = kc;
    const double x = (g.cosr - z * g.z0) * g.xa;
    const double ysq = 1 - z * z - x * x;
    if (ysq > 0) {
        // dphi = atan2(sqrt ysq, x) from the same ysq and x as the FP64
        // construction: with R^2 = x^2 + ysq, the smaller of sin^2(dphi/2)
        // and cos^2(dphi/2) is m = ysq / (2 R (R + |x|)) <= 1/2 (no
        // cancellation), and dphi = 2 asin(sqrt m), or pi minus it for x < 0
        // (relative error below 5e-7, inside the bound the floor decisions
        // are checked against)
        const float xf = (float)x, yf = (float)ysq;
        const float R = fast_sqrt(fmaf(xf, xf, yf));
        const float m = yf * fast_rcp(2.0f * R * (R + fabsf(xf)));
        const float a2 = 2.0f * asin_sqrt(m);
        const float df = xf >= 0.0f ? a2 : 3.14159274f - a2;
        double dphi = (double)df;
        double ulo = fnr * kInvTwoPi * (g.phi - dphi) - shift;
        double uhi = fnr * kInvTwoPi * (g.phi + dphi) - shift;
        const double margin = fnr * kInvTwoPi * dphi * 1e-6 + 1e-9;
        const double flo = ulo - floor(ulo), fhi = uhi - floor(uhi);
        if (flo < margin || flo > 1.0 - margin || fhi < margin || fhi > 1.0 - margin) {
            dphi = atan2(sqrt(ysq), x);
            ulo = fnr * kInvTwoPi * (g.phi - dphi) - shift;
            uhi = fnr * kInvTwoPi * (g.phi + dphi) - shift;
        }
        if (dphi > 0) {
            const int ip_lo = (int)floor(ulo) + 1;
            const int ip_hi = (int)floor(uhi);
            if (ip_lo <= ip_hi) {
                o.col0 = ip_lo;
                o.cnt = ip_hi - ip_lo + 1;
            }
        }
    }
    const double s = sqrt((1 - z) * (1 + z));
    double A = (1.0 - (z * g.z0 + s * g.s0)) * 0.5;
    if (A < 0) A = 0;
    o.A = (float)A;
    o.B = (float)(s * g.s0);
    if (g.pv == 0) {
        ring_quad(o.A, o.B, g.K4, o.e0, o.e1, o.e2);
    } else {
        o.e0 = o.A;
        o.e1 = o.B;
        o.e2 = 0.0f;
    }
    o.step = (float)kTwoPi / (float)o.nr;
    o.offm = (float)(((double)kc + shift) - a) * o.step;
    o.kc_rel = kc - o.col0;
    return o;
}

// Address parameters of a ring whose first footprint pixel has flat index
// pre_start: pixel(k) = k + abase - (k >= kwrap ? nr : 0).  Columns run
// col0 .. col0+cnt-1 unwrapped, with col0 in [-nr/2-1, nr) and cnt <= nr.
DMANN_HD void ring_addr(const RingRec &rr, int pre_start, int &abase, int &kwrap)
{
    abase = rr.startpix + rr.col0 - pre_start;
    if (rr.col0 < 0) {
        abase += rr.nr;
        kwrap = pre_start - rr.col0;
    } else {
        kwrap = pre_start + rr.nr - rr.col0;
    }
}

DMANN_HD int ring_pixel(int abase, int kwrap, int nr, int k) { return k + abase - (k >= kwrap ? nr : 0); }

DMANN_HD float fast_exp2(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_exp2f(x);
#else
    return exp2f(x);
#endif
}

// Gaussian weight of a pixel at longitude offset x from the particle, in a
// ring at colatitude theta_r; theta is the particle's colatitude:
//   h = hav(psi) = A + B sin^2(x/2),  A = sin^2((theta_r-theta)/2),
//   B = sin(theta_r) sin(theta);  weight = exp(-0.5 (psi/r)^2 / 0.333)
//   (src/genmap.cpp:137-143) = exp2(K4 (psi/2)^2).
// MODE 0 (psi <= 0.5, |x| <= 0.5): (psi/2)^2 = h g(h) by polynomials.
// MODE 2 (psi <= 0.5, any x in [-pi - step, pi + step]): sin(x/2) from its
//   Taylor series through the y^13 term (y = x/2, |y| <= pi/2 + step/2:
//   truncation below 7e-10 relative), squared; then as MODE 0.
// MODE 1 (any geometry): with hp = sin^2(x/2), cp = cos^2(x/2),
//   h = A cp + E hp and 1 - h = Ac cp + D hp (E, D = sin^2, cos^2 of
//   (theta_r+theta)/2; Ac = 1 - A): sums of non-negative terms, so
//   psi/2 = atan2(sqrt h, sqrt(1-h)) keeps its relative accuracy up to the
//   antipode.  The general form takes (A, Ac, E, D) in (A, B, Ac, D) with
//   E passed in B.
template <int MODE>
DMANN_HD float weight32(float x, float A, float B, float Ac, float D, float K4)
{
    if (MODE == 0) {
        // sin^2(x/2) = u (1/4 - u/48 + u^2/1440 - u^3/80640), u = x^2, |x| <= 0.5
        const float u = x * x;
        float p = fmaf(u, -1.0f / 80640.0f, 1.0f / 1440.0f);
        p = fmaf(u, p, -1.0f / 48.0f);
        p = fmaf(u, p, 0.25f);
        const float hphi = u * p;
        const float h = fmaf(B, hphi, A);
        // asin^2(sqrt h)/h = sum_n 2^(2n) h^(n-1) / (2 n^2 C(2n,n)), h <= 0.0612
        float gq = fmaf(h, 128.0f / 2079.0f, 128.0f / 1575.0f);
        gq = fmaf(h, gq, 4.0f / 35.0f);
        gq = fmaf(h, gq, 8.0f / 45.0f);
        gq = fmaf(h, gq, 1.0f / 3.0f);
        gq = fmaf(h, gq, 1.0f);
        return fast_exp2(K4 * (h * gq));
    } else if (MODE == 2) {
        const float y = 0.5f * x;
        const float y2 = y * y;
        float p = fmaf(y2, 1.0f / 6227020800.0f, -1.0f / 39916800.0f);
        p = fmaf(y2, p, 1.0f / 362880.0f);
        p = fmaf(y2, p, -1.0f / 5040.0f);
        p = fmaf(y2, p, 1.0f / 120.0f);
        p = fmaf(y2, p, -1.0f / 6.0f);
        p = fmaf(y2, p, 1.0f);
        const float sh = y * p;
        const float h = fmaf(B, sh * sh, A);
        float gq = fmaf(h, 128.0f / 2079.0f, 128.0f / 1575.0f);
        gq = fmaf(h, gq, 4.0f / 35.0f);
        gq = fmaf(h, gq, 8.0f / 45.0f);
        gq = fmaf(h, gq, 1.0f / 3.0f);
        gq = fmaf(h, gq, 1.0f);
        return fast_exp2(K4 * (h * gq));
    } else {
        // sin^2 and cos^2 of y = x/2 (|y| <= 3 pi / 4: |x| <= pi + step,
        // step <= pi / 2): with r = |y|, or pi/2 - |y| (a two-part pi/2,
        // exact subtraction) when |y| > pi/4, so |r| <= pi/4 and the Taylor
        // series of sin r through r^11 and cos r through r^12 are exact to
        // 2e-11; the squares do not need the signs.
        const float ay = fabsf(0.5f * x);
        const bool big = ay > 0.78539816f;
        const float r = big ? (1.57079637f - ay) + -4.37113883e-8f : ay;
        const float v = r * r;
        float sp = fmaf(v, -1.0f / 39916800.0f, 1.0f / 362880.0f);
        sp = fmaf(v, sp, -1.0f / 5040.0f);
        sp = fmaf(v, sp, 1.0f / 120.0f);
        sp = fmaf(v, sp, -1.0f / 6.0f);
        sp = fmaf(v, sp, 1.0f);
        const float sr = r * sp;
        float cq = fmaf(v, 1.0f / 479001600.0f, -1.0f / 3628800.0f);
        cq = fmaf(v, cq, 1.0f / 40320.0f);
        cq = fmaf(v, cq, -1.0f / 720.0f);
        cq = fmaf(v, cq, 1.0f / 24.0f);
        cq = fmaf(v, cq, -0.5f);
        const float cr = fmaf(v, cq, 1.0f);
        const float sy = big ? cr : sr, cy = big ? sr : cr;
        const float hp = sy * sy, cp = cy * cy;
        const float h = fmaf(B, hp, A * cp);
        const float c = fmaf(D, hp, Ac * cp);
        // psi/2 = atan2(sqrt h, sqrt c), h + c = 1: asin(sqrt m) of
        // m = min(h, c) <= 1/2, or pi/2 minus it
        const float as = asin_sqrt(fminf(h, c));
        const float a = h <= c ? as : 1.57079637f - as;
        return fast_exp2(K4 * (a * a));
    }
}

// The fast form with the series lengths of fast_variant (PV 2 is
// weight32<0>; PV 1 drops the terms below 2e-8 relative), with the sin^2
// coefficients folded into the ring's B once per ring segment (ring_coef):
//   h = A + u (B/4 - u B/48 + u^2 B/1440),  u = x^2.
// PV 0 evaluates the ring's exponent polynomial (RingLite e0..e2): a pixel
// costs 4 VALU operations + exp2.
struct RingCoef {
    float A, c1, c2, c3;
};
DMANN_HD RingCoef ring_coef(float A, float B)
{
    RingCoef c;
    c.A = A;
    c.c1 = B * 0.25f;
    c.c2 = B * (-1.0f / 48.0f);
    c.c3 = B * (1.0f / 1440.0f);
    return c;
}
template <int PV>
DMANN_HD float weight_coef(float x, const RingCoef &c, float K4)
{
    if constexpr (PV == 2) {
        return weight32<0>(x, c.A, c.c1 * 4.0f, 0.0f, 0.0f, K4);
    } else {
        const float u = x * x;
        const float p = fmaf(u, fmaf(u, c.c3, c.c2), c.c1);
        const float h = fmaf(u, p, c.A);
        float gq = fmaf(h, K4 * (8.0f / 45.0f), K4 * (1.0f / 3.0f));
        gq = fmaf(h, gq, K4);
        return fast_exp2(h * gq);
    }
}

// Weight of a pixel at offset x on a ring with terms (e0, e1, e2) (RingLite)
// in series form PV.
template <int PV>
DMANN_HD float ring_term_weight(float x, float e0, float e1, float e2, float K4)
{
    if constexpr (PV == 0) {
        const float u = x * x;
        return fast_exp2(fmaf(u, fmaf(u, e2, e1), e0));
    } else {
        return weight_coef<PV>(x, ring_coef(e0, e1), K4);
    }
}

// Pixels per segment of the grouped kernel's ring walks.
#ifndef DMANN_GROUP_SEG
#define DMANN_GROUP_SEG 7
#endif
constexpr int kSeg = DMANN_GROUP_SEG;

// Longitude offset of pixel t of a RingLite ring as the grouped kernel
// forms it: from the first pixel of t's segment, then t's place in it.
DMANN_HD float seg_x(const RingLite &rl, int t)
{
    const int t0 = t / kSeg * kSeg;
    const float x0 = fmaf((float)(t0 - rl.kc_rel), rl.step, rl.offm);
    return fmaf((float)(t - t0), rl.step, x0);
}

DMANN_HD float lite_weight32(const PGeo &g, const RingLite &rl, int t)
{
    const float x = seg_x(rl, t);
    switch (g.pv) {
    case 0: return ring_term_weight<0>(x, rl.e0, rl.e1, rl.e2, g.K4);
    case 1: return ring_term_weight<1>(x, rl.e0, rl.e1, rl.e2, g.K4);
    default: return ring_term_weight<2>(x, rl.e0, rl.e1, rl.e2, g.K4);
    }
}

DMANN_HD float ring_weight32(const PGeo &g, const RingRec &rr, int t)
{
    const float kf = (float)(t - rr.kc_rel);
    const float x = fmaf(kf, rr.step, rr.offm);
    // the sky-covering kernel's forms: full series, sin series, or the
    // general form
    if (g.mode == 1) return weight32<1>(x, rr.A, rr.E, rr.Ac, rr.D, g.K4);
    if (g.mode == 2) return weight32<2>(x, rr.A, rr.B, 0.0f, 0.0f, g.K4);
    return weight32<0>(x, rr.A, rr.B, 0.0f, 0.0f, g.K4);
}

} // namespace dmann
