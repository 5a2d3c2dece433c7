// HEALPix RING geometry used by the projector, for host and device.
//
// Closed forms of Gorski et al. 2005 (ApJ 622:759), in the forms HEALPix C++
// 3.00 uses for Healpix_Base::ring_above / ring2z / get_ring_info_small /
// loc2pix and the non-inclusive query_disc ring walk, which the reference CPU
// generator calls (src/genmap.cpp:40,106,119).  The reference CUDA helpers
// (src/chealpixrender.cu:202-314) compute the same quantities with known
// off-by-one errors (ip > nl4, p > npix) and are not followed.
//
// Pixel indices are int32: nside <= 8192 gives npix <= 805306368 < 2^31.
// The translation unit is compiled with -ffp-contract=off so these FP64
// expressions round exactly as the x86-64 oracle does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DMANN_HD __host__ __device__ __forceinline__

namespace dmann {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 6.283185307179586476925286766559005768394;
constexpr double kHalfPi = 1.570796326794896619231321691639751442099;
constexpr double kInvTwoPi = 1.0 / 6.283185307179586476925286766559005768394;
constexpr double kTwoThird = 2.0 / 3.0;
constexpr int kMaxNside = 8192;

DMANN_HD int64_t npix_of(int nside) { return 12LL * nside * nside; }

// Ring index (1..4nside-1) of the lowest ring with z_ring > z (0 above ring 1).
DMANN_HD int ring_above(int nside, double z)
{
    double az = fabs(z);
    if (az <= kTwoThird) return (int)(nside * (2 - 1.5 * z));
    int iring = (int)(nside * sqrt(3 * (1 - az)));
    return (z > 0) ? iring : 4 * nside - iring - 1;
}

// Factors of ring2z for one nside, computed once per kernel.
struct RingGeo {
    int nside;
    double fact1, fact2;
};

DMANN_HD RingGeo ring_geo(int nside)
{
    RingGeo G;
    G.nside = nside;
    G.fact2 = 4. / (double)npix_of(nside);
    G.fact1 = (double)(2 * nside) * G.fact2;
    return G;
}

// z = cos(theta) of the pixel centres of ring `ring`.
DMANN_HD double ring2z(const RingGeo &G, int ring)
{
    const int nside = G.nside;
    const double fact1 = G.fact1, fact2 = G.fact2;
    int64_t r = ring;
    if (r < nside) return 1 - (double)(r * r) * fact2;
    if (r <= 3 * nside) return (double)(2 * nside - r) * fact1;
    r = 4 * nside - r;
    return (double)(r * r) * fact2 - 1;
}

DMANN_HD double ring2z(int nside, int ring) { return ring2z(ring_geo(nside), ring); }

// First pixel, pixel count and half-pixel phase shift of a ring.
DMANN_HD void ring_info(int nside, int ring, int &startpix, int &ringpix, int &shifted)
{
    if (ring < nside) {
        shifted = 1;
        ringpix = 4 * ring;
        startpix = 2 * ring * (ring - 1);
    } else if (ring < 3 * nside) {
        shifted = ((ring - nside) & 1) == 0;
        ringpix = 4 * nside;
        startpix = 2 * nside * (nside - 1) + (ring - nside) * 4 * nside;
    } else {
        int nr = 4 * nside - ring;
        shifted = 1;
        ringpix = 4 * nr;
        startpix = (int)(npix_of(nside) - 2LL * nr * (nr + 1));
    }
}

DMANN_HD double fmodulo(double v1, double v2)
{
    if (v1 >= 0) return (v1 < v2) ? v1 : fmod(v1, v2);
    double tmp = fmod(v1, v2) + v2;
    return (tmp == v2) ? 0. : tmp;
}

// RING pixel containing the direction (z = cos theta, phi).
DMANN_HD int loc2pix(int nside, double z, double phi, double sth, bool have_sth)
{
    const int ns = nside;
    double za = fabs(z);
    double tt = fmodulo(phi * (1.0 / kHalfPi), 4.0);
    if (za <= kTwoThird) {
        int nl4 = 4 * ns;
        double temp1 = ns * (0.5 + tt);
        double temp2 = ns * z * 0.75;
        int jp = (int)(temp1 - temp2);
        int jm = (int)(temp1 + temp2);
        int ir = ns + 1 + jp - jm;
        int kshift = 1 - (ir & 1);
        int t1 = jp + jm - ns + kshift + 1 + nl4 + nl4;
        int ip = (t1 >> 1) % nl4;
        return 2 * ns * (ns - 1) + (ir - 1) * nl4 + ip;
    }
    double tp = tt - (double)(int)tt;
    double tmp = ((za < 0.99) || (!have_sth)) ? ns * sqrt(3 * (1 - za))
                                              : ns * sth / sqrt((1. + za) / 3.);
    int jp = (int)(tp * tmp);
    int jm = (int)((1.0 - tp) * tmp);
    int ir = jp + jm + 1;
    int ip = (int)(tt * ir);
    return (z > 0) ? 2 * ir * (ir - 1) + ip : (int)(npix_of(nside) - 2LL * ir * (ir + 1)) + ip;
}

// Healpix_Base::ang2pix (RING) for a pointing (theta, phi).
DMANN_HD int ang2pix(int nside, double theta, double phi)
{
    if ((theta < 0.01) || (theta > 3.14159 - 0.01)) return loc2pix(nside, cos(theta), phi, sin(theta), true);
    return loc2pix(nside, cos(theta), phi, 0., false);
}

// Bits 0..11 of v spread to the even bit positions 0..22.
DMANN_HD unsigned spread_bits12(unsigned v)
{
    v &= 0xfffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// NESTED pixel index at nside nc (a power of two, at most 4096) of the
// direction (theta, phi): base face and the bit-interleaved
// position in it, the HEALPix NESTED scheme (Gorski et al. 2005, ApJ 622,
// 759).  Consecutive indices form compact 2-D blocks on the sky (a Z-order
// curve in each base face).  Only the grouping's locality depends on it.
DMANN_HD unsigned nest_cell(int nc, double theta, double phi)
{
    const double z = cos(theta), za = fabs(z);
    const double tt = fmodulo(phi * (1.0 / kHalfPi), 4.0);
    int face, ix, iy;
    if (za <= kTwoThird) {
        const double t1 = nc * (0.5 + tt), t2 = nc * z * 0.75;
        const int jp = (int)(t1 - t2), jm = (int)(t1 + t2);
        const int ifp = jp / nc, ifm = jm / nc;
        face = ifp == ifm ? (ifp | 4) : (ifp < ifm ? ifp : ifm + 8);
        ix = jm & (nc - 1);
        iy = nc - (jp & (nc - 1)) - 1;
    } else {
        const int ntt = (int)tt < 3 ? (int)tt : 3;
        const double tp = tt - ntt;
        const double tmp = nc * sqrt(3.0 * (1.0 - za));
        int jp = (int)(tp * tmp), jm = (int)((1.0 - tp) * tmp);
        jp = jp < nc - 1 ? jp : nc - 1;
        jm = jm < nc - 1 ? jm : nc - 1;
        if (z >= 0) {
            face = ntt;
            ix = nc - jm - 1;
            iy = nc - jp - 1;
        } else {
            face = ntt + 8;
            ix = jp;
            iy = jm;
        }
    }
    return (unsigned)face * (unsigned)nc * (unsigned)nc + (spread_bits12((unsigned)ix) | (spread_bits12((unsigned)iy) << 1));
}

} // namespace dmann
