// Known-answer harness around the reference's own translation units.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// src/parameters.cpp, src/datatypes.cpp, src/flux/flux_ann_nosom.cpp,
// src/datareader.cpp and src/tipsy_io.c read in place from the reference
// tree, into oracle/_ref/libref_kat.so.  Tests load it (when it has been
// built) to pin the oracle's Parameters, getflux and native-reader
// restatements against the reference's compiled code.  Nothing here is part
// of the product and nothing of the reference is copied: this file only calls
// the reference's public interfaces (parameters.h, flux.h, datareader.h).
#include <rpc/types.h>
#include <rpc/xdr.h>

#include <cstdint>
#include <cstring>
#include <string>

#include "datareader.h"
#include "datatypes.h"
#include "flux.h"
#include "parameters.h"

extern "C" {

// Layout of `out` (doubles):
//  0..8   rotmatrix[i][j] row-major      9..11  params.opos
// 12..14  params.cpos                   15,16  otheta, ophi
// 17..24  codeunits: mass_to_cgs, mass_to_Msun, length_to_Mpc, length_to_cgs,
//         time_to_cgs, velocity_to_cgs, density_to_cgs, annihilation_flux_to_cgs
// 25      natconst.rho_crit_in_cgs      26 natconst.Delta_vir
// 27..36  params.particle_masses[10]
// 37      map.Nside   38 map.Npix   39 map.dOmega   40 memParts   41 testNum
// 42      isNative    43 isMask
int ref_params(const char *ini, double *out)
{
    Parameters par{std::string(ini)};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[3 * i + j] = par.rotmatrix[i][j];
    for (int i = 0; i < 3; i++) {
        out[9 + i] = par.params.opos[i];
        out[12 + i] = par.params.cpos[i];
    }
    out[15] = par.params.otheta;
    out[16] = par.params.ophi;
    out[17] = par.codeunits.mass_to_cgs;
    out[18] = par.codeunits.mass_to_Msun;
    out[19] = par.codeunits.length_to_Mpc;
    out[20] = par.codeunits.length_to_cgs;
    out[21] = par.codeunits.time_to_cgs;
    out[22] = par.codeunits.velocity_to_cgs;
    out[23] = par.codeunits.density_to_cgs;
    out[24] = par.codeunits.annihilation_flux_to_cgs;
    out[25] = par.natconst.rho_crit_in_cgs;
    out[26] = par.natconst.Delta_vir;
    for (int i = 0; i < 10; i++) out[27 + i] = par.params.particle_masses[i];
    out[37] = par.map.Nside;
    out[38] = par.map.Npix;
    out[39] = par.map.dOmega;
    out[40] = par.memParts;
    out[41] = par.testNum;
    out[42] = par.isNative ? 1 : 0;
    out[43] = par.isMask ? 1 : 0;
    return 0;
}

// getflux of the nosom plug-in for a 12-float particle record at distance d.
double ref_getflux(const char *ini, const float *rec12, double d)
{
    Parameters par{std::string(ini)};
    DMParticle p;
    std::memcpy(&p, rec12, sizeof(DMParticle));
    return getflux(&par, p, d);
}

// Reads a tipsy file + smooth side files (density, hsmooth, sigma-v) through
// DataReader's tipsy constructor and chunked buffer, applying `test` and the
// optional mask, into out (cap records of 12 floats).  Returns the count.
long long ref_read_tipsy(const char *part, const char *dens, const char *hsm, const char *sig, int test,
                         int membuf, const char *mask, float *out, long long cap)
{
    const bool use_mask = mask && *mask;
    DataReader r(std::string(part), std::string(dens), std::string(hsm), std::string(sig), use_mask,
                 std::string(use_mask ? mask : ""));
    r.setTest(test);
    r.setBuf(membuf);
    r.open();
    long long n = 0;
    while (r.hasNext()) {
        const int m = r.getMemparts();
        for (int i = 0; i < m; i++, n++)
            if (n < cap) std::memcpy(out + 12 * n, r.getBuf() + i, sizeof(DMParticle));
        r.loadBuffer();
    }
    r.close();
    return n;
}

// Reads a native particle file through DataReader's chunked buffer
// (membuf particles per chunk), applying `test` and the optional mask,
// into out (cap records of 12 floats).  Returns the number of records.
long long ref_read_native(const char *path, int test, int membuf, const char *mask, float *out,
                          long long cap)
{
    bool isMask = mask && mask[0];
    DataReader reader{std::string(path), isMask, std::string(isMask ? mask : "")};
    reader.setTest(test);
    reader.setBuf((unsigned)membuf);
    reader.open();
    long long n = 0;
    while (reader.hasNext()) {
        DMParticle *buf = reader.getBuf();
        for (int i = 0; i < reader.getMemparts(); i++) {
            if (n < cap) std::memcpy(out + 12 * n, &buf[i], sizeof(DMParticle));
            n++;
        }
        reader.loadBuffer();
    }
    reader.close();
    return n;
}

int ref_sizeof_particle() { return (int)sizeof(DMParticle); }
}
