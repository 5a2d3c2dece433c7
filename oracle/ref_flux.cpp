// Known-answer harness around one of the reference's flux plug-ins.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file with
// src/parameters.cpp, src/datatypes.cpp and ONE of src/flux/flux_*.cpp
// (each defines the same `getflux`, chosen at link time in the reference,
// src/Makefile:130-160), read in place from the reference tree, into
// oracle/_ref/libref_flux_<kind>.so.  Tests load them (when built) to pin
// the oracle's and the product's flux restatements.  Only the reference's
// public interfaces (parameters.h, flux.h) are called.
#include <rpc/types.h>
#include <rpc/xdr.h>

#include <cstring>
#include <string>

#include "datatypes.h"
#include "flux.h"
#include "parameters.h"

extern "C" double ref_flux(const char *ini, const float *rec12, double d)
{
    Parameters par{std::string(ini)};
    DMParticle p;
    std::memcpy(&p, rec12, sizeof(DMParticle));
    return getflux(&par, p, d);
}
