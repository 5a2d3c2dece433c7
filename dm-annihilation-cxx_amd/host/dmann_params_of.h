// C-ABI map constants of a Parameters object, from its public members only
// (src/parameters.h:6-45), so that the drop-in generator compiles against
// the reference's own Parameters class as well as this package's mirror.
#pragma once

#include <cfloat>
#include <cmath>

#include "dmann.h"

template <class Par>
inline void dmann_params_of(const Par &par, int precision, dmann_params *p)
{
    for (int i = 0; i < 3; i++) p->opos[i] = par.params.opos[i];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) p->rotmatrix[3 * i + j] = par.rotmatrix[i][j];
    // unit factor of the nosom plug-in, src/flux/flux_ann_nosom.cpp:13-16
    p->flux_factor = std::pow(std::pow((par.natconst.c_in_cgs), 2) / (par.units.eV_in_cgs * 1.0e9), 2) /
                     (par.units.Mpc_in_cgs * 1.0e-3) * par.codeunits.annihilation_flux_to_cgs;
    // high-resolution cut of src/genmap.cpp:52-58 (a float minimum over the
    // positive particle masses) and the skip test of :82-85
    float hires = FLT_MAX;
    for (int i = 0; i < 10; i++)
        if (par.params.particle_masses[i] > 0 && par.params.particle_masses[i] < hires)
            hires = (float)par.params.particle_masses[i];
    p->mass_cut = (double)hires * 1.1;
    p->nside = par.map.Nside;
    p->precision = precision;
    p->flux_kind = DMANN_FLUX_ANN_NOSOM;
    p->reserved = 0;
    // unit factor of src/flux/flux_decay.cpp:11-14
    const double decay_flux_to_cgs = (par.codeunits.mass_to_cgs / std::pow(par.codeunits.length_to_cgs, 2));
    p->decay_factor = std::pow((par.natconst.c_in_cgs), 2) / (par.units.eV_in_cgs * 1.0e9) /
                      (par.units.Mpc_in_cgs * 1.0e-3) * decay_flux_to_cgs;
    p->vel_to_cgs = par.codeunits.velocity_to_cgs;
    p->c_in_cgs = par.natconst.c_in_cgs;
    for (int i = 0; i < 3; i++) p->cpos[i] = par.params.cpos[i];
    p->footprint_pad = 0.0;
}
