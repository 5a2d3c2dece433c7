// Data model of the drop-in, with the reference's class names and fields
// (src/datatypes.h:29-257) so that reference callers compile unchanged
// against it.  Only the types the map path touches are kept.
#pragma once

#include <string>

#include "dmann.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#ifndef MAX_NUM_MASS_RES
#define MAX_NUM_MASS_RES 10
#endif

const double PI = M_PI;

typedef double MAPTYPE; // map scalar (src/datatypes.h:29)

// 12 floats, 48 bytes: the native particle record (src/datatypes.h:46-105).
class DMParticle {
public:
    float mass;
    float dens;
    float hsmooth;
    float posx;
    float posy;
    float posz;
    float velx;
    float vely;
    float velz;
    float sigmav;
    float eps;
    float phi;
};
static_assert(sizeof(DMParticle) == sizeof(dmann_particle), "DMParticle must match dmann_particle");

class Halo {
public:
    double Mvir_in_Msun = 0, M200_in_Msun = 0, M200crit_in_Msun = 0;
    double Rvir_in_Mpc = 0, R200_in_Mpc = 0, R200crit_in_Mpc = 0;
    double Vmax_in_kms = 0, RVmax_in_kpc = 0, rconverg_in_kpc = 0;
    double params_NFW[3] = {0, 0, 0};
    double params_GNFW[3] = {0, 0, 0};
    double params_Einasto[3] = {0, 0, 0};
    double shape_r = 0, shape = 0;
};

class Info {
public:
    double Omega_M = 0, h100 = 0, n_s = 0, sigma_8 = 0, Lbox_in_Mpc = 0, mass_to_Msun = -1;
    double particle_masses[MAX_NUM_MASS_RES] = {0};
    long particle_numbers[MAX_NUM_MASS_RES] = {0};
    double centerpos[3] = {0, 0, 0};
    double centervel[3] = {0, 0, 0};
    Halo haloInfo;
};

class Codeunits {
public:
    double mass_to_cgs = 0, mass_to_Msun = 0, length_to_Mpc = 0, length_to_cgs = 0;
    double time_to_cgs = 0, velocity_to_cgs = 0, density_to_cgs = 0, annihilation_flux_to_cgs = 0;
};

class Units {
public:
    static const double Msun_in_cgs; // 1.9889212e+33
    static const double Mpc_in_cgs;  // 3.08568025e+24
    static const double eV_in_cgs;   // 1.602e-12
};

class Natconst {
public:
    double h100 = 0;
    static const double G_in_cgs; // 6.67259e-8
    double rho_crit_in_cgs = 0;
    double Omega_m = 0;
    double Delta_vir = 0;
    double Rvir_MW_in_Mpc = 0;
    static const double c_in_cgs; // 2.99792458e10
};

class Dm {
public:
    double M_in_GeV = 0, M_in_cgs = 0, sigma_v_in_cgs = 0;
};

class Params {
public:
    double z = 0;
    double cpos[3] = {0, 0, 0};
    double cvel[3] = {0, 0, 0};
    double opos[3] = {0, 0, 0};
    double otheta = 0, ophi = 0;
    double Lbox_in_Mpc = 0;
    int particle_numbers[MAX_NUM_MASS_RES] = {0};
    double particle_masses[MAX_NUM_MASS_RES] = {0};
    double particle_masses_in_Msun[MAX_NUM_MASS_RES] = {0};
};

class Map {
public:
    std::string projection;
    int Nside = 0;
    int Npix = 0;
    double dOmega = 0;
    double theta0 = 0;
    void setNside(int nside)
    {
        Nside = nside;
        Npix = 12 * Nside * Nside;
        dOmega = 4 * PI / Npix;
        theta0 = 0.0;
    }
};
