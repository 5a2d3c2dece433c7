// Parameters of one map: ini parsing, unit chain, observer and rotation.
// Arithmetic follows src/parameters.cpp:42-82 (units), :111-152 (observer),
// :183-289 (rotation) operation by operation, so the constants are
// bit-identical to the reference build (checked against it in the tests).
#include "parameters.h"
#include "dmann_params_of.h"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <vector>

const double Units::Msun_in_cgs = 1.9889212e+33;
const double Units::Mpc_in_cgs = 3.08568025e+24;
const double Units::eV_in_cgs = 1.602e-12;
const double Natconst::G_in_cgs = 6.67259e-8;
const double Natconst::c_in_cgs = 2.99792458e10;

Parameters::Parameters() { initialize(); }

Parameters::Parameters(std::string info_file)
{
    if (!parse(info_file)) {
        printf("Info file incorrect!!!\n");
        exit(1);
    }
}

bool Parameters::load(const std::string &info_file, Parameters &out) { return out.parse(info_file); }

void Parameters::initialize()
{
    // src/parameters.cpp:291-308 and the constructor preamble :311-319
    map.setNside(512);
    memParts = 10000000;
    testNum = -1;
    isNative = true;
    isMask = false;
    partFile = densFile = hsmFile = dspFile = nativeDatafile = outputFileName = maskFileName = "";
    redshift = 0;
    for (int i = 0; i < MAX_NUM_MASS_RES; i++) {
        info.particle_masses[i] = 0;
        info.particle_numbers[i] = 0;
    }
    info.mass_to_Msun = -1.0;
    for (int i = 0; i < 3; i++) align_vector[i] = obs_pos[i] = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) rotmatrix[i][j] = 0.0;
}

static void take3(double *dst, char *val, char **save)
{
    for (int k = 0; k < 3 && val; k++) {
        dst[k] = atof(val);
        if (k < 2) val = strtok_r(nullptr, " ", save);
    }
}

bool Parameters::parse(const std::string &info_file)
{
    initialize();
    std::ifstream in(info_file.c_str(), std::ios_base::in);
    if (!in.is_open()) return false;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::vector<char> buf(line.begin(), line.end());
        buf.push_back('\0');
        char *save = nullptr;
        char *tok = strtok_r(buf.data(), " ", &save);
        if (!tok || tok[0] == '#') continue;
        const std::string key = tok;
        char *val = strtok_r(nullptr, " ", &save); // "="
        val = val ? strtok_r(nullptr, " ", &save) : nullptr;
        if (!val) continue;
        if (key == "Omega_M") info.Omega_M = atof(val);
        else if (key == "h100") info.h100 = atof(val);
        else if (key == "n_s") info.n_s = atof(val);
        else if (key == "sigma_8") info.sigma_8 = atof(val);
        else if (key == "Lbox_in_Mpc") info.Lbox_in_Mpc = atof(val);
        else if (key == "mass_to_Msun") info.mass_to_Msun = atof(val);
        else if (key == "particle_masses") {
            int i = 0;
            do {
                info.particle_masses[i++] = atof(val);
                val = strtok_r(nullptr, " ", &save);
            } while (val && i < MAX_NUM_MASS_RES);
        } else if (key == "particle_numbers") {
            int i = 0;
            do {
                info.particle_numbers[i++] = atol(val);
                val = strtok_r(nullptr, " ", &save);
            } while (val && i < MAX_NUM_MASS_RES);
        } else if (key == "centerpos") take3(info.centerpos, val, &save);
        else if (key == "centervel") take3(info.centervel, val, &save);
        else if (key == "Mvir_in_Msun") info.haloInfo.Mvir_in_Msun = atof(val);
        else if (key == "M200_in_Msun") info.haloInfo.M200_in_Msun = atof(val);
        else if (key == "M200crit_in_Msun") info.haloInfo.M200crit_in_Msun = atof(val);
        else if (key == "Rvir_in_Mpc") info.haloInfo.Rvir_in_Mpc = atof(val);
        else if (key == "R200_in_Mpc") info.haloInfo.R200_in_Mpc = atof(val);
        else if (key == "R200crit_in_Mpc") info.haloInfo.R200crit_in_Mpc = atof(val);
        else if (key == "Vmax_in_kms") info.haloInfo.Vmax_in_kms = atof(val);
        else if (key == "RVmax_in_kpc") info.haloInfo.RVmax_in_kpc = atof(val);
        else if (key == "rconverg_in_kpc") info.haloInfo.rconverg_in_kpc = atof(val);
        else if (key == "params_NFW") take3(info.haloInfo.params_NFW, val, &save);
        else if (key == "params_GNFW") take3(info.haloInfo.params_GNFW, val, &save);
        else if (key == "params_Einasto") take3(info.haloInfo.params_Einasto, val, &save);
        else if (key == "align_vec") take3(align_vector, val, &save);
        else if (key == "obs_pos") take3(obs_pos, val, &save);
        else if (key == "nside") map.setNside(atoi(val));
        else if (key == "test") testNum = atoi(val);
        else if (key == "datafile") nativeDatafile = val;
        else if (key == "partfile") { partFile = val; isNative = false; }
        else if (key == "densfile") { densFile = val; isNative = false; }
        else if (key == "hsmfile") { hsmFile = val; isNative = false; }
        else if (key == "vdspfile") { dspFile = val; isNative = false; }
        else if (key == "outputmap") outputFileName = val;
        else if (key == "mem_buffer") memParts = atoi(val);
        else if (key == "mask") { maskFileName = val; isMask = true; }
    }
    setupUnits();
    setupOthers();
    setupHalo();
    setupParams();
    setupRotation();
    return true;
}

void Parameters::setupUnits()
{
    natconst.h100 = info.h100;
    natconst.rho_crit_in_cgs = 1.8788309e-29 * pow(info.h100, 2);
    natconst.Omega_m = info.Omega_M;
    natconst.Rvir_MW_in_Mpc = 0.275;
    // Bryan & Norman 1998 virial overdensity at the (zero) redshift
    const double Om = natconst.Omega_m;
    const double Omz = Om * pow((1.0 + redshift), 3) / (Om * pow((1.0 + redshift), 3) + (1.0 - Om));
    const double x = Omz - 1.0;
    natconst.Delta_vir = (18.0 * PI * PI + 82.0 * x - 39.0 * x * x) / (1.0 + x);
    natconst.rho_crit_in_cgs =
        3.0 * pow((100.0 * natconst.h100 / (units.Mpc_in_cgs * 1e-5)), 2) / (8 * PI * natconst.G_in_cgs);

    codeunits.length_to_Mpc = info.Lbox_in_Mpc;
    codeunits.length_to_cgs = info.Lbox_in_Mpc * units.Mpc_in_cgs;
    codeunits.time_to_cgs = pow((natconst.rho_crit_in_cgs * natconst.G_in_cgs), (-0.5));
    codeunits.density_to_cgs = natconst.rho_crit_in_cgs;
    if (info.mass_to_Msun == -1) {
        // mass unit from critical density times box volume
        codeunits.mass_to_cgs = codeunits.density_to_cgs * pow(codeunits.length_to_cgs, 3);
        codeunits.mass_to_Msun = codeunits.mass_to_cgs / units.Msun_in_cgs;
        info.mass_to_Msun = codeunits.mass_to_Msun;
    } else {
        codeunits.mass_to_Msun = info.mass_to_Msun;
        codeunits.mass_to_cgs = codeunits.mass_to_Msun * units.Msun_in_cgs;
    }
    codeunits.velocity_to_cgs = codeunits.length_to_cgs / codeunits.time_to_cgs;
    codeunits.annihilation_flux_to_cgs =
        (codeunits.density_to_cgs * codeunits.mass_to_cgs / pow(codeunits.length_to_cgs, 2));
}

void Parameters::setupHalo()
{
    halo = info.haloInfo;
    halo.shape_r = 0.0;
    halo.shape = 0.0;
}

void Parameters::setupOthers()
{
    dm.M_in_GeV = 46.0;
    dm.sigma_v_in_cgs = 5.0e-26;
    dm.M_in_cgs = dm.M_in_GeV * 1e9 * units.eV_in_cgs / pow(natconst.c_in_cgs, 2);
}

void Parameters::setupParams()
{
    params.z = redshift;
    for (int i = 0; i < 3; i++) {
        params.cpos[i] = info.centerpos[i];
        params.cvel[i] = info.centervel[i];
        params.opos[i] = obs_pos[i];
    }
    params.Lbox_in_Mpc = info.Lbox_in_Mpc;
    for (int i = 0, j = 0; i < MAX_NUM_MASS_RES; i++) {
        params.particle_numbers[i] = 0;
        params.particle_masses[i] = 0;
        params.particle_masses_in_Msun[i] = 0;
        if (info.particle_masses[i] > 0) {
            params.particle_numbers[j] = (int)info.particle_numbers[i];
            params.particle_masses[j] = info.particle_masses[i];
            params.particle_masses_in_Msun[j] = info.particle_masses[i] * codeunits.mass_to_cgs / units.Msun_in_cgs;
            j++;
        }
    }
    const double dx = params.opos[0] - params.cpos[0];
    const double dy = params.opos[1] - params.cpos[1];
    const double dz = params.opos[2] - params.cpos[2];
    const double radius = sqrt(pow(dx, 2.0) + pow(dy, 2.0) + pow(dz, 2.0));
    double ophi = atan2(dy, dx);
    if (ophi < 0.0) ophi += 2.0 * PI;
    params.otheta = acos(dz / radius);
    params.ophi = ophi;
}

// Rotation taking the halo-centre direction to -x about r x e1 (arbitrary
// axis rotation), then about x so that the align vector is horizontal.
void Parameters::setupRotation()
{
    const double ot = params.otheta, op = params.ophi;
    double u = 0, v = -cos(ot), w = sin(ot) * sin(op);
    const double rr = sqrt(u * u + v * v + w * w);
    u = u / rr;
    v = v / rr;
    w = w / rr;
    const double c = -sin(ot) * cos(op);
    const double t = 1.0 - c;
    const double s = rr;
    double R[3][3];
    if (rr > 0.0) {
        R[0][0] = u * u * t + c;
        R[0][1] = u * v * t + w * s;
        R[0][2] = u * w * t - v * s;
        R[1][0] = v * u * t - w * s;
        R[1][1] = v * v * t + c;
        R[1][2] = v * w * t + u * s;
        R[2][0] = w * u * t + v * s;
        R[2][1] = w * v * t - u * s;
        R[2][2] = w * w * t + c;
    } else {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) R[i][j] = i == j ? 1.0 : 0.0;
        if (cos(op) > 0.0) R[0][0] = R[1][1] = -1.0;
    }
    const double dist = sqrt(pow(params.opos[0] - params.cpos[0], 2.0) + pow(params.opos[1] - params.cpos[1], 2.0) +
                             pow(params.opos[2] - params.cpos[2], 2.0));
    double a[3];
    for (int i = 0; i < 3; i++) a[i] = params.cpos[i] + dist * align_vector[i] - params.opos[i];
    const double yt = a[0] * R[0][1] + a[1] * R[1][1] + a[2] * R[2][1];
    const double zt = a[0] * R[0][2] + a[1] * R[1][2] + a[2] * R[2][2];
    const double g = -fabs(atan2(zt, yt));
    const double X[3][3] = {{1.0, 0.0, 0.0}, {0.0, cos(g), sin(g)}, {0.0, -sin(g), cos(g)}};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double acc = 0;
            for (int k = 0; k < 3; k++) acc += R[i][k] * X[k][j];
            rotmatrix[i][j] = acc;
        }
}

static std::string join(const double *d, int n)
{
    std::string s;
    for (int i = 0; i < n; i++) {
        std::stringstream ss;
        ss << d[i];
        s += ss.str() + " ";
    }
    return s;
}

void Parameters::printParameters()
{
    using std::cout;
    using std::endl;
    cout << "+========================Parameters========================+" << endl;
    cout << "h100 " << natconst.h100 << endl;
    cout << "rho_crit_in_cgs " << natconst.rho_crit_in_cgs << endl;
    cout << "Omega_m " << natconst.Omega_m << endl;
    cout << "Delta_vir " << natconst.Delta_vir << endl;
    cout << "opos " << join(params.opos, 3) << endl;
    cout << "cpos " << join(params.cpos, 3) << endl;
    cout << "otheta " << params.otheta << endl;
    cout << "ophi " << params.ophi << endl;
    cout << "particle_masses " << join(params.particle_masses, 10) << endl;
    cout << "mass_to_cgs " << codeunits.mass_to_cgs << endl;
    cout << "length_to_cgs " << codeunits.length_to_cgs << endl;
    cout << "annihilation_flux_to_cgs " << codeunits.annihilation_flux_to_cgs << endl;
    cout << "Nside " << map.Nside << endl;
    cout << "Npix " << map.Npix << endl;
    cout << "dOmega " << map.dOmega << endl;
    for (int i = 0; i < 3; i++) cout << "rotmatrix " << join(rotmatrix[i], 3) << endl;
    printf("Particle buffer size: %d\n", memParts);
    cout << "Data File (Native): " << nativeDatafile << endl;
    if (isMask) cout << "MaskFile: " << maskFileName << endl;
    cout << "Output Fits: " << outputFileName << endl;
    cout << "+==========================================================+" << endl;
}

// ---------------------------------------------------------------- C ABI

extern "C" int dmann_params_from_ini(const char *ini, int precision, dmann_params *out, int *mem_buffer,
                                     int *test_num)
{
    if (!ini || !out) return DMANN_ERR_ARG;
    Parameters par;
    if (!Parameters::load(ini, par)) return DMANN_ERR_ARG;
    dmann_params_of(par, precision, out);
    if (mem_buffer) *mem_buffer = par.memParts;
    if (test_num) *test_num = par.testNum;
    return DMANN_OK;
}
