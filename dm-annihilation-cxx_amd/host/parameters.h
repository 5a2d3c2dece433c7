// Parameters: ini reader and derived constants of one map, with the public
// interface of the reference class (src/parameters.h:4-57).
#pragma once

#include <string>

#include "datatypes.h"

class Parameters {
public:
    // Reads `info_file` (key = value lines, '#' comments) and derives units,
    // observer, rotation.  Prints "Info file incorrect!!!" and exits(1) when
    // the file cannot be opened, as src/parameters.cpp:471-474 does.
    Parameters(std::string info_file);
    // Non-exiting form used by the C ABI: returns false when unreadable.
    static bool load(const std::string &info_file, Parameters &out);
    Parameters();

    Info info;
    Units units;
    Natconst natconst;
    Halo halo;
    Codeunits codeunits;
    Params params;
    Dm dm;
    Map map;
    double rotmatrix[3][3];

    int memParts;
    int testNum;

    void printParameters();

    std::string partFile, densFile, hsmFile, dspFile;
    std::string nativeDatafile;
    std::string maskFileName;
    std::string outputFileName;

    bool isNative;
    bool isMask;

private:
    double align_vector[3];
    double obs_pos[3];

    void setupUnits();
    void setupParams();
    void setupHalo();
    void setupRotation();
    void setupOthers();
    void initialize();
    bool parse(const std::string &info_file);

    double redshift;
};
