// MapGenerator bookkeeping and FITS output (src/mapgenerator.cpp:18-74).
// Uses only the public Parameters members and the library's C ABI, so it
// builds against the reference's headers too (the reference's own
// mapgenerator.cpp needs HEALPix C++ for its FITS writer).
#include <rpc/types.h>
#include <rpc/xdr.h>

#include "mapgenerator.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>

#include "dmann.h"

MapGenerator::MapGenerator(Parameters *par, DataReader *reader)
    : par_(par), reader_(reader), isFinished_(false), map_(nullptr)
{
}

MapGenerator::~MapGenerator()
{
    if (map_ != nullptr) free(map_); // map_ is malloc'd, as the reference frees it
}

// Writes the RING float64 map to par_->outputFileName (the argument is
// ignored, as in src/mapgenerator.cpp:40); an existing file is replaced,
// after a 'y' on stdin when isVerbose.
void MapGenerator::saveToFits(std::string, bool isVerbose)
{
    if (!isFinished()) {
        printf("Map not generated yet");
        return;
    }
    const std::string fits = par_->outputFileName;
    std::ifstream probe(fits.c_str());
    if (probe) {
        probe.close();
        std::cout << "File exists! Owerite..." << std::endl;
        char t = 0;
        if (isVerbose) std::cin >> t;
        if (!(t == 'y' || !isVerbose)) return;
        remove(fits.c_str());
    } else {
        std::cout << "Creating fits!..." << std::endl;
    }
    if (dmann_write_healpix_fits(fits.c_str(), map_, (int64_t)par_->map.Npix, par_->map.Nside) != DMANN_OK)
        fprintf(stderr, "cannot write %s\n", fits.c_str());
}

bool MapGenerator::isFinished() { return isFinished_; }

MAPTYPE *MapGenerator::getMap() { return map_; }

int MapGenerator::getNside() { return par_->map.Nside; }
