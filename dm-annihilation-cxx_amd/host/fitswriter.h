// Self-contained HEALPix FITS map writer/reader (no cfitsio, no HEALPix).
//
// Writes what MapGenerator::saveToFits produces through HEALPix C++
// (src/mapgenerator.cpp:30-62, write_Healpix_map_to_fits with
// PLANCK_FLOAT64): an empty primary HDU and one BINTABLE extension with a
// single float64 column "signal" of repeat count 1024 (npix >= 1024 and
// divisible by 1024; otherwise 1), keywords PIXTYPE='HEALPIX',
// ORDERING='RING', NSIDE, FIRSTPIX=0, LASTPIX=npix-1, INDXSCHM='IMPLICIT'.
// Data are big-endian IEEE doubles; every HDU is padded to 2880 bytes
// (FITS standard 4.0).
#pragma once

#include <string>
#include <vector>

// Returns true on success; err receives the reason otherwise.
bool write_healpix_fits(const std::string &path, const double *map, long long npix, int nside,
                        std::string *err = nullptr);
// Reads a map written by write_healpix_fits (or any single-column float64
// RING HEALPix table).  Returns true on success.
bool read_healpix_fits(const std::string &path, std::vector<double> &map, int *nside, std::string *err = nullptr);
