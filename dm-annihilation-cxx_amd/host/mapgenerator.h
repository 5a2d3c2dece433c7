// MapGenerator with the public interface of the reference class
// (src/mapgenerator.h:8-26).  start() is defined in genmap_hip.cpp, the HIP
// drop-in for src/genmap_cuda_render.cpp / src/genmap.cpp.
#pragma once

#include <string>

#include "datareader.h"
#include "parameters.h"

class MapGenerator {
public:
    MapGenerator(Parameters *par, DataReader *reader);
    ~MapGenerator();
    void start();
    void saveToFits(std::string fits_file, bool isVerbose = false);
    bool isFinished();
    MAPTYPE *getMap();
    int getNside();

private:
    Parameters *par_;
    DataReader *reader_;
    bool isFinished_;
    MAPTYPE *map_;
};
