// annmap_hip: command-line driver with the reference's interface
// (src/main.cpp:16-73): -info <ini> [-verbose].
//
// Differences: only the native particle format is read (the tipsy/XDR
// inputs are upstream of the map path), and the external `map2tga`
// conversion that the reference shells out to is not run.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "datareader.h"
#include "mapgenerator.h"
#include "parameters.h"

int main(int argc, const char **argv)
{
    int m = 1;
    bool verbose = false;
    std::string info_file = "VL2_info.ini";
    while (m < argc) {
        std::string arg = argv[m];
        if (arg == "-info" && m + 1 < argc) {
            info_file = argv[m + 1];
            m += 1;
        } else if (arg == "-verbose") {
            verbose = true;
        } else {
            std::cout << "Usage:" << std::endl;
            std::cout << "-info name: setup info file path. Default: VL2_info.ini" << std::endl;
            std::cout << "-help: watching help" << std::endl;
            exit(0);
        }
        m++;
    }
    if (verbose) std::cout << "INFO FILE: " << info_file << std::endl;
    Parameters par(info_file);
    if (verbose) par.printParameters();
    // src/main.cpp:41-53: native file, or tipsy + smooth side files
    DataReader *reader = par.isNative ? new DataReader(par.nativeDatafile, par.isMask, par.maskFileName)
                                      : new DataReader(par.partFile, par.densFile, par.hsmFile, par.dspFile,
                                                       par.isMask, par.maskFileName);
    reader->setTest(par.testNum);
    reader->setBuf(par.memParts);
    reader->open();
    MapGenerator generator(&par, reader);
    printf("Start generating dark matter map:...\n");
    generator.start();
    reader->close();
    generator.saveToFits(par.outputFileName, verbose);
    printf("Finished....\nMap saved as fits file in %s.\n", par.outputFileName.c_str());
    delete reader;
    return 0;
}
