// annmap_hip: command-line front end of the map generator.
//
// Command-line contract of the reference's map executables (src/main.cpp):
//   -info <file>   the parameter (ini) file; VL2_info.ini when absent
//   -verbose       echo the file name and the parameters, and ask before an
//                  existing output map is replaced
//   anything else  print the usage text and exit with status 0
// The run reads the particles the parameter file names (native file, or
// tipsy + smooth side files, with its test truncation and mask) in chunks of
// mem_buffer, renders the map on the GPU and writes it as a HEALPix FITS
// file to the file's outputmap.  The reference then shells out to an
// external map2tga image converter; that step is not part of the map path
// and is not run.
#include <cstdio>
#include <iostream>
#include <memory>
#include <string>

#include "datareader.h"
#include "mapgenerator.h"
#include "parameters.h"

namespace {

struct CommandLine {
    std::string info_file = "VL2_info.ini";
    bool verbose = false;
};

void print_usage()
{
    std::cout << "Usage:" << std::endl
              << "-info name: setup info file path. Default: VL2_info.ini" << std::endl
              << "-help: watching help" << std::endl;
}

// False when an argument is not understood (the usage text is due).
bool read_command_line(int argc, const char **argv, CommandLine &cl)
{
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-verbose") {
            cl.verbose = true;
        } else if (a == "-info" && i + 1 < argc) {
            cl.info_file = argv[++i];
        } else {
            return false;
        }
    }
    return true;
}

// The particle source the parameters name.
DataReader *particle_reader(const Parameters &par)
{
    if (par.isNative) return new DataReader(par.nativeDatafile, par.isMask, par.maskFileName);
    return new DataReader(par.partFile, par.densFile, par.hsmFile, par.dspFile, par.isMask, par.maskFileName);
}

} // namespace

int main(int argc, const char **argv)
{
    CommandLine cl;
    if (!read_command_line(argc, argv, cl)) {
        print_usage();
        return 0;
    }
    if (cl.verbose) std::cout << "INFO FILE: " << cl.info_file << std::endl;

    Parameters par(cl.info_file);
    if (cl.verbose) par.printParameters();

    std::unique_ptr<DataReader> reader(particle_reader(par));
    reader->setTest(par.testNum);
    reader->setBuf(par.memParts);
    reader->open();

    MapGenerator generator(&par, reader.get());
    printf("Start generating dark matter map:...\n");
    generator.start();
    reader->close();

    generator.saveToFits(par.outputFileName, cl.verbose);
    printf("Finished....\nMap saved as fits file in %s.\n", par.outputFileName.c_str());
    return 0;
}
