// DataReader: chunked particle reader with the public interface of the
// reference class (src/datareader.h:16-112), for both of its formats:
//  * native: int32 N, then N x 48-byte DMParticle records
//    (src/datareader.cpp:113-121,236-237);
//  * tipsy: an XDR (big-endian) tipsy file -- 32-byte header (double time;
//    int nbodies, ndim, nsph, ndark, nstar, pad), nsph gas records of 12
//    floats, then ndark dark records (mass, pos[3], vel[3], eps, phi) --
//    plus three side files written by "smooth" (int32 nbodies, then one
//    native float per body, gas first) for density, smoothing length and
//    velocity dispersion (src/datareader.cpp:45-69,122-191,238-261).
// Mask file: int32 N then N bytes; a zero byte marks the particle with
// dens = -1 (src/datareader.cpp:263-271).
#pragma once

#include <cstdio>
#include <string>

#include "datatypes.h"

class DataReader {
public:
    // native particle file
    DataReader(std::string path, bool isMask = false, std::string maskfile = "");
    // tipsy file + smooth side files (density, hsmooth, sigma-v); with
    // isSingleFile the side files are not read (dens = hsmooth = sigmav = 0)
    DataReader(std::string partfile, std::string densfile, std::string hsmoothfile, std::string sigmafile,
               bool isMask = false, std::string maskfile = "", bool isSingleFile = false);
    ~DataReader();
    DataReader(const DataReader &) = delete;
    DataReader &operator=(const DataReader &) = delete;

    int getPartNum() { return testnum_ != -1 ? testnum_ : (int)partNumbers_; }
    void setTest(int tn);
    void setBuf(unsigned int mem);
    // Opens the file(s) and loads the first chunk.  Prints the reason and
    // exits(1) on a missing or malformed file or a count mismatch, as the
    // reference does; tryOpen() returns false with the reason instead.
    bool open();
    bool tryOpen(std::string *err = nullptr);
    void loadBuffer();
    DMParticle *getBuf() { return buffer_; }
    int getMemparts() { return (int)memParts_; }
    unsigned int getBufSize() { return memParts_; }
    bool hasNext() { return readCursor_ < partNumbers_; }
    bool isOpen() { return fp_ != nullptr; }
    // true when a chunk ended early (truncated file)
    bool truncated() const { return truncated_; }
    void close();

private:
    bool openNative(std::string *err);
    bool openTipsy(std::string *err);
    unsigned readTipsy(unsigned n);

    std::string filePath_, maskPath_;
    std::string densPath_, hsmPath_, sigPath_;
    bool isMask_;
    bool isNative_ = true;
    bool isSingleFile_ = false;
    bool truncated_ = false;
    FILE *fp_ = nullptr;
    FILE *mask_ = nullptr;
    FILE *dens_ = nullptr, *hsm_ = nullptr, *sig_ = nullptr;
    unsigned int partNumbers_ = 0;
    unsigned int readCursor_ = 0;
    unsigned int memBuffer_ = 10000000;
    unsigned int memParts_ = 0;
    int testnum_ = -1;
    DMParticle *buffer_ = nullptr;
};
