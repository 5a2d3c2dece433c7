// DataReader: chunked reader of the native particle file, with the public
// interface of the reference class (src/datareader.h:16-112).
// Native format: int32 N, then N x 48-byte DMParticle records
// (src/datareader.cpp:113-121,236-237).  Mask file: int32 N then N bytes;
// a zero byte marks the particle with dens = -1 (src/datareader.cpp:263-271).
// The tipsy/XDR constructor of the reference is not provided: that format is
// upstream of the map path.
#pragma once

#include <cstdio>
#include <string>

#include "datatypes.h"

class DataReader {
public:
    DataReader(std::string path, bool isMask = false, std::string maskfile = "");
    ~DataReader();

    int getPartNum() { return testnum_ != -1 ? testnum_ : (int)partNumbers_; }
    void setTest(int tn);
    void setBuf(unsigned int mem);
    // Opens the file and loads the first chunk.  Prints the reason and
    // exits(1) on a missing file or a mask count mismatch, as the reference
    // does (src/datareader.cpp:116-119,263-271); tryOpen() returns false.
    bool open();
    bool tryOpen(std::string *err = nullptr);
    void loadBuffer();
    DMParticle *getBuf() { return buffer_; }
    int getMemparts() { return (int)memParts_; }
    unsigned int getBufSize() { return memParts_; }
    bool hasNext() { return readCursor_ < partNumbers_; }
    bool isOpen() { return fp_ != nullptr; }
    void close();

private:
    std::string filePath_, maskPath_;
    bool isMask_;
    FILE *fp_ = nullptr;
    FILE *mask_ = nullptr;
    unsigned int partNumbers_ = 0;
    unsigned int readCursor_ = 0;
    unsigned int memBuffer_ = 10000000;
    unsigned int memParts_ = 0;
    int testnum_ = -1;
    DMParticle *buffer_ = nullptr;
};
