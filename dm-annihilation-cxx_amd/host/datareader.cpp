// Native particle-file reader (chunked), see datareader.h.
#include "datareader.h"

#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

DataReader::DataReader(std::string path, bool isMask, std::string maskfile)
    : filePath_(path), maskPath_(maskfile), isMask_(isMask)
{
}

DataReader::~DataReader() { close(); }

void DataReader::setTest(int tn) { testnum_ = tn; }

void DataReader::setBuf(unsigned int mem)
{
    delete[] buffer_;
    buffer_ = nullptr;
    memParts_ = 0;
    memBuffer_ = mem;
}

bool DataReader::tryOpen(std::string *err)
{
    close();
    readCursor_ = 0;
    memParts_ = 0;
    fp_ = fopen(filePath_.c_str(), "rb");
    if (!fp_) {
        if (err) *err = "Datafile error: " + filePath_ + " !";
        return false;
    }
    int n = 0;
    if (fread(&n, sizeof(int), 1, fp_) != 1) {
        if (err) *err = "Datafile error: " + filePath_ + " has no header";
        close();
        return false;
    }
    partNumbers_ = (unsigned)n;
    std::cerr << "Particles: " << partNumbers_ << std::endl;
    if (isMask_) {
        mask_ = fopen(maskPath_.c_str(), "rb");
        int np = -1;
        if (!mask_ || fread(&np, sizeof(int), 1, mask_) != 1 || np != (int)partNumbers_) {
            if (err) *err = "Particle numbers in mask file not match.";
            close();
            return false;
        }
    }
    if (testnum_ != -1 && (int)partNumbers_ > testnum_) partNumbers_ = testnum_;
    buffer_ = new DMParticle[memBuffer_ ? memBuffer_ : 1];
    loadBuffer();
    return true;
}

bool DataReader::open()
{
    std::string err;
    if (!tryOpen(&err)) {
        fprintf(stderr, "%s\n", err.c_str());
        exit(1);
    }
    return true;
}

// Advances to the next chunk of at most memBuffer_ particles.
void DataReader::loadBuffer()
{
    readCursor_ += memParts_;
    const unsigned rest = partNumbers_ > readCursor_ ? partNumbers_ - readCursor_ : 0;
    memParts_ = rest > memBuffer_ ? memBuffer_ : rest;
    if (!fp_ || memParts_ == 0) return;
    const size_t got = fread(buffer_, sizeof(DMParticle), memParts_, fp_);
    if (got != memParts_) {
        // a short file ends the stream where its data ends
        memParts_ = (unsigned)got;
        partNumbers_ = readCursor_ + memParts_;
    }
    if (isMask_ && mask_) {
        std::vector<char> m(memParts_);
        if (fread(m.data(), 1, memParts_, mask_) == memParts_)
            for (unsigned i = 0; i < memParts_; i++)
                if (m[i] == 0) buffer_[i].dens = -1;
    }
}

void DataReader::close()
{
    delete[] buffer_;
    buffer_ = nullptr;
    if (fp_) fclose(fp_);
    if (mask_) fclose(mask_);
    fp_ = nullptr;
    mask_ = nullptr;
}
