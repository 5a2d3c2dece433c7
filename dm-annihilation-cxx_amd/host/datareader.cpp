// Chunked particle reader, native and tipsy formats (see datareader.h).
#include "datareader.h"

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

#include "dmann.h"

namespace {

// XDR is big-endian IEEE; the side files and the native format are in the
// host's own (little-endian) byte order.
uint32_t be32(const unsigned char *b)
{
    return ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | (uint32_t)b[3];
}
float be_float(const unsigned char *b)
{
    const uint32_t u = be32(b);
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
int be_int(const unsigned char *b) { return (int)be32(b); }
double be_double(const unsigned char *b)
{
    const uint64_t u = ((uint64_t)be32(b) << 32) | be32(b + 4);
    double d;
    std::memcpy(&d, &u, 8);
    return d;
}

constexpr long kTipsyHeader = 32;    // double + 6 ints
constexpr long kTipsyGas = 12 * 4;   // gas record
constexpr size_t kTipsyDark = 9 * 4; // mass, pos[3], vel[3], eps, phi

} // namespace

DataReader::DataReader(std::string path, bool isMask, std::string maskfile)
    : filePath_(path), maskPath_(maskfile), isMask_(isMask)
{
}

DataReader::DataReader(std::string partfile, std::string densfile, std::string hsmoothfile, std::string sigmafile,
                       bool isMask, std::string maskfile, bool isSingleFile)
    : filePath_(partfile), maskPath_(maskfile), densPath_(densfile), hsmPath_(hsmoothfile), sigPath_(sigmafile),
      isMask_(isMask), isNative_(false), isSingleFile_(isSingleFile)
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

bool DataReader::openNative(std::string *err)
{
    fp_ = fopen(filePath_.c_str(), "rb");
    if (!fp_) {
        if (err) *err = "Datafile error: " + filePath_ + " !";
        return false;
    }
    int n = 0;
    if (fread(&n, sizeof(int), 1, fp_) != 1) {
        if (err) *err = "Datafile error: " + filePath_ + " has no header";
        return false;
    }
    partNumbers_ = (unsigned)n;
    std::cerr << "Particles: " << partNumbers_ << std::endl;
    return true;
}

bool DataReader::openTipsy(std::string *err)
{
    fp_ = fopen(filePath_.c_str(), "rb");
    if (!fp_) {
        if (err) *err = "Problem opening file: " + filePath_;
        return false;
    }
    unsigned char h[kTipsyHeader];
    if (fread(h, 1, sizeof h, fp_) != sizeof h) {
        if (err) *err = "Problem reading tipsy header.";
        return false;
    }
    const double time = be_double(h);
    const int nbodies = be_int(h + 8), nsph = be_int(h + 16), ndark = be_int(h + 20), nstar = be_int(h + 24);
    if (nbodies < 0 || nsph < 0 || ndark < 0 || nstar < 0) {
        if (err) *err = "Problem reading tipsy header.";
        return false;
    }
    std::cerr << "Time: " << time << std::endl;
    std::cerr << "# of bodies: " << nbodies << std::endl;
    std::cerr << "# of gas part: " << nsph << std::endl;
    std::cerr << "# of star part: " << nstar << std::endl;
    std::cerr << "# of DM part: " << ndark << std::endl;
    if (nsph != 0) fseek(fp_, kTipsyHeader + (long)nsph * kTipsyGas, SEEK_SET);
    partNumbers_ = (unsigned)ndark;
    if (!isSingleFile_) {
        const char *what[3] = {"density", "hsmooth", "sigmav"};
        const std::string *path[3] = {&densPath_, &hsmPath_, &sigPath_};
        FILE **fh[3] = {&dens_, &hsm_, &sig_};
        for (int k = 0; k < 3; k++) {
            *fh[k] = fopen(path[k]->c_str(), "rb");
            int np = -1;
            if (!*fh[k] || fread(&np, sizeof np, 1, *fh[k]) != 1 || np != nbodies) {
                if (err) *err = std::string("Particle numbers in ") + what[k] + " not match.";
                return false;
            }
            if (nsph != 0) fseek(*fh[k], (long)nsph * (long)sizeof(float), SEEK_CUR);
        }
    }
    return true;
}

bool DataReader::tryOpen(std::string *err)
{
    close();
    readCursor_ = 0;
    memParts_ = 0;
    truncated_ = false;
    const bool ok = isNative_ ? openNative(err) : openTipsy(err);
    if (!ok) {
        close();
        return false;
    }
    if (isMask_) {
        mask_ = fopen(maskPath_.c_str(), "rb");
        int np = -1;
        if (!mask_ || fread(&np, sizeof(int), 1, mask_) != 1 || np != (int)partNumbers_) {
            if (err) *err = "Particle numbers in mask file not match.";
            close();
            return false;
        }
    }
    std::cerr << "Load to buffer ... " << std::endl;
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

// n dark particles of the tipsy file and their smooth values into buffer_.
unsigned DataReader::readTipsy(unsigned n)
{
    std::vector<unsigned char> raw((size_t)n * kTipsyDark);
    const unsigned got = (unsigned)(fread(raw.data(), kTipsyDark, n, fp_));
    std::vector<float> side[3];
    FILE *fh[3] = {dens_, hsm_, sig_};
    for (int k = 0; k < 3; k++) {
        side[k].assign(n, 0.0f);
        if (!isSingleFile_ && fh[k] && fread(side[k].data(), sizeof(float), got, fh[k]) != got) truncated_ = true;
    }
    for (unsigned i = 0; i < got; i++) {
        const unsigned char *b = raw.data() + (size_t)i * kTipsyDark;
        DMParticle &p = buffer_[i];
        p.mass = be_float(b);
        p.posx = be_float(b + 4);
        p.posy = be_float(b + 8);
        p.posz = be_float(b + 12);
        p.velx = be_float(b + 16);
        p.vely = be_float(b + 20);
        p.velz = be_float(b + 24);
        p.eps = be_float(b + 28);
        p.phi = be_float(b + 32);
        p.dens = side[0][i];
        p.hsmooth = side[1][i];
        p.sigmav = side[2][i];
    }
    return got;
}

// Advances to the next chunk of at most memBuffer_ particles.
void DataReader::loadBuffer()
{
    readCursor_ += memParts_;
    const unsigned rest = partNumbers_ > readCursor_ ? partNumbers_ - readCursor_ : 0;
    memParts_ = rest > memBuffer_ ? memBuffer_ : rest;
    if (!fp_ || memParts_ == 0) return;
    const unsigned got =
        isNative_ ? (unsigned)fread(buffer_, sizeof(DMParticle), memParts_, fp_) : readTipsy(memParts_);
    if (got != memParts_) {
        // a short file ends the stream where its data ends
        truncated_ = true;
        memParts_ = got;
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
    FILE **fs[5] = {&fp_, &mask_, &dens_, &hsm_, &sig_};
    for (FILE **f : fs) {
        if (*f) fclose(*f);
        *f = nullptr;
    }
}

// C ABI: the particles an ini file names, read as the map generator reads
// them (see include/dmann.h).
#include "parameters.h"

extern "C" int64_t dmann_read_particles(const char *ini, dmann_particle *out, int64_t cap)
{
    if (!ini || cap < 0) return -DMANN_ERR_ARG;
    Parameters par;
    if (!Parameters::load(ini, par)) return -DMANN_ERR_ARG;
    DataReader r = par.isNative ? DataReader(par.nativeDatafile, par.isMask, par.maskFileName)
                                : DataReader(par.partFile, par.densFile, par.hsmFile, par.dspFile, par.isMask,
                                             par.maskFileName);
    r.setTest(par.testNum);
    r.setBuf(par.memParts > 0 ? (unsigned)par.memParts : 1u);
    std::string err;
    if (!r.tryOpen(&err)) {
        fprintf(stderr, "%s\n", err.c_str());
        return -DMANN_ERR_ARG;
    }
    int64_t n = 0;
    while (r.hasNext()) {
        const int m = r.getMemparts();
        if (m <= 0) break;
        if (out && n < cap) {
            const int64_t k = m < cap - n ? m : cap - n;
            std::memcpy(out + n, r.getBuf(), (size_t)k * sizeof(DMParticle));
        }
        n += m;
        r.loadBuffer();
    }
    const bool bad = r.truncated();
    r.close();
    return bad ? -DMANN_ERR_DOMAIN : n;
}

// C ABI: halo file rows (xc yc zc radius), read as floats in order as
// src/halocore/halocore.cpp:17-41 reads them (only complete rows count).
#include <fstream>

extern "C" int64_t dmann_read_halo_file(const char *path, float *out, int64_t cap)
{
    if (!path) return -1;
    std::ifstream in(path);
    if (!in.good()) return -1;
    int64_t n = 0;
    float v[4];
    while (in >> v[0] >> v[1] >> v[2] >> v[3]) {
        if (out && n < cap)
            for (int k = 0; k < 4; k++) out[4 * n + k] = v[k];
        n++;
    }
    return n;
}
