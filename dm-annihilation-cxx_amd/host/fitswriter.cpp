// HEALPix FITS writer/reader, see fitswriter.h.
#include "fitswriter.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cstdlib>

namespace {

const size_t kBlock = 2880;

std::string card(const char *key, const std::string &value, const char *comment = nullptr)
{
    char buf[81];
    if (comment)
        snprintf(buf, sizeof buf, "%-8.8s= %20s / %s", key, value.c_str(), comment);
    else
        snprintf(buf, sizeof buf, "%-8.8s= %20s", key, value.c_str());
    std::string s(buf);
    s.resize(80, ' ');
    return s;
}

std::string sval(const std::string &v)
{
    std::string q = "'" + v;
    while (q.size() < 9) q += ' ';
    q += "'";
    char buf[32];
    snprintf(buf, sizeof buf, "%-20s", q.c_str());
    return std::string(buf);
}

std::string ival(long long v) { return std::to_string(v); }

void pad(std::string &h)
{
    while (h.size() % kBlock) h += ' ';
}

void put_be64(unsigned char *dst, double v)
{
    uint64_t u;
    std::memcpy(&u, &v, 8);
    for (int i = 0; i < 8; i++) dst[i] = (unsigned char)(u >> (56 - 8 * i));
}

double get_be64(const unsigned char *src)
{
    uint64_t u = 0;
    for (int i = 0; i < 8; i++) u = (u << 8) | src[i];
    double v;
    std::memcpy(&v, &u, 8);
    return v;
}

} // namespace

bool write_healpix_fits(const std::string &path, const double *map, long long npix, int nside, std::string *err)
{
    const long long rep = (npix >= 1024 && npix % 1024 == 0) ? 1024 : 1;
    std::string h;
    h += card("SIMPLE", "T", "file does conform to FITS standard");
    h += card("BITPIX", "8", "number of bits per data pixel");
    h += card("NAXIS", "0", "number of data axes");
    h += card("EXTEND", "T", "FITS dataset may contain extensions");
    h += std::string("END").append(77, ' ');
    pad(h);
    std::string x;
    x += card("XTENSION", sval("BINTABLE"), "binary table extension");
    x += card("BITPIX", "8", "8-bit bytes");
    x += card("NAXIS", "2", "2-dimensional binary table");
    x += card("NAXIS1", ival(8 * rep), "width of table in bytes");
    x += card("NAXIS2", ival(npix / rep), "number of rows in table");
    x += card("PCOUNT", "0", "size of special data area");
    x += card("GCOUNT", "1", "one data group (required keyword)");
    x += card("TFIELDS", "1", "number of fields in each row");
    x += card("TTYPE1", sval("signal"), "label for field   1");
    x += card("TFORM1", sval(std::to_string(rep) + "D"), "data format of field: 8-byte DOUBLE");
    x += card("TUNIT1", sval("unknown"), "physical unit of field");
    x += card("PIXTYPE", sval("HEALPIX"), "HEALPIX pixelisation");
    x += card("ORDERING", sval("RING"), "Pixel ordering scheme, either RING or NESTED");
    x += card("NSIDE", ival(nside), "Resolution parameter for HEALPIX");
    x += card("FIRSTPIX", "0", "First pixel # (0 based)");
    x += card("LASTPIX", ival(npix - 1), "Last pixel # (0 based)");
    x += card("INDXSCHM", sval("IMPLICIT"), "Indexing: IMPLICIT or EXPLICIT");
    x += std::string("END").append(77, ' ');
    pad(x);
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) {
        if (err) *err = "cannot create " + path;
        return false;
    }
    bool ok = fwrite(h.data(), 1, h.size(), f) == h.size() && fwrite(x.data(), 1, x.size(), f) == x.size();
    const size_t chunk = 1 << 16;
    std::vector<unsigned char> buf(chunk * 8);
    for (long long i = 0; ok && i < npix; i += chunk) {
        const size_t m = (size_t)((npix - i) < (long long)chunk ? npix - i : chunk);
        for (size_t k = 0; k < m; k++) put_be64(&buf[8 * k], map[i + k]);
        ok = fwrite(buf.data(), 8, m, f) == m;
    }
    const size_t data = (size_t)npix * 8;
    if (ok && data % kBlock) {
        std::vector<unsigned char> z(kBlock - data % kBlock, 0);
        ok = fwrite(z.data(), 1, z.size(), f) == z.size();
    }
    ok = (fclose(f) == 0) && ok;
    if (!ok && err) *err = "write failed for " + path;
    return ok;
}

static bool read_header(FILE *f, std::vector<std::string> &cards)
{
    char blk[kBlock];
    for (;;) {
        if (fread(blk, 1, kBlock, f) != kBlock) return false;
        for (size_t i = 0; i < kBlock; i += 80) {
            std::string c(blk + i, 80);
            cards.push_back(c);
            if (c.compare(0, 8, "END     ") == 0) return true;
        }
    }
}

static bool key_value(const std::vector<std::string> &cards, const char *key, std::string &val)
{
    for (const auto &c : cards) {
        std::string k = c.substr(0, 8);
        while (!k.empty() && k.back() == ' ') k.pop_back();
        if (k == key && c.size() > 10 && c[8] == '=') {
            std::string v = c.substr(10);
            size_t slash = v.find('/');
            if (v.find('\'') != std::string::npos) {
                size_t a = v.find('\''), b = v.find('\'', a + 1);
                val = v.substr(a + 1, b - a - 1);
                while (!val.empty() && val.back() == ' ') val.pop_back();
            } else {
                val = v.substr(0, slash);
                size_t s = val.find_first_not_of(' '), e = val.find_last_not_of(' ');
                val = s == std::string::npos ? "" : val.substr(s, e - s + 1);
            }
            return true;
        }
    }
    return false;
}

bool read_healpix_fits(const std::string &path, std::vector<double> &map, int *nside, std::string *err)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) {
        if (err) *err = "cannot open " + path;
        return false;
    }
    std::vector<std::string> prim, ext;
    bool ok = read_header(f, prim) && read_header(f, ext);
    std::string v;
    long long naxis1 = 0, naxis2 = 0;
    int ns = 0;
    if (ok && key_value(ext, "NAXIS1", v)) naxis1 = atoll(v.c_str());
    if (ok && key_value(ext, "NAXIS2", v)) naxis2 = atoll(v.c_str());
    if (ok && key_value(ext, "NSIDE", v)) ns = atoi(v.c_str());
    ok = ok && naxis1 > 0 && naxis2 > 0 && naxis1 % 8 == 0;
    if (ok) {
        const long long n = naxis1 / 8 * naxis2;
        std::vector<unsigned char> raw((size_t)n * 8);
        ok = fread(raw.data(), 8, (size_t)n, f) == (size_t)n;
        if (ok) {
            map.resize((size_t)n);
            for (long long i = 0; i < n; i++) map[(size_t)i] = get_be64(&raw[(size_t)i * 8]);
        }
    }
    fclose(f);
    if (nside) *nside = ns;
    if (!ok && err) *err = "malformed HEALPix FITS file " + path;
    return ok;
}

// ---------------------------------------------------------------- C ABI

#include "dmann.h"

extern "C" int dmann_write_healpix_fits(const char *path, const double *map, int64_t npix, int nside)
{
    if (!path || !map || npix <= 0 || nside < 1) return DMANN_ERR_ARG;
    return write_healpix_fits(path, map, (long long)npix, nside) ? DMANN_OK : DMANN_ERR_ARG;
}

extern "C" int dmann_read_healpix_fits(const char *path, double *map, int64_t cap, int *nside, int64_t *npix)
{
    if (!path) return DMANN_ERR_ARG;
    std::vector<double> m;
    int ns = 0;
    if (!read_healpix_fits(path, m, &ns)) return DMANN_ERR_ARG;
    if (nside) *nside = ns;
    if (npix) *npix = (int64_t)m.size();
    if (map) {
        if ((int64_t)m.size() > cap) return DMANN_ERR_CAPACITY;
        std::memcpy(map, m.data(), m.size() * sizeof(double));
    }
    return DMANN_OK;
}
