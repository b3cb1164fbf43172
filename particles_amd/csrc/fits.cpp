// Minimal FITS primary-image reader and writer for the host driver: the
// field files the reference reads (sources/fitsio.F90:173-269, through
// cfitsio) and the counter / time / state files it writes
// (sources/fitsio.F90:281-540).  Format per the FITS standard 4.0: 2880-byte
// blocks of 80-character header cards, big-endian data.  Files may be
// gzip-compressed (read through zlib, which also reads plain files).
//
// Differences from the reference's I/O, on purpose: a BITPIX -64 field is
// read as FP64 (the reference reads every field through the single-precision
// call, fitsio.F90:255-260); the 1-D writer writes its data once
// (fitsio.F90:403,410 write it twice).
#include "particles_amd_fits.h"

#include <zlib.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_fits_err;

int ferr(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_fits_err = buf;
    return 1;
}

struct Header {
    int bitpix = 0, naxis = 0;
    int64_t dims[8] = {0};
    double bscale = 1.0, bzero = 0.0;
    int64_t header_bytes = 0;
};

std::string card_key(const char *card)
{
    std::string k(card, 8);
    while (!k.empty() && k.back() == ' ') k.pop_back();
    return k;
}

bool card_value(const char *card, std::string &v)
{
    if (card[8] != '=' || card[9] != ' ') return false;
    std::string s(card + 10, 70);
    const size_t slash = s.find('/');
    if (slash != std::string::npos && s.find('\'') == std::string::npos) s = s.substr(0, slash);
    size_t a = s.find_first_not_of(' '), b = s.find_last_not_of(' ');
    if (a == std::string::npos) return false;
    v = s.substr(a, b - a + 1);
    return true;
}

int read_header(gzFile f, const char *path, Header &h)
{
    char card[80];
    bool end = false, simple = false;
    int64_t ncards = 0;
    while (!end) {
        const int got = gzread(f, card, 80);
        if (got != 80) return ferr("%s: truncated FITS header", path);
        ncards++;
        const std::string key = card_key(card);
        std::string v;
        if (key == "END") {
            end = true;
        } else if (card_value(card, v)) {
            if (key == "SIMPLE") simple = v == "T";
            else if (key == "BITPIX") h.bitpix = atoi(v.c_str());
            else if (key == "NAXIS") h.naxis = atoi(v.c_str());
            else if (key.rfind("NAXIS", 0) == 0 && key.size() > 5) {
                const int i = atoi(key.c_str() + 5);
                if (i >= 1 && i <= 8) h.dims[i - 1] = strtoll(v.c_str(), nullptr, 10);
            } else if (key == "BSCALE") h.bscale = strtod(v.c_str(), nullptr);
            else if (key == "BZERO") h.bzero = strtod(v.c_str(), nullptr);
        }
    }
    if (!simple) return ferr("%s: not a FITS primary image (SIMPLE = T missing)", path);
    if (h.naxis < 0 || h.naxis > 8) return ferr("%s: NAXIS = %d not supported", path, h.naxis);
    switch (h.bitpix) {
    case 8: case 16: case 32: case 64: case -32: case -64: break;
    default: return ferr("%s: BITPIX = %d not supported", path, h.bitpix);
    }
    // skip the rest of the last header block
    const int64_t used = ncards * 80;
    const int64_t pad = (2880 - used % 2880) % 2880;
    std::vector<char> skip((size_t)pad);
    if (pad && gzread(f, skip.data(), (unsigned)pad) != (int)pad)
        return ferr("%s: truncated FITS header block", path);
    h.header_bytes = used + pad;
    return 0;
}

void swap_bytes(unsigned char *p, int64_t n, int w)
{
    for (int64_t i = 0; i < n; i++, p += w)
        for (int a = 0, b = w - 1; a < b; a++, b--) {
            const unsigned char t = p[a];
            p[a] = p[b];
            p[b] = t;
        }
}

bool big_endian_host()
{
    const uint16_t one = 1;
    return *(const unsigned char *)&one == 0;
}

// fixed-format header card "KEY     = value / comment"
void put_card(std::string &hdr, const char *key, const std::string &value, const char *comment)
{
    char card[81];
    if (comment && *comment)
        snprintf(card, sizeof card, "%-8.8s= %20s / %-47.47s", key, value.c_str(), comment);
    else
        snprintf(card, sizeof card, "%-8.8s= %20s%50s", key, value.c_str(), "");
    hdr.append(card, 80);
}

int write_image(const char *path, int bitpix, int naxis, const int64_t *dims, const void *data,
                int nkeys, const char *const *keys, const double *values,
                const char *const *comments)
{
    if (!path || !dims || naxis < 1 || naxis > 8) return ferr("fits write: bad arguments");
    if (*path == '!') path++;  // the cfitsio "overwrite" prefix; files are always replaced
    int64_t n = 1;
    for (int i = 0; i < naxis; i++) {
        if (dims[i] < 0) return ferr("%s: negative dimension", path);
        n *= dims[i];
    }
    if (n > 0 && !data) return ferr("%s: null data", path);
    std::string hdr;
    put_card(hdr, "SIMPLE", "T", "conforms to FITS standard");
    put_card(hdr, "BITPIX", std::to_string(bitpix), "array data type");
    put_card(hdr, "NAXIS", std::to_string(naxis), "number of array dimensions");
    for (int i = 0; i < naxis; i++) {
        const std::string k = "NAXIS" + std::to_string(i + 1);
        put_card(hdr, k.c_str(), std::to_string(dims[i]), nullptr);
    }
    put_card(hdr, "EXTEND", "T", nullptr);
    for (int i = 0; i < nkeys; i++) {
        char v[32];
        snprintf(v, sizeof v, "%.8E", values[i]);  // 9 significant digits
        put_card(hdr, keys[i], v, comments ? comments[i] : nullptr);
    }
    char end[81];
    snprintf(end, sizeof end, "%-80s", "END");
    hdr.append(end, 80);
    hdr.append((2880 - hdr.size() % 2880) % 2880, ' ');

    FILE *f = fopen(path, "wb");
    if (!f) return ferr("%s: cannot open for writing", path);
    bool ok = fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size();
    const int w = bitpix < 0 ? -bitpix / 8 : bitpix / 8;
    const size_t chunk = 1 << 20;
    std::vector<unsigned char> buf;
    for (int64_t i = 0; ok && i < n; i += (int64_t)chunk) {
        const int64_t m = n - i < (int64_t)chunk ? n - i : (int64_t)chunk;
        buf.assign((const unsigned char *)data + i * w, (const unsigned char *)data + (i + m) * w);
        if (!big_endian_host()) swap_bytes(buf.data(), m, w);
        ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    }
    const int64_t bytes = n * w;
    const int64_t pad = (2880 - bytes % 2880) % 2880;
    if (ok && pad) {
        std::vector<char> z((size_t)pad, 0);
        ok = fwrite(z.data(), 1, z.size(), f) == z.size();
    }
    ok = (fclose(f) == 0) && ok;
    if (!ok) return ferr("%s: write failed", path);
    return 0;
}

}  // namespace

extern "C" {

const char *ptc_fits_last_error(void) { return g_fits_err.c_str(); }

int ptc_fits_dims(const char *path, int *naxis, int64_t dims[8], int *bitpix)
{
    if (!path || !naxis || !dims) return ferr("ptc_fits_dims: null argument");
    gzFile f = gzopen(path, "rb");
    if (!f) return ferr("%s: cannot open", path);
    Header h;
    const int rc = read_header(f, path, h);
    gzclose(f);
    if (rc) return rc;
    *naxis = h.naxis;
    for (int i = 0; i < 8; i++) dims[i] = h.dims[i];
    if (bitpix) *bitpix = h.bitpix;
    g_fits_err.clear();
    return 0;
}

int ptc_fits_read(const char *path, int64_t n, double *out)
{
    if (!path || (n > 0 && !out)) return ferr("ptc_fits_read: null argument");
    gzFile f = gzopen(path, "rb");
    if (!f) return ferr("%s: cannot open", path);
    Header h;
    int rc = read_header(f, path, h);
    if (rc) { gzclose(f); return rc; }
    int64_t total = h.naxis > 0 ? 1 : 0;
    for (int i = 0; i < h.naxis; i++) total *= h.dims[i];
    if (total != n) {
        gzclose(f);
        return ferr("%s: holds %lld values, %lld requested", path, (long long)total, (long long)n);
    }
    const int w = h.bitpix < 0 ? -h.bitpix / 8 : h.bitpix / 8;
    const int64_t chunk = 1 << 20;
    std::vector<unsigned char> buf((size_t)(chunk * w));
    for (int64_t i = 0; i < n; i += chunk) {
        const int64_t m = n - i < chunk ? n - i : chunk;
        const int64_t bytes = m * w;
        if (gzread(f, buf.data(), (unsigned)bytes) != (int)bytes) {
            gzclose(f);
            return ferr("%s: truncated data", path);
        }
        if (!big_endian_host()) swap_bytes(buf.data(), m, w);
        for (int64_t j = 0; j < m; j++) {
            const unsigned char *p = buf.data() + j * w;
            double v = 0.0;
            switch (h.bitpix) {
            case 8: v = *(const uint8_t *)p; break;
            case 16: v = *(const int16_t *)p; break;
            case 32: v = *(const int32_t *)p; break;
            case 64: v = (double)*(const int64_t *)p; break;
            case -32: v = *(const float *)p; break;
            default: v = *(const double *)p; break;
            }
            out[i + j] = h.bzero + h.bscale * v;
        }
    }
    gzclose(f);
    g_fits_err.clear();
    return 0;
}

int ptc_fits_write_f64(const char *path, int naxis, const int64_t *dims, const double *data,
                       int nkeys, const char *const *keys, const double *values,
                       const char *const *comments)
{
    if (nkeys < 0 || (nkeys > 0 && (!keys || !values))) return ferr("fits write: bad keys");
    const int rc = write_image(path, -64, naxis, dims, data, nkeys, keys, values, comments);
    if (!rc) g_fits_err.clear();
    return rc;
}

int ptc_fits_write_i64(const char *path, int naxis, const int64_t *dims, const int64_t *data)
{
    const int rc = write_image(path, 64, naxis, dims, data, 0, nullptr, nullptr, nullptr);
    if (!rc) g_fits_err.clear();
    return rc;
}

}  // extern "C"
