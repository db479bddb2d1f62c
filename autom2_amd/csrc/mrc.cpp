// mrc.cpp -- MRC file I/O for the TiltRec CLI.  Whole-file contiguous
// reads and writes replace the reference's row-at-a-time MPI-IO calls.
#include "mrc.h"

#include <cerrno>
#include <cfloat>
#include <cstdio>
#include <cstring>

namespace trec_cli {

namespace {

struct File {
    FILE *f = nullptr;
    ~File()
    {
        if (f) fclose(f);
    }
};

size_t mode_bytes(int mode)
{
    switch (mode) {
    case MRC_BYTE: return 1;
    case MRC_SHORT:
    case MRC_USHORT:
    case MRC_HALF: return 2;
    case MRC_FLOAT: return 4;
    default: return 0;
    }
}

float half_to_float(uint16_t h)
{
    const uint32_t s = (uint32_t)(h >> 15) << 31;
    uint32_t e = (h >> 10) & 0x1f, m = h & 0x3ff, bits;
    if (e == 0) {
        if (m == 0) {
            bits = s;
        } else {  // subnormal: renormalise
            e = 127 - 15 + 1;
            while (!(m & 0x400)) {
                m <<= 1;
                --e;
            }
            bits = s | (e << 23) | ((m & 0x3ff) << 13);
        }
    } else if (e == 31) {
        bits = s | 0x7f800000u | (m << 13);
    } else {
        bits = s | ((e + 127 - 15) << 23) | (m << 13);
    }
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

}  // namespace

bool mrc_read_header(const std::string &path, MrcHeader &h, std::string &err)
{
    File fh;
    fh.f = fopen(path.c_str(), "rb");
    if (!fh.f) {
        err = "cannot open " + path + ": " + strerror(errno);
        return false;
    }
    if (fread(&h, sizeof h, 1, fh.f) != 1) {
        err = "short MRC header in " + path;
        return false;
    }
    if (h.nx <= 0 || h.ny <= 0 || h.nz <= 0) {
        err = "bad MRC size in " + path;
        return false;
    }
    if (!mode_bytes(h.mode)) {
        err = "File type unknown! (MRC mode " + std::to_string(h.mode) + ")";
        return false;
    }
    if (h.next < 0) {
        err = "bad extended header size in " + path;
        return false;
    }
    return true;
}

bool mrc_read_float(const std::string &path, MrcHeader &h, std::vector<float> &data, std::string &err, ByteRule bytes)
{
    if (!mrc_read_header(path, h, err)) return false;
    File fh;
    fh.f = fopen(path.c_str(), "rb");
    if (!fh.f) {
        err = "cannot open " + path;
        return false;
    }
    const size_t n = (size_t)h.nx * h.ny * h.nz;
    const size_t es = mode_bytes(h.mode);
    if (fseeko(fh.f, (off_t)sizeof(MrcHeader) + h.next, SEEK_SET) != 0) {
        err = "seek failed in " + path;
        return false;
    }
    std::vector<unsigned char> raw(n * es);
    if (fread(raw.data(), es, n, fh.f) != n) {
        err = "truncated MRC data in " + path;
        return false;
    }
    data.resize(n);
    switch (h.mode) {
    case MRC_BYTE:
        if (bytes == BYTE_UNSIGNED_SCALED)
            for (size_t i = 0; i < n; ++i) data[i] = raw[i] / 255.0f;
        else
            for (size_t i = 0; i < n; ++i) data[i] = (float)(signed char)raw[i];
        break;
    case MRC_SHORT:
        for (size_t i = 0; i < n; ++i) {
            int16_t v;
            memcpy(&v, &raw[2 * i], 2);
            data[i] = v;
        }
        break;
    case MRC_USHORT:
        for (size_t i = 0; i < n; ++i) {
            uint16_t v;
            memcpy(&v, &raw[2 * i], 2);
            data[i] = v;
        }
        break;
    case MRC_HALF:
        for (size_t i = 0; i < n; ++i) {
            uint16_t v;
            memcpy(&v, &raw[2 * i], 2);
            data[i] = half_to_float(v);
        }
        break;
    default:
        memcpy(data.data(), raw.data(), n * 4);
        break;
    }
    return true;
}

MrcHeader mrc_new_header(int nx, int ny, int nz)
{
    MrcHeader h;
    memset(&h, 0, sizeof h);
    h.mode = MRC_FLOAT;
    h.mx = h.my = h.mz = 1;
    h.xlen = h.ylen = h.zlen = 1;
    h.alpha = h.beta = h.gamma = 90;
    h.mapc = 1;
    h.mapr = 2;
    h.maps = 3;
    h.amin = 0;
    h.amax = 255;
    h.amean = 128;
    h.creatid = 1000;
    h.nx = nx;
    h.ny = ny;
    h.nz = nz;
    return h;
}

bool mrc_write_float(const std::string &path, MrcHeader h, const float *data, std::string &err)
{
    const size_t sec = (size_t)h.nx * h.ny;
    const float *mid = data + sec * (size_t)(int)(h.nz * .5);
    float amin = FLT_MAX, amax = FLT_MIN;
    long double amean = 0;
    for (size_t i = sec; i--;) {
        if (mid[i] > amax) amax = mid[i];
        if (mid[i] < amin) amin = mid[i];
        amean += mid[i];
    }
    amean /= sec;
    h.amin = amin;
    h.amax = amax;
    h.amean = (float)amean;
    h.nlabl = 0;
    h.mode = MRC_FLOAT;
    h.next = 0;
    File fh;
    fh.f = fopen(path.c_str(), "wb");
    if (!fh.f) {
        err = "cannot create " + path + ": " + strerror(errno);
        return false;
    }
    const size_t n = sec * h.nz;
    if (fwrite(&h, sizeof h, 1, fh.f) != 1 || fwrite(data, 4, n, fh.f) != n) {
        err = "write failed for " + path;
        return false;
    }
    if (fclose(fh.f) != 0) {
        fh.f = nullptr;
        err = "close failed for " + path;
        return false;
    }
    fh.f = nullptr;
    return true;
}

bool mrc_write_bytes(const std::string &path, MrcHeader h, const unsigned char *data, std::string &err)
{
    const size_t sec = (size_t)h.nx * h.ny;
    const unsigned char *mid = data + sec * (size_t)(int)(h.nz * .5);
    float amin = FLT_MAX, amax = FLT_MIN;
    long double amean = 0;
    for (size_t i = sec; i--;) {
        const float v = mid[i];
        if (v > amax) amax = v;
        if (v < amin) amin = v;
        amean += v;
    }
    amean /= sec;
    h.amin = amin;
    h.amax = amax;
    h.amean = (float)amean;
    h.nlabl = 0;
    h.mode = MRC_BYTE;
    h.next = 0;
    File fh;
    fh.f = fopen(path.c_str(), "wb");
    if (!fh.f) {
        err = "cannot create " + path + ": " + strerror(errno);
        return false;
    }
    const size_t n = sec * h.nz;
    if (fwrite(&h, sizeof h, 1, fh.f) != 1 || fwrite(data, 1, n, fh.f) != n) {
        err = "write failed for " + path;
        return false;
    }
    if (fclose(fh.f) != 0) {
        fh.f = nullptr;
        err = "close failed for " + path;
        return false;
    }
    fh.f = nullptr;
    return true;
}

bool mrc_write_with_header(const std::string &path, const MrcHeader &h, const float *data, std::string &err)
{
    File fh;
    fh.f = fopen(path.c_str(), "wb");
    if (!fh.f) {
        err = "cannot create " + path + ": " + strerror(errno);
        return false;
    }
    const size_t n = (size_t)h.nx * h.ny * h.nz;
    std::vector<char> ext(h.next > 0 ? h.next : 0, 0);
    if (fwrite(&h, sizeof h, 1, fh.f) != 1 || fwrite(ext.data(), 1, ext.size(), fh.f) != ext.size() ||
        fwrite(data, 4, n, fh.f) != n) {
        err = "write failed for " + path;
        return false;
    }
    if (fclose(fh.f) != 0) {
        fh.f = nullptr;
        err = "close failed for " + path;
        return false;
    }
    fh.f = nullptr;
    return true;
}

}  // namespace trec_cli
