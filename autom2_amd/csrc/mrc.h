// mrc.h -- MRC file I/O for the TiltRec CLI (host only).
//
// The 1024-byte header has the field layout of TiltRec's MRCheader
// (src/TiltRec/src/mrcmx/mrcheader.h:24-101, packed), which is the MRC
// format's own layout.  Reading follows MrcStackM::ReadBlock
// (src/TiltRec/src/mrcmx/mrcstack.cpp:354-459): data start at 1024 + next,
// mode 0 is read as signed char (savedate, :255-287), 1 as short, 2 as
// float.  Beyond the reference, the readers take the rest of MRC2014: modes
// 6 (unsigned short), 12 (half) and 101 (4-bit), and big-endian files (by
// the machine stamp, or by whichever byte order gives a plausible header
// when there is none).  Writing follows InitializeHeader / WriteToFile /
// UpdateHeader (:14-142); on request it writes mode 12 and an MRC2014
// header instead.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace trec_cli {

#pragma pack(push, 1)
struct MrcHeader {
    int32_t nx, ny, nz, mode;
    int32_t nxstart, nystart, nzstart;
    int32_t mx, my, mz;
    float xlen, ylen, zlen;
    float alpha, beta, gamma;
    int32_t mapc, mapr, maps;
    float amin, amax, amean;
    int16_t ispg, nsymbt;
    int32_t next;
    int16_t creatid;
    char blank[30];
    int16_t nint, nreal, sub, zfac;
    float min2, max2, min3, max3, min4, max4;
    int16_t idtype, lens, nd1, nd2, vd1, vd2;
    float tiltangles[6];
    float xorg, yorg, zorg;
    char cmap[4];
    char stamp[4];
    float rms;
    int32_t nlabl;
    char labels[10][80];
};
#pragma pack(pop)
static_assert(sizeof(MrcHeader) == 1024, "MRC header is 1024 bytes");

enum { MRC_BYTE = 0, MRC_SHORT = 1, MRC_FLOAT = 2, MRC_USHORT = 6, MRC_HALF = 12, MRC_NIBBLE = 101 };

// Reads the header in native byte order; returns false with a message on
// failure.  *swapped (may be NULL) tells whether the file is big-endian.
bool mrc_read_header(const std::string &path, MrcHeader &h, std::string &err, bool *swapped = nullptr);

// Reads the data as stored, in native byte order: nx * ny * nz elements of
// the mode's type (mode 101 unpacked to one unsigned byte per pixel).
bool mrc_read_raw(const std::string &path, MrcHeader &h, std::vector<unsigned char> &raw, std::string &err);

// How mode 0 (bytes) becomes float: TiltRec's reader takes signed char
// (mrcstack.cpp:255-287); mrctools' MrcStack takes unsigned char / 255
// (src/mrctools/src/mrcimg/mrc2img.cpp:60-75).
enum ByteRule { BYTE_SIGNED = 0, BYTE_UNSIGNED_SCALED = 1 };

// Reads all sections as float, [nz][ny][nx].
bool mrc_read_float(const std::string &path, MrcHeader &h, std::vector<float> &data, std::string &err,
                    ByteRule bytes = BYTE_SIGNED);

// --f2b output: mode 0 (InitializeHeader(0), mrcstack.cpp:57-86) and bytes
// [nz][ny][nx]; amin/amax/amean from section nz/2 read as bytes (the
// reference's UpdateHeader reads that section as floats, which for a byte
// file takes the wrong bytes).
bool mrc_write_bytes(const std::string &path, MrcHeader h, const unsigned char *data, std::string &err);

// Writes header h as given, h.next zero bytes of extended header, and
// h.nx * h.ny * h.nz floats.
bool mrc_write_with_header(const std::string &path, const MrcHeader &h, const float *data, std::string &err);

// InitializeHeader (mrcstack.cpp:57-86) with the given size, mode 2.
MrcHeader mrc_new_header(int nx, int ny, int nz);

// Writes header and float data ([nz][ny][nx]); then sets amin/amax/amean
// from section nz/2 as UpdateHeader does (mrcstack.cpp:104-142: amax
// starts at FLT_MIN, the mean is accumulated in long double) and nlabl = 0.
bool mrc_write_float(const std::string &path, MrcHeader h, const float *data, std::string &err);

// MRC2014 header fields: the "MAP " stamp, a little-endian machine stamp,
// NVERSION 20140, sampling mx/my/mz = nx/ny/nz and cell lengths of `pixel`
// (Angstrom, 1 when unknown) per voxel.
void mrc_set_2014(MrcHeader &h, float pixel);

// Writes mode 2 (float) or 12 (half) data.  Statistics: with mrc2014, min,
// max, mean and RMS deviation of the whole map and one label; otherwise
// UpdateHeader's, as mrc_write_float.
bool mrc_write_volume(const std::string &path, MrcHeader h, const void *data, int mode, bool mrc2014,
                      std::string &err);

// Element offset of section nz / 2, whose statistics UpdateHeader puts in
// the header (mrcstack.cpp:104-142), and that update from the section alone.
size_t mrc_mid_section(const MrcHeader &h);
void update_header_mid(MrcHeader &h, const void *section, int mode);

// The same file as mrc_write_volume without mrc2014, written in pieces: the
// header from section nz / 2 (which the caller fetches first), then the
// elements in order, so a volume can go out while the rest of it is still
// being copied from the device.
class MrcStreamWriter {
  public:
    bool open(const std::string &path, MrcHeader h, int mode, const void *mid_section, std::string &err);
    bool write(const void *data, size_t bytes, std::string &err);
    bool close(std::string &err);
    ~MrcStreamWriter();

  private:
    FILE *f_ = nullptr;
    std::string path_;
};

}  // namespace trec_cli
