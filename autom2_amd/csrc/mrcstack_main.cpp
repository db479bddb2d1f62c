// mrcstack_main.cpp -- mrcstack-compatible front end of libtiltrec_hip.so:
// applies an IMOD-style .xf to every section of an MRC stack on one GPU
// (trec_apply_xf), producing the aligned stack TiltRec reads.
//
// Options follow src/mrctools/src/mrcstack/opts.h:20-95: -i/--input,
// -o/--output, -x/--xform, -m/--mode (2: Lanczos-4 over 8 x 8 pixels, any
// other value: bicubic over 4 x 4; AuTom2's workflow passes no -m,
// CommandChainManager.qml:85-98).  The output follows main.cpp:24-84: the
// input header with mode 2 (float) and amin / amean / amax taken from the
// middle input section after a 3-sigma cutoff (GetImgAvgSdvCutoffMinMax,
// src/mrctools/src/mrcimg/img_util.h:353-368), the extended header kept at
// its length, then the warped sections.
//
// Differences, all where the reference's behaviour is undefined or fails:
//   - the extended header area is written as zeros (the reference writes an
//     uninitialised buffer, mrc2img.cpp:219-231);
//   - a .xf with fewer lines than sections is an error (the reference keeps
//     the previous line's values, and reads uninitialised ones for the
//     first);
//   - MRC modes 6 (unsigned short) and 12 (half) are read (the reference
//     leaves such sections zero, mrc2img.cpp:46-126);
//   - failures exit non-zero.
#include <hip/hip_runtime.h>
#include <getopt.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "mrc.h"
#include "tiltrec.h"

using namespace trec_cli;

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void usage()
{
    printf("[-i INPUT FILENAME][-x TRANSFORM][-o OUTPUT FILENAME]([-m MODE])\n"
           "  MODE is the interpolation modes to transform images.\n"
           "  MODE can be 1 for bicubic interpolation over 4x4 pixel neighborhood, 2 for lanczos "
           "interpolation over 8x8 pixel neighborhood.\n");
}

int die(int code, const char *fmt, const char *a = "")
{
    fprintf(stderr, "mrcstack error: ");
    fprintf(stderr, fmt, a);
    fprintf(stderr, "\n");
    return code;
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

// GetImgAvgSdvCutoffMinMax with cutoff 3: mean and standard deviation in
// double (meanStdDev), min and max, then max lowered to mean + 3 sd and min
// raised to mean - 3 sd.
void section_stats(const float *s, size_t n, double &avg, double &mn, double &mx)
{
    double sum = 0, sq = 0;
    double lo = s[0], hi = s[0];
    for (size_t i = 0; i < n; ++i) {
        const double v = s[i];
        sum += v;
        sq += v * v;
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
    const double scale = 1. / n;
    const double mean = sum * scale;
    const double sd = std::sqrt(std::max(sq * scale - mean * mean, 0.));
    mx = hi > mean + 3 * sd ? mean + 3 * sd : hi;
    mn = lo > mean - 3 * sd ? lo : mean - 3 * sd;
    avg = mean;
}

}  // namespace

int main(int argc, char **argv)
{
    std::string input, output, xform;
    int mode = 1, device = 0;
    static struct option longopts[] = {{"help", no_argument, nullptr, 'h'},
                                       {"output", required_argument, nullptr, 'o'},
                                       {"xform", required_argument, nullptr, 'x'},
                                       {"input", required_argument, nullptr, 'i'},
                                       {"mode", required_argument, nullptr, 'm'},
                                       {"device", required_argument, nullptr, 'd'},
                                       {nullptr, 0, nullptr, 0}};
    if (argc == 1) {
        usage();
        return 2;
    }
    int ch;
    while ((ch = getopt_long(argc, argv, "hi:o:x:m:d:", longopts, nullptr)) != -1) {
        switch (ch) {
        case 'h': usage(); return 0;
        case 'i': input = optarg; break;
        case 'o': output = optarg; break;
        case 'x': xform = optarg; break;
        case 'm': {
            char *end = nullptr;
            mode = (int)strtol(optarg, &end, 10);
            if (!*optarg || *end) return die(2, "Invalid argument '%s'.", optarg);
            break;
        }
        case 'd': device = atoi(optarg); break;
        default: usage(); return 2;
        }
    }
    if (input.empty() || output.empty() || xform.empty()) {
        usage();
        return 2;
    }
    printf("MRC transformation.\n");
    const double t0 = now_s();
    MrcHeader h;
    std::vector<float> stack;
    std::string err;
    if (!mrc_read_float(input, h, stack, err, BYTE_UNSIGNED_SCALED)) return die(4, "%s", err.c_str());
    const int W = h.nx, H = h.ny, P = h.nz;
    std::vector<float> xf((size_t)6 * P);
    {
        std::ifstream in(xform);
        if (!in.good()) return die(4, "cannot read transform file %s", xform.c_str());
        for (int i = 0; i < 6 * P; ++i)
            if (!(in >> xf[i])) return die(4, "transform file %s has fewer lines than the stack has sections", xform.c_str());
    }
    printf("Transformation parameters: %s\n", xform.c_str());
    const size_t sec = (size_t)W * H;
    double avg, mn, mx;
    section_stats(stack.data() + sec * (P / 2), sec, avg, mn, mx);
    const double t_read = now_s();

    if (hipSetDevice(device) != hipSuccess) return die(5, "HIP device %s is not available", std::to_string(device).c_str());
    DevBuf din, dout;
    if (hipMalloc(&din.p, 4 * sec * P) != hipSuccess || hipMalloc(&dout.p, 4 * sec * P) != hipSuccess)
        return die(5, "device allocation failed");
    if (hipMemcpy(din.p, stack.data(), 4 * sec * P, hipMemcpyHostToDevice) != hipSuccess)
        return die(5, "copy to device failed");
    trec_context *ctx = nullptr;
    trec_status s = trec_context_create(device, &ctx);
    if (s != TREC_OK) return die(5, "context: %s", trec_status_string(s));
    const double t_up = now_s();
    s = trec_apply_xf(ctx, W, H, P, xf.data(), mode, (const float *)din.p, (float *)dout.p, nullptr);
    if (s == TREC_OK && hipDeviceSynchronize() != hipSuccess) s = TREC_ERR_DEVICE;
    if (s != TREC_OK) {
        std::string msg = std::string(trec_status_string(s)) + ": " + trec_last_error(ctx);
        trec_context_destroy(ctx);
        return die(6, "transformation failed: %s", msg.c_str());
    }
    trec_context_destroy(ctx);
    const double t_xf = now_s();
    if (hipMemcpy(stack.data(), dout.p, 4 * sec * P, hipMemcpyDeviceToHost) != hipSuccess)
        return die(5, "copy from device failed");
    h.mode = MRC_FLOAT;
    h.amin = (float)mn;
    h.amean = (float)avg;
    h.amax = (float)mx;
    if (!mrc_write_with_header(output, h, stack.data(), err)) return die(4, "%s", err.c_str());
    const double t_end = now_s();
    printf("%d sections of %d x %d, %s: read %.3f s, upload %.3f s, transform %.3f s, write %.3f s\n", P, W, H,
           mode == 2 ? "Lanczos-4" : "bicubic", t_read - t0, t_up - t_read, t_xf - t_up, t_end - t_xf);
    return 0;
}
