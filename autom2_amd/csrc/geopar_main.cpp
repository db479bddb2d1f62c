// geopar_main.cpp -- GeoParGen-compatible front end of libtiltrec_hip.so:
// estimates the tilt, thickness and z centre of the specimen slab in a
// trial reconstruction on one GPU (trec_geopar).
//
// Usage follows src/GeoParGen/src/main/main.cpp:96-135: GeoParGen VOLUME
// [x|y]; "x" (the default) takes the views at every column, "y" every
// section.  The last three lines are the ones AuTom2's GUI parses
// (CommandChainManager.qml:829-853): "average angle : A", "average thick :
// T", "average z : Z", printed as std::cout prints a float.
//
// Under mpirun -n N the reference splits the views over ranks and gathers
// them on rank 0 (main.cpp:137-216); the results do not depend on N.  Here
// rank 0 does every view on one GPU and the other ranks exit at once.
//
// Differences, all where the reference fails or reads undefined values:
//   - failures exit non-zero (the reference returns 0 after printing);
//   - without arguments it prints its usage (the reference opens an empty
//     path: its default path is assigned to a shadowing local);
//   - MRC modes 6 and 12 are read (the reference leaves them zero);
//   - a volume holding NaN or infinity is an error.
//   - --device K picks the GPU (an addition).
#include <hip/hip_runtime.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mrc.h"
#include "tiltrec.h"

using namespace trec_cli;

namespace {

int launcher_rank()
{
    const char *vars[] = {"OMPI_COMM_WORLD_RANK", "PMI_RANK", "PMIX_RANK", "MV2_COMM_WORLD_RANK"};
    for (const char *v : vars)
        if (const char *s = getenv(v)) return atoi(s);
    return 0;
}

int launcher_size()
{
    const char *vars[] = {"OMPI_COMM_WORLD_SIZE", "PMI_SIZE", "PMIX_SIZE", "MV2_COMM_WORLD_SIZE"};
    for (const char *v : vars)
        if (const char *s = getenv(v)) return atoi(s);
    return 1;
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

int main(int argc, char **argv)
{
    const int rank = launcher_rank();
    if (rank > 0) {
        printf("processor %d: the views are computed on rank 0\n", rank);
        return 0;
    }
    std::vector<std::string> pos;
    int device = 0;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--device") && i + 1 < argc) device = atoi(argv[++i]);
        else pos.push_back(argv[i]);
    }
    char host[256] = "localhost";
    gethostname(host, sizeof host - 1);
    printf("MPI Init success! from processor %s, rank %d out of %d processors\n", host, rank, launcher_size());
    if (pos.empty() || pos.size() > 2) {
        printf("invalid parameter! \nusage: GeoParGen VOLUME.mrc [x|y] [--device K]\n");
        return 2;
    }
    const std::string path = pos[0], what = pos.size() == 2 ? pos[1] : "x";
    MrcHeader h;
    std::vector<float> vol;
    std::string err;
    if (!mrc_read_float(path, h, vol, err, BYTE_UNSIGNED_SCALED)) {
        printf("file read failed! %s\n", path.c_str());
        fprintf(stderr, "GeoParGen error: %s\n", err.c_str());
        return 4;
    }
    printf("file read success\n");
    printf("begin generate %s view ... \n", what.c_str());
    if (what != "x" && what != "y") {
        printf("invalid parameter! \n");
        return 2;
    }
    printf("generate %s View ...\n", what == "x" ? "X" : "Y");
    const int axis = what == "x" ? TREC_GP_VIEW_X : TREC_GP_VIEW_Y;

    if (hipSetDevice(device) != hipSuccess) {
        fprintf(stderr, "GeoParGen error: HIP device %d is not available\n", device);
        return 5;
    }
    const size_t bytes = sizeof(float) * vol.size();
    DevBuf d;
    if (hipMalloc(&d.p, bytes) != hipSuccess || hipMemcpy(d.p, vol.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
        fprintf(stderr, "GeoParGen error: device allocation or copy failed\n");
        return 5;
    }
    trec_context *ctx = nullptr;
    trec_status s = trec_context_create(device, &ctx);
    if (s != TREC_OK) {
        fprintf(stderr, "GeoParGen error: context: %s\n", trec_status_string(s));
        return 5;
    }
    float avg[3];
    s = trec_geopar(ctx, (const float *)d.p, h.nx, h.ny, h.nz, axis, 0, nullptr, nullptr, avg, nullptr, nullptr);
    if (s != TREC_OK) {
        fprintf(stderr, "GeoParGen error: %s: %s\n", trec_status_string(s), trec_last_error(ctx));
        trec_context_destroy(ctx);
        return 6;
    }
    trec_context_destroy(ctx);
    printf("average angle : %g\n", avg[0]);
    printf("average thick : %g\n", avg[1]);
    printf("average z : %g\n", avg[2]);
    return 0;
}
