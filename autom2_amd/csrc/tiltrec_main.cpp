// tiltrec_main.cpp -- TiltRec-compatible command-line front end of
// libtiltrec_hip.so.
//
// Installed (or symlinked) as TiltRec-cuda / TiltRecZ-cuda it takes the
// argv AuTom2's workflow passes (CommandChainManager.qml:161-206 composes
// "./TiltRec[Z]-cuda --input .. --output .. --tiltfile .. --geometry
// off,pitch,zshift,thick --method SIRT,iters,relax"), reads the MRC tilt
// stack and the tilt file, reconstructs on one GPU through the C ABI and
// writes the MRC volume in the binary's orientation: TiltRec-* (X, T, Y),
// TiltRecZ-* (X, Y, T) (main-cuda.cpp:266-268, main-cudaz.cpp:172).
//
// The reference runs under `mpirun -n N` and does all GPU work on rank 0
// (main-cuda.cpp:303-348); this program does the same: a process whose
// launcher rank is not 0 exits 0 at once.  Unlike the reference, every
// failure exits non-zero (the reference returns 0 after most failures,
// main-cuda.cpp:406-411), so the workflow sees it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "mrc.h"
#include "opts.h"
#include "tiltrec.h"

using namespace trec_cli;

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int launcher_rank()
{
    const char *vars[] = {"OMPI_COMM_WORLD_RANK", "PMI_RANK", "PMIX_RANK", "MV2_COMM_WORLD_RANK"};
    for (const char *v : vars)
        if (const char *s = getenv(v)) return atoi(s);
    return 0;
}

// ReadAngles (main-cuda.cpp:51-72): floats until the first parse failure.
bool read_angles(const std::string &path, std::vector<float> &angles)
{
    std::ifstream in(path);
    if (!in.good()) return false;
    float v;
    while (in >> v) angles.push_back(v);
    return true;
}

std::string base_name(const char *p)
{
    std::string s(p);
    size_t k = s.find_last_of('/');
    return k == std::string::npos ? s : s.substr(k + 1);
}

int die(int code, const char *fmt, const char *a = "", const char *b = "")
{
    fprintf(stderr, "TiltRec error: ");
    fprintf(stderr, fmt, a, b);
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

// Copies `bytes` of device memory into the file in 64 MiB pieces through two
// pinned buffers: piece k + 1 comes down while piece k is written.  false
// with err empty: no pinned memory, the caller copies the volume whole.
bool stream_out(MrcStreamWriter &w, const unsigned char *src, size_t bytes, std::string &err)
{
    constexpr size_t CH = (size_t)64 << 20;
    unsigned char *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipStream_t st = nullptr;
    bool ok = hipStreamCreate(&st) == hipSuccess;
    for (int b = 0; b < 2 && ok; ++b)
        ok = hipHostMalloc((void **)&buf[b], CH, hipHostMallocDefault) == hipSuccess &&
             hipEventCreateWithFlags(&ev[b], hipEventDisableTiming) == hipSuccess;
    auto issue = [&](size_t off, int b) {
        return hipMemcpyAsync(buf[b], src + off, std::min(CH, bytes - off), hipMemcpyDeviceToHost, st) == hipSuccess &&
               hipEventRecord(ev[b], st) == hipSuccess;
    };
    bool wrote = false;
    if (ok && bytes) ok = issue(0, 0);
    for (size_t off = 0; ok && off < bytes; off += CH) {
        const int b = (int)((off / CH) & 1);
        ok = hipEventSynchronize(ev[b]) == hipSuccess;
        if (ok && off + CH < bytes) ok = issue(off + CH, b ^ 1);
        if (ok) {
            wrote = true;
            if (!w.write(buf[b], std::min(CH, bytes - off), err)) return false;
        }
    }
    if (st) (void)hipStreamSynchronize(st);
    for (int b = 0; b < 2; ++b) {
        if (buf[b]) (void)hipHostFree(buf[b]);
        if (ev[b]) (void)hipEventDestroy(ev[b]);
    }
    if (st) (void)hipStreamDestroy(st);
    (void)hipGetLastError();
    if (!ok && wrote) err = "copy from device failed";
    return ok;
}

// memcpy over four threads (one host thread copies at about 4 GB/s)
void copy4(unsigned char *dst, const unsigned char *src, size_t n)
{
    constexpr int NT = 4;
    const size_t part = (n / NT + 4095) & ~(size_t)4095;
    std::thread t[NT - 1];
    int started = 0;
    for (int k = 1; k < NT && (size_t)k * part < n; ++k, ++started)
        t[k - 1] = std::thread([=] { std::memcpy(dst + k * part, src + k * part, std::min(part, n - k * part)); });
    std::memcpy(dst, src, std::min(part, n));
    for (int k = 0; k < started; ++k) t[k].join();
}

// Copies `bytes` of host memory to the device in 64 MiB pieces through two
// pinned buffers: piece k + 1 is copied into one while piece k goes up from
// the other.  Without pinned memory it copies the whole at once.
bool stream_in(void *dst, const unsigned char *src, size_t bytes)
{
    constexpr size_t CH = (size_t)64 << 20;
    unsigned char *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipStream_t st = nullptr;
    bool pinned = hipStreamCreate(&st) == hipSuccess;
    for (int b = 0; b < 2 && pinned; ++b)
        pinned = hipHostMalloc((void **)&buf[b], CH, hipHostMallocDefault) == hipSuccess &&
                 hipEventCreateWithFlags(&ev[b], hipEventDisableTiming) == hipSuccess;
    bool ok = true;
    if (pinned) {
        for (size_t off = 0; ok && off < bytes; off += CH) {
            const int b = (int)((off / CH) & 1);
            const size_t n = std::min(CH, bytes - off);
            // buffer b's previous piece (two pieces ago) has gone up
            if (off >= 2 * CH) ok = hipEventSynchronize(ev[b]) == hipSuccess;
            if (!ok) break;
            copy4(buf[b], src + off, n);
            ok = hipMemcpyAsync(static_cast<unsigned char *>(dst) + off, buf[b], n, hipMemcpyHostToDevice, st) ==
                     hipSuccess &&
                 hipEventRecord(ev[b], st) == hipSuccess;
        }
        if (st) ok = hipStreamSynchronize(st) == hipSuccess && ok;
    }
    for (int b = 0; b < 2; ++b) {
        if (buf[b]) (void)hipHostFree(buf[b]);
        if (ev[b]) (void)hipEventDestroy(ev[b]);
    }
    if (st) (void)hipStreamDestroy(st);
    (void)hipGetLastError();
    if (!pinned) return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    return ok;
}

}  // namespace

int main(int argc, char **argv)
{
    const int rank = launcher_rank();
    if (rank > 0) {
        printf("rank %d: the reconstruction runs on rank 0\n", rank);
        return 0;
    }
    Options o;
    const int r = get_opts(argc, argv, o);
    if (r == 0) return 0;
    if (r < 0) return 2;

    const std::string prog = base_name(argv[0]);
    const bool zbin = prog.rfind("TiltRecZ", 0) == 0;
    const int layout = (zbin || o.axis == "z" || o.axis == "Z") ? TREC_LAYOUT_Z : TREC_LAYOUT_Y;
    int rules = TREC_RULES_CUDA;
    if (o.rules == "mpi" || (o.rules.empty() && prog.find("-mpi") != std::string::npos)) rules = TREC_RULES_MPI;
    if (o.dry_run) {
        std::string j = opts_json(o);
        j.pop_back();
        printf("%s,\"layout\":\"%s\",\"resolved_rules\":\"%s\"}\n", j.c_str(), layout == TREC_LAYOUT_Z ? "z" : "y",
               rules == TREC_RULES_MPI ? "mpi" : "cuda");
        return 0;
    }
    int kind;
    if (o.method == "BPT") kind = TREC_BPT;
    else if (o.method == "FBP") kind = TREC_FBP;
    else if (o.method == "WBP") kind = TREC_WBP;
    else if (o.method == "SIRT") kind = TREC_SIRT;
    else if (o.method == "SART") kind = TREC_SART;
    else if (o.method == "ADMM") kind = TREC_ADMM;
    else return die(3, "method %s is not provided", o.method.c_str());

    const double t0 = now_s();
    MrcHeader hin;
    std::vector<unsigned char> stack;
    std::string err;
    // The device comes up (runtime, buffers, context) on a thread of its
    // own while the stack is read from the file.
    if (!mrc_read_header(o.input, hin, err, nullptr)) return die(4, "%s", err.c_str());
    const size_t nstack = (size_t)hin.nz * hin.ny * hin.nx, nvol = (size_t)hin.nx * hin.ny * o.thickness;
    DevBuf dstack, dvol, draw;
    trec_context *ctx = nullptr;
    struct CtxGuard {
        trec_context *&c;
        ~CtxGuard()
        {
            if (c) trec_context_destroy(c);
        }
    } ctx_guard{ctx};  // destroyed after the thread below has joined
    auto drop_ctx = [&] {
        if (ctx) trec_context_destroy(ctx);
        ctx = nullptr;
    };
    int dev_fail = 0;  // 1: no device, 2: allocation, 3: context
    std::thread dev_up([&] {
        if (hipSetDevice(o.device) != hipSuccess) {
            dev_fail = 1;
            return;
        }
        if (hipMalloc(&dstack.p, 4 * nstack) != hipSuccess || hipMalloc(&dvol.p, 4 * nvol) != hipSuccess) {
            dev_fail = 2;
            return;
        }
        if (trec_context_create(o.device, &ctx) != TREC_OK) dev_fail = 3;
    });
    struct Joiner {
        std::thread &t;
        ~Joiner()
        {
            if (t.joinable()) t.join();
        }
    } joiner{dev_up};
    if (!mrc_read_raw(o.input, hin, stack, err)) return die(4, "%s", err.c_str());
    std::vector<float> angles;
    if (!read_angles(o.angle, angles)) return die(4, "cannot read tilt file %s", o.angle.c_str());
    if ((int)angles.size() < hin.nz)
        return die(4, "tilt file %s has fewer angles than the stack has projections", o.angle.c_str());
    angles.resize(hin.nz);
    const int X = hin.nx, Y = hin.ny, P = hin.nz, T = o.thickness;
    printf("TiltRec-hip: %d projections of %d x %d -> %d x %d x %d (%s), method %s\n", P, X, Y, X, Y, T,
           layout == TREC_LAYOUT_Z ? "TiltRecZ orientation" : "TiltRec orientation", o.method.c_str());
    fflush(stdout);  // a caller watching the pipe sees the run start

    MrcHeader hi{};
    std::vector<unsigned char> init;
    if (!o.initial.empty()) {
        if (!mrc_read_raw(o.initial, hi, init, err)) return die(4, "%s", err.c_str());
        const bool ok = layout == TREC_LAYOUT_Y ? (hi.nx == X && hi.ny == T && hi.nz == Y)
                                                : (hi.nx == X && hi.ny == Y && hi.nz == T);
        if (!ok) return die(4, "initial volume %s does not match the output size", o.initial.c_str());
    }
    const double t_read = now_s();

    trec_geom g{X, Y, P, T, o.offset, o.pitch, o.zshift, rules};
    const int storage = o.storage == "f16" ? TREC_STORE_F16 : (o.storage == "bf16" ? TREC_STORE_BF16 : TREC_STORE_F32);
    trec_method m{kind, o.iteration, o.gamma, storage, layout, o.cgiter, o.soft};
    dev_up.join();
    if (dev_fail == 1 || hipSetDevice(o.device) != hipSuccess)
        return die(5, "HIP device %s is not available", std::to_string(o.device).c_str());
    if (dev_fail == 2) return die(5, "device allocation failed");
    if (dev_fail == 3) return die(5, "context: cannot be created");
    trec_status s = TREC_OK;
    // the file's own elements go up; the GPU widens them to float
    // (MRC mode 0 as signed bytes, mrcstack.cpp:255-287)
    // (through pinned pieces, stream_in)
    auto upload = [&](const std::vector<unsigned char> &raw, int mode, void *dst, size_t n) -> bool {
        if (mode == MRC_FLOAT) return stream_in(dst, raw.data(), 4 * n);
        if (!draw.p && hipMalloc(&draw.p, 2 * std::max(nstack, nvol)) != hipSuccess) return false;
        if (!stream_in(draw.p, raw.data(), raw.size())) return false;
        const int dt = mode == MRC_BYTE ? TREC_DT_I8
                     : mode == MRC_NIBBLE ? TREC_DT_U8
                     : mode == MRC_SHORT ? TREC_DT_I16
                     : mode == MRC_USHORT ? TREC_DT_U16 : TREC_DT_F16;
        return trec_convert(ctx, draw.p, dt, dst, TREC_DT_F32, n, nullptr) == TREC_OK &&
               hipDeviceSynchronize() == hipSuccess;
    };
    if (!upload(stack, hin.mode, dstack.p, nstack) || (!init.empty() && !upload(init, hi.mode, dvol.p, nvol))) {
        drop_ctx();
        return die(5, "copy to device failed");
    }
    stack.clear();
    stack.shrink_to_fit();
    const double t_up = now_s();
    s = trec_reconstruct(ctx, &g, angles.data(), (const float *)dstack.p, init.empty() ? nullptr : (const float *)dvol.p,
                         &m, (float *)dvol.p, nullptr);
    if (s == TREC_OK && hipDeviceSynchronize() != hipSuccess) s = TREC_ERR_DEVICE;
    if (s != TREC_OK) {
        std::string msg = std::string(trec_status_string(s)) + ": " + trec_last_error(ctx);
        drop_ctx();
        return die(6, "reconstruction failed: %s", msg.c_str());
    }
    const double t_rec = now_s();
    MrcHeader hout = layout == TREC_LAYOUT_Y ? mrc_new_header(X, T, Y) : mrc_new_header(X, Y, T);
    if (o.mrc2014) mrc_set_2014(hout, hin.mx > 0 && hin.xlen > 0 ? hin.xlen / hin.mx : 1.f);
    // The volume (float or half on the device) to the file: without MRC2014
    // statistics the header needs only section nz / 2, so the volume streams
    // out through pinned memory; otherwise, or without pinned memory, it is
    // copied whole first.
    auto write_out = [&](const void *dsrc, int mode) -> int {
        const size_t es = mode == MRC_HALF ? 2 : 4;
        const unsigned char *src = static_cast<const unsigned char *>(dsrc);
        if (!o.mrc2014) {
            std::vector<unsigned char> mid((size_t)hout.nx * hout.ny * es);
            if (hipMemcpy(mid.data(), src + es * mrc_mid_section(hout), mid.size(), hipMemcpyDeviceToHost) !=
                hipSuccess)
                return die(5, "copy from device failed");
            MrcStreamWriter w;
            if (!w.open(o.output, hout, mode, mid.data(), err)) return die(4, "%s", err.c_str());
            if (stream_out(w, src, es * nvol, err)) return w.close(err) ? 0 : die(4, "%s", err.c_str());
            if (!err.empty()) return die(err == "copy from device failed" ? 5 : 4, "%s", err.c_str());
        }
        std::unique_ptr<unsigned char[]> vol(new (std::nothrow) unsigned char[es * nvol]);
        if (!vol) return die(5, "host allocation failed");
        if (hipMemcpy(vol.get(), src, es * nvol, hipMemcpyDeviceToHost) != hipSuccess)
            return die(5, "copy from device failed");
        if (!mrc_write_volume(o.output, hout, vol.get(), mode, o.mrc2014, err)) return die(4, "%s", err.c_str());
        return 0;
    };
    if (o.f2b) {
        // byte output over the whole volume's mean and deviation, as
        // TiltRecZ-cuda writes it (recon_z.cu:253-262)
        DevBuf dbytes;
        float stats[2] = {0, 0};
        if (hipMalloc(&dbytes.p, nvol) != hipSuccess) {
            drop_ctx();
            return die(5, "device allocation failed");
        }
        s = trec_float_to_byte(ctx, (const float *)dvol.p, nvol, (unsigned char *)dbytes.p, stats, nullptr);
        if (s != TREC_OK) {
            std::string msg = std::string(trec_status_string(s)) + ": " + trec_last_error(ctx);
            drop_ctx();
            return die(6, "byte conversion failed: %s", msg.c_str());
        }
        drop_ctx();
        std::vector<unsigned char> bytes(nvol);
        if (hipMemcpy(bytes.data(), dbytes.p, nvol, hipMemcpyDeviceToHost) != hipSuccess)
            return die(5, "copy from device failed");
        printf("f2b: mean %g, standard deviation %g\n", stats[0], stats[1]);
        if (!mrc_write_bytes(o.output, hout, bytes.data(), err)) return die(4, "%s", err.c_str());
    } else if (o.output_mode == MRC_HALF) {
        // half output: the GPU rounds to nearest even, half the bytes come down
        DevBuf dhalf;
        if (hipMalloc(&dhalf.p, 2 * nvol) != hipSuccess) {
            drop_ctx();
            return die(5, "device allocation failed");
        }
        s = trec_convert(ctx, dvol.p, TREC_DT_F32, dhalf.p, TREC_DT_F16, nvol, nullptr);
        drop_ctx();
        if (s != TREC_OK) return die(6, "half conversion failed: %s", trec_status_string(s));
        if (const int rc = write_out(dhalf.p, MRC_HALF)) return rc;
    } else {
        drop_ctx();
        if (const int rc = write_out(dvol.p, MRC_FLOAT)) return rc;
    }
    const double t_end = now_s();
    printf("read %.3f s, upload %.3f s, reconstruct %.3f s, write %.3f s, total %.3f s\n", t_read - t0,
           t_up - t_read, t_rec - t_up, t_end - t_rec, t_end - t0);
    return 0;
}
