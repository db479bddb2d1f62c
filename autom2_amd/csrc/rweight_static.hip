// rweight_static.hip -- trec_filter_rweight_static: the R-weighting filter
// (ApplyFilterInplace mode 2, src/TiltRec/src/filter/filter_prj.cpp:652-802)
// for callers that cannot allocate device memory.
//
// trec_filter_rweight builds the filter's circulant kernel on the host and
// copies it into the context's workspace through pinned staging, both of
// which may allocate.  This variant builds the kernel on the device in a
// fixed, file-scope device array (`scratch`) and then runs the same filter
// kernels, so a call allocates nothing.  The array is held by every device
// the library runs on until the process ends; its device address is taken
// once per device, with that device current, when a context is created
// there (trec_context_create calls static_scratch_init).
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "trec_host.h"
#include "trec_internal.h"

namespace trec {

namespace {

// Longest circulant kernel: N = nx + nx / 10 <= RW_STATIC_MAX_N, i.e. rows
// of up to 7448 pixels, the widest rows the filter kernels hold in one CU's
// LDS (launch_rweight).
constexpr int RW_STATIC_MAX_N = 8192;
// Layout of the array: cos table [N] double | weights [N / 2 + 1] double |
// kernel [N] float.
constexpr size_t RW_CT_OFF = 0;
constexpr size_t RW_W_OFF = RW_CT_OFF + sizeof(double) * RW_STATIC_MAX_N;
constexpr size_t RW_HK_OFF = RW_W_OFF + sizeof(double) * (RW_STATIC_MAX_N / 2 + 1 + 31);
constexpr size_t RW_SCRATCH_BYTES = RW_HK_OFF + sizeof(float) * RW_STATIC_MAX_N;

static __device__ __align__(256) char scratch[RW_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};

// The host's rweight_kernel (filter_prj.cpp:674-736 as a circular
// convolution), evaluated on the device in the same double operations and
// order: ct[m] = cos(2 pi m / N); W[0] = 0.2, W[k] = k for k <= K,
// K exp(-(k - K)^2 / sig^2) beyond.
__global__ __launch_bounds__(256) void k_rw_tables(int N, int K, double sig, double *__restrict__ ct,
                                                   double *__restrict__ W)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < N) ct[i] = cos(2.0 * M_PI * i / N);
    if (i < N / 2 + 1) {
        const int k = i;
        W[k] = k == 0 ? 0.2 : (k <= K ? (double)k : (double)K * exp(-((double)(k - K) * (k - K)) / (sig * sig)));
    }
}

// hk[d] = (W[0] + sum_k 2 W[k] ct[d k mod N] (+ W[N/2] (-1)^d for even N)) / X,
// each sum in ascending k as the host adds it.
__global__ __launch_bounds__(256) void k_rw_kernel(int N, int X, const double *__restrict__ ct,
                                                   const double *__restrict__ W, float *__restrict__ hk)
{
#pragma clang fp contract(off)
    const int d = blockIdx.x * 256 + threadIdx.x;
    if (d >= N) return;
    double acc = W[0];
    int idx = 0;  // d * k mod N, advanced by d
    for (int k = 1; k <= (N - 1) / 2; ++k) {
        idx += d;
        if (idx >= N) idx -= N;
        acc += 2.0 * W[k] * ct[idx];
    }
    if (N % 2 == 0) acc += W[N / 2] * ((d % 2) ? -1.0 : 1.0);
    hk[d] = (float)(acc / X);
}

}  // namespace

// Takes the device address of `scratch` on `device` (which must be current)
// once per device.
hipError_t static_scratch_init(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

}  // namespace trec

using namespace trec;

extern "C" trec_status trec_filter_rweight_static(trec_context *ctx, float *rows, int nx, long long nrows,
                                                  void *stream)
{
    if (!ctx) return TREC_ERR_INVALID;
    clear_error(ctx);
    if (!rows || nx < 2 || nrows < 0) return fail(ctx, TREC_ERR_INVALID, "bad filter arguments");
    const int N = nx + nx / 10;
    if (N > RW_STATIC_MAX_N)
        return fail(ctx, TREC_ERR_INVALID,
                    "rows of %d pixels need a filter kernel of %d taps; the no-allocation filter holds at most %d "
                    "(rows up to 7448 pixels)",
                    nx, N, RW_STATIC_MAX_N);
    if (nrows == 0) return TREC_OK;
    const int device = context_device(ctx);
    if (hipSetDevice(device) != hipSuccess) return fail(ctx, TREC_ERR_DEVICE, "hipSetDevice(%d) failed", device);
    void *ptr = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (device >= 0 && device < MAX_DEVICES) ptr = g_scratch_ptr[device];
    }
    if (!ptr) return fail(ctx, TREC_ERR_DEVICE, "no scratch address for device %d", device);
    char *base = static_cast<char *>(ptr);
    double *ct = reinterpret_cast<double *>(base + RW_CT_OFF);
    double *W = reinterpret_cast<double *>(base + RW_W_OFF);
    float *hk = reinterpret_cast<float *>(base + RW_HK_OFF);
    const int K = (int)floorf((float)N * 0.05f);
    const double sig = (double)((float)N * 0.5f);
    hipStream_t st = (hipStream_t)stream;
    const unsigned nb = (unsigned)((N + 255) / 256);
    hipLaunchKernelGGL(k_rw_tables, dim3(nb), dim3(256), 0, st, N, K, sig, ct, W);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_rw_kernel, dim3(nb), dim3(256), 0, st, N, nx, ct, W, hk);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = launch_rweight(rows, hk, N, nx, nrows, 0, 1, 1, nx, st);
    if (e != hipSuccess) return fail(ctx, TREC_ERR_DEVICE, "no-allocation R-weighting: %s", hipGetErrorString(e));
    return TREC_OK;
}
