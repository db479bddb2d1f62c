// kernels.hip -- CDNA4 (gfx950) kernels of libtiltrec_hip.so.
//
// Row-blocked fast path (pitch 0, |a1| >= 0.2): every volume row y is an
// independent XZ problem fed by row y of every tilt, so each thread carries
// RB = 16 rows and evaluates the geometry once for all of them.  The kernels
// are deterministic: there are no float atomics, and every sum is taken in a
// fixed order whatever the launch mix, so a job gives the same bits alone or
// batched with others.
//
//   forward   pixel-driven gather of TiltRec's voxel splat (Reproject,
//             main-mpi.cpp:286-344; CuReprojectKernel, basic.cu:237-264).
//             Voxel x with detector coordinate u adds (1 - frac u) to pixel
//             floor(u) and frac u to pixel floor(u) + 1, i.e. weight
//             max(0, 1 - |u - i|) to pixel i; a thread owns pixel i of one
//             tilt and walks the z layers of its chunk.
//   combine   residual p - R / Rwt over the z-chunk partials
//             (CuCalcProjectionDiffKernel, basic.cu:142-158).
//   back      voxel-driven gather over all tilts with the exact double
//             ValCoef geometry (main-mpi.cpp:204-226, 526-562), fused with
//             the SIRT update (basic.cu:177-190) or the BPT / SART updates
//             (basic.cu:383-415, 417-438).
//
// General path (any pitch): natural layouts, 4-tap bilinear splat and
// gather as the reference does it, float atomics in the splat.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "trec_internal.h"

namespace trec {

// ------------------------------------------------------------------ helpers

// Work-groups are dispatched to the 8 XCDs round-robin by linear id.  This
// maps the linear id so that each XCD walks one contiguous range of the work
// list; neighbouring work items then share that XCD's L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned n)
{
    const unsigned q = n >> 3, r = n & 7u;
    const unsigned xcd = b & 7u, k = b >> 3;
    return xcd * q + (xcd < r ? xcd : r) + k;
}

// Locates the job that owns work-group `b` of a multi-job launch.
__device__ __forceinline__ int find_job(const LaunchList &L, unsigned b)
{
    int j = 0;
    while (j + 1 < L.n && b >= L.off[j + 1]) ++j;
    return j;
}

// Row-block storage: RB values per (pixel or voxel), in float, half or bf16.
template <typename PT>
struct Rows;

template <>
struct Rows<float> {
    __device__ static __forceinline__ void load(const void *base, size_t cell, float (&v)[RB])
    {
        const float4 *p = reinterpret_cast<const float4 *>(static_cast<const float *>(base) + cell * RB);
#pragma unroll
        for (int k = 0; k < RB / 4; ++k) {
            float4 t = p[k];
            v[4 * k + 0] = t.x;
            v[4 * k + 1] = t.y;
            v[4 * k + 2] = t.z;
            v[4 * k + 3] = t.w;
        }
    }
    __device__ static __forceinline__ void store(void *base, size_t cell, const float (&v)[RB])
    {
        float4 *p = reinterpret_cast<float4 *>(static_cast<float *>(base) + cell * RB);
#pragma unroll
        for (int k = 0; k < RB / 4; ++k) p[k] = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
    }
    __device__ static __forceinline__ void put4(void *base, size_t idx, float a, float b, float c, float d)
    {
        *reinterpret_cast<float4 *>(static_cast<float *>(base) + idx) = make_float4(a, b, c, d);
    }
};

template <typename H>
struct Rows16 {
    // 16-bit storage: RB values = 32 bytes = two 16-byte loads.
    __device__ static __forceinline__ void load(const void *base, size_t cell, float (&v)[RB])
    {
        const uint4 *p = reinterpret_cast<const uint4 *>(static_cast<const H *>(base) + cell * RB);
#pragma unroll
        for (int k = 0; k < RB / 8; ++k) {
            uint4 t = p[k];
            const H *h = reinterpret_cast<const H *>(&t);
#pragma unroll
            for (int e = 0; e < 8; ++e) v[8 * k + e] = (float)h[e];
        }
    }
    __device__ static __forceinline__ void store(void *base, size_t cell, const float (&v)[RB])
    {
        uint4 *p = reinterpret_cast<uint4 *>(static_cast<H *>(base) + cell * RB);
#pragma unroll
        for (int k = 0; k < RB / 8; ++k) {
            uint4 t;
            H *h = reinterpret_cast<H *>(&t);
#pragma unroll
            for (int e = 0; e < 8; ++e) h[e] = (H)v[8 * k + e];
            p[k] = t;
        }
    }
    __device__ static __forceinline__ void put4(void *base, size_t idx, float a, float b, float c, float d)
    {
        H t[4] = {(H)a, (H)b, (H)c, (H)d};
        *reinterpret_cast<uint2 *>(static_cast<H *>(base) + idx) = *reinterpret_cast<uint2 *>(t);
    }
};

template <>
struct Rows<_Float16> : Rows16<_Float16> {};
template <>
struct Rows<__bf16> : Rows16<__bf16> {};

// Exact TiltRec-mpi geometry (ValCoef, main-mpi.cpp:204-226) at pitch 0:
// u = ((a0 + a1 * X) + a3 * Z) + ox in double, taps floor(u) and floor(u)+1,
// fraction rounded to float as Weight::x_min_del is.
__device__ __forceinline__ void exact_tap(const TiltConst &t, double Xc, double Zc, int &g, float &dx)
{
    double u = __dadd_rn(__dadd_rn(__dadd_rn(t.a0, __dmul_rn(t.a1, Xc)), __dmul_rn(t.a3, Zc)), t.ox);
    double f = floor(u);
    g = (int)f;
    dx = (float)__dsub_rn(u, f);
}

// ------------------------------------------------------------------ layouts

// stack [P][Y][X] float -> proj [nyb][P][X][RB] (storage type), / scale.
template <typename PT>
__global__ __launch_bounds__(WG) void k_block_stack(const JobDesc *__restrict__ jd, const float *__restrict__ stack)
{
    const JobDesc &J = *jd;
    __shared__ float tile[RB][65];
    const int nxt = (J.X + 63) / 64;
    unsigned b = blockIdx.x;
    const int xt = b % nxt;
    b /= nxt;
    const int th = b % J.P;
    const int yb = b / J.P;
    const int x0 = xt * 64;
    const float inv = 1.0f / J.scale;
    for (int k = threadIdx.x; k < RB * 64; k += WG) {
        const int r = k / 64, xx = k % 64;
        const int y = yb * RB + r, x = x0 + xx;
        float v = 0.f;
        if (y < J.Y && x < J.X) v = stack[((size_t)th * J.Y + y) * J.X + x] * inv;
        tile[r][xx] = v;
    }
    __syncthreads();
    const int xx = threadIdx.x / 4, q = threadIdx.x % 4;
    const int x = x0 + xx;
    if (x < J.X) {
        const size_t idx = (((size_t)yb * J.P + th) * J.X + x) * RB + q * 4;
        Rows<PT>::put4(J.proj, idx, tile[q * 4][xx], tile[q * 4 + 1][xx], tile[q * 4 + 2][xx], tile[q * 4 + 3][xx]);
    }
}

// volume (layout Y [Y][T][X] or Z [T][Y][X]) -> vol [nyb][T][X][RB], / scale.
__global__ __launch_bounds__(WG) void k_block_volume(const JobDesc *__restrict__ jd, const float *__restrict__ src,
                                                     int layout)
{
    const JobDesc &J = *jd;
    __shared__ float tile[RB][65];
    const int nxt = (J.X + 63) / 64;
    unsigned b = blockIdx.x;
    const int xt = b % nxt;
    b /= nxt;
    const int z = b % J.T;
    const int yb = b / J.T;
    const int x0 = xt * 64;
    const float inv = 1.0f / J.scale;
    for (int k = threadIdx.x; k < RB * 64; k += WG) {
        const int r = k / 64, xx = k % 64;
        const int y = yb * RB + r, x = x0 + xx;
        float v = 0.f;
        if (y < J.Y && x < J.X) {
            const size_t o = layout == TREC_LAYOUT_Y ? ((size_t)y * J.T + z) * J.X + x : ((size_t)z * J.Y + y) * J.X + x;
            v = src[o] * inv;
        }
        tile[r][xx] = v;
    }
    __syncthreads();
    const int xx = threadIdx.x / 4, q = threadIdx.x % 4;
    const int x = x0 + xx;
    if (x < J.X) {
        const size_t idx = (((size_t)yb * J.T + z) * J.X + x) * RB + q * 4;
        *reinterpret_cast<float4 *>(J.vol + idx) =
            make_float4(tile[q * 4][xx], tile[q * 4 + 1][xx], tile[q * 4 + 2][xx], tile[q * 4 + 3][xx]);
    }
}

// vol [nyb][T][X][RB] -> volume (layout Y or Z), * scale.
__global__ __launch_bounds__(WG) void k_unblock_volume(const JobDesc *__restrict__ jd, float *__restrict__ dst,
                                                       int layout)
{
    const JobDesc &J = *jd;
    __shared__ float tile[RB][65];
    const int nxt = (J.X + 63) / 64;
    unsigned b = blockIdx.x;
    const int xt = b % nxt;
    b /= nxt;
    const int z = b % J.T;
    const int yb = b / J.T;
    const int x0 = xt * 64;
    {
        const int xx = threadIdx.x / 4, q = threadIdx.x % 4;
        const int x = x0 + xx;
        if (x < J.X) {
            const float4 v = *reinterpret_cast<const float4 *>(J.vol + (((size_t)yb * J.T + z) * J.X + x) * RB + q * 4);
            tile[q * 4][xx] = v.x;
            tile[q * 4 + 1][xx] = v.y;
            tile[q * 4 + 2][xx] = v.z;
            tile[q * 4 + 3][xx] = v.w;
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < RB * 64; k += WG) {
        const int r = k / 64, xx = k % 64;
        const int y = yb * RB + r, x = x0 + xx;
        if (y < J.Y && x < J.X) {
            const size_t o = layout == TREC_LAYOUT_Y ? ((size_t)y * J.T + z) * J.X + x : ((size_t)z * J.Y + y) * J.X + x;
            dst[o] = tile[r][xx] * J.scale;
        }
    }
}

__global__ void k_zero(float *__restrict__ p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0.f;
}

// ------------------------------------------------------------------ weights

// inv_c[z][x] = 1 / C (or 0), C = sum over tilts of the in-bounds tap weights
// of voxel (x, z), accumulated as UpdateWeightsByProjDiff does
// (main-mpi.cpp:547-559): c = 0; c += 1 - dx; c += dx; wtvol += c.
// C does not depend on the row at pitch 0.
__global__ __launch_bounds__(WG) void k_voxel_weights(const JobDesc *__restrict__ jd)
{
    const JobDesc &J = *jd;
    const int idx = blockIdx.x * WG + threadIdx.x;
    if (idx >= J.X * J.T) return;
    const int z = idx / J.X, x = idx % J.X;
    const double Xc = (double)((float)x - (float)J.tc[0].ox);
    const double Zc = (double)((float)z - (float)J.tc[0].oz);
    float C = 0.f;
    for (int th = 0; th < J.P; ++th) {
        int g;
        float dx;
        exact_tap(J.tc[th], Xc, Zc, g, dx);
        float c = 0.f;
        if ((unsigned)g < (unsigned)J.X) c += 1.f - dx;
        if ((unsigned)(g + 1) < (unsigned)J.X) c += dx;
        C += c;
    }
    J.inv_c[idx] = (C > J.thr_upd) ? 1.0f / C : 0.f;
}

// ------------------------------------------------------------------ forward

unsigned forward_blocks(const JobDesc &h, int ntheta)
{
    const unsigned npb = (h.X + FP_PIX - 1) / FP_PIX;
    return npb * (unsigned)ntheta * (unsigned)h.nzc * (unsigned)h.nyb;
}

// Pixel-driven forward projection of one z chunk.
//   FUSED: nzc == 1, write the residual p - R / Rwt directly.
//   ONES:  splat of ones (Rwt); writes inv_rwt[t][i] from row block 0.
// For pixel i of tilt t the voxels of layer z with weight > 0 satisfy
// i - 1 < u(x, z) < i + 1.  With s(z) = (i - 1 - c0 - a3 z) / a1 the first
// is xf = floor(s) + 1, at distance d0 = a1 (1 - frac s) - 1 from i, and
// candidates xf + j sit at d0 + a1 j, j < ncand = ceil(2 / a1).
template <typename PT, bool FUSED, bool ONES>
__global__ __launch_bounds__(WG) void k_forward(const LaunchList L, unsigned nblk)
{
    const unsigned b0 = xcd_remap(blockIdx.x, nblk);
    const int j = find_job(L, b0);
    const JobDesc &J = *L.job[j];
    unsigned lb = b0 - L.off[j];
    const int ntheta = L.ntheta[j] ? L.ntheta[j] : J.P;
    const unsigned npb = (J.X + FP_PIX - 1) / FP_PIX;
    const int pb = lb % npb;
    lb /= npb;
    const int th = L.theta0[j] + (int)(lb % ntheta);
    lb /= ntheta;
    const int zc = lb % J.nzc;
    const int yb = lb / J.nzc;
    const int i = pb * FP_PIX + threadIdx.x;
    if (i >= J.X) return;
    const TiltConst &t = J.tc[th];
    const int X = J.X;
    const int z0 = zc * J.zc_len;
    const int z1 = min(J.T, z0 + J.zc_len);
    const double sbase = ((double)(i - 1) - t.c0) * t.inv_a1;
    const float a1f = t.a1f, a1m1 = t.a1f - 1.0f;
    const int ncand = t.ncand;

    float acc[ONES ? 1 : RB];
#pragma unroll
    for (int r = 0; r < (ONES ? 1 : RB); ++r) acc[r] = 0.f;

    const float *vbase = J.vol + (size_t)yb * J.T * X * RB;
    for (int z = z0; z < z1; ++z) {
        const double s = fma(t.slope, (double)z, sbase);
        const double fl = floor(s);
        const float frac = (float)(s - fl);
        const int xf = (int)fl + 1;
        const float d0 = fmaf(-a1f, frac, a1m1);
        const float *vz = vbase + (size_t)z * X * RB;
        for (int c = 0; c < ncand; ++c) {
            const int x = xf + c;
            const float d = fmaf(a1f, (float)c, d0);
            const bool ok = (unsigned)x < (unsigned)X;
            const float w = ok ? fmaxf(0.f, 1.f - fabsf(d)) : 0.f;
            if constexpr (ONES) {
                acc[0] += w;
            } else {
                float v[RB];
                Rows<float>::load(vz, ok ? x : 0, v);
#pragma unroll
                for (int r = 0; r < RB; ++r) acc[r] = fmaf(w, v[r], acc[r]);
            }
        }
    }

    if constexpr (ONES) {
        if (yb == 0) {
            J.inv_rwt[th * X + i] = (acc[0] > J.thr_res) ? 1.0f / acc[0] : 0.f;
            if (J.rwt) J.rwt[th * X + i] = acc[0];
        }
    } else if constexpr (FUSED) {
        const size_t cell = ((size_t)yb * J.P + th) * X + i;
        const float inv = J.inv_rwt[th * X + i];
        float p[RB];
        Rows<PT>::load(J.proj, cell, p);
#pragma unroll
        for (int r = 0; r < RB; ++r) acc[r] = fmaf(-acc[r], inv, p[r]);
        Rows<PT>::store(J.res, cell, acc);
    } else {
        const size_t cell = (((size_t)zc * J.nyb + yb) * J.P + th) * X + i;
        Rows<float>::store(J.part, cell, acc);
    }
}

unsigned combine_blocks(const JobDesc &h, int ntheta)
{
    const unsigned cells = (unsigned)h.nyb * (unsigned)ntheta * (unsigned)h.X;
    return (cells + WG - 1) / WG;
}

// res = p - (sum of the z-chunk partials, in chunk order) * inv_rwt.
template <typename PT>
__global__ __launch_bounds__(WG) void k_combine(const LaunchList L, unsigned nblk)
{
    const unsigned b0 = xcd_remap(blockIdx.x, nblk);
    const int j = find_job(L, b0);
    const JobDesc &J = *L.job[j];
    const unsigned lb = b0 - L.off[j];
    const int ntheta = L.ntheta[j] ? L.ntheta[j] : J.P;
    const size_t k = (size_t)lb * WG + threadIdx.x;
    const size_t ncell = (size_t)J.nyb * ntheta * J.X;
    if (k >= ncell) return;
    const int i = k % J.X;
    const int th = L.theta0[j] + (int)((k / J.X) % ntheta);
    const int yb = k / ((size_t)J.X * ntheta);
    const size_t cell = ((size_t)yb * J.P + th) * J.X + i;
    const size_t zstride = (size_t)J.nyb * J.P * J.X;
    float acc[RB];
    Rows<float>::load(J.part, cell, acc);
    for (int zc = 1; zc < J.nzc; ++zc) {
        float v[RB];
        Rows<float>::load(J.part, cell + zc * zstride, v);
#pragma unroll
        for (int r = 0; r < RB; ++r) acc[r] += v[r];
    }
    const float inv = J.inv_rwt[th * J.X + i];
    float p[RB];
    Rows<PT>::load(J.proj, cell, p);
#pragma unroll
    for (int r = 0; r < RB; ++r) acc[r] = fmaf(-acc[r], inv, p[r]);
    Rows<PT>::store(J.res, cell, acc);
}

// ------------------------------------------------------------------ back

unsigned back_blocks(const JobDesc &h)
{
    const unsigned nxb = (h.X + BP_X - 1) / BP_X;
    const unsigned nzb = (h.T + BP_Z - 1) / BP_Z;
    return nxb * nzb * (unsigned)h.nyb;
}

// Voxel-driven back-projection: a wave holds 64 consecutive x of one z, a
// thread RB rows.  MODE selects the update (BackMode).  For BACK_SIRT and
// BACK_SART the image is the residual, for BACK_BPT the (filtered) stack.
template <typename PT, int MODE>
__global__ __launch_bounds__(WG) void k_back(const LaunchList L, unsigned nblk)
{
    const unsigned b0 = xcd_remap(blockIdx.x, nblk);
    const int j = find_job(L, b0);
    const JobDesc &J = *L.job[j];
    unsigned lb = b0 - L.off[j];
    const unsigned nzb = (J.T + BP_Z - 1) / BP_Z;
    const unsigned nxb = (J.X + BP_X - 1) / BP_X;
    const int zb = lb % nzb;
    lb /= nzb;
    const int xb = lb % nxb;
    const int yb = lb / nxb;
    const int x = xb * BP_X + (threadIdx.x % BP_X);
    const int z = zb * BP_Z + (threadIdx.x / BP_X);
    if (x >= J.X || z >= J.T) return;
    const int X = J.X;
    const double Xc = (double)((float)x - (float)J.tc[0].ox);
    const double Zc = (double)((float)z - (float)J.tc[0].oz);
    const void *img = (MODE == BACK_BPT) ? J.proj : J.res;
    const int th0 = L.theta0[j];
    const int th1 = L.ntheta[j] ? th0 + L.ntheta[j] : J.P;

    float val[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) val[r] = 0.f;
    float csart = 0.f;

    for (int th = th0; th < th1; ++th) {
        int g;
        float dx;
        exact_tap(J.tc[th], Xc, Zc, g, dx);
        const bool v0 = (unsigned)g < (unsigned)X;
        const bool v1 = (unsigned)(g + 1) < (unsigned)X;
        const float w0 = v0 ? 1.f - dx : 0.f;
        const float w1 = v1 ? dx : 0.f;
        const size_t row = ((size_t)yb * J.P + th) * X;
        float r0[RB], r1[RB];
        Rows<PT>::load(img, row + (v0 ? g : 0), r0);
        Rows<PT>::load(img, row + (v1 ? g + 1 : 0), r1);
        if constexpr (MODE == BACK_SIRT) {
#pragma unroll
            for (int r = 0; r < RB; ++r) val[r] = fmaf(w1, r1[r], fmaf(w0, r0[r], val[r]));
        } else {
            // c = 0; c += 1 - dx; c += dx (BilinearValue order)
            float c = 0.f;
            c += w0;
            c += w1;
            if constexpr (MODE == BACK_SART) csart = c;
            if (c > J.thr_upd) {
                const float ic = 1.0f / c;
#pragma unroll
                for (int r = 0; r < RB; ++r) val[r] = fmaf(fmaf(w1, r1[r], w0 * r0[r]), ic, val[r]);
            }
        }
    }

    float *vp = J.vol + (((size_t)yb * J.T + z) * X + x) * RB;
    float v[RB];
    Rows<float>::load(vp, 0, v);
    if constexpr (MODE == BACK_SIRT) {
        const float f = J.gamma * J.inv_c[z * X + x];
#pragma unroll
        for (int r = 0; r < RB; ++r) v[r] = fmaf(val[r], f, v[r]);
    } else if constexpr (MODE == BACK_SART) {
        if (!(csart > J.thr_upd)) return;
#pragma unroll
        for (int r = 0; r < RB; ++r) v[r] = fmaf(val[r], J.gamma, v[r]);
    } else {
#pragma unroll
        for (int r = 0; r < RB; ++r) v[r] += val[r];
    }
    Rows<float>::store(vp, 0, v);
}

// ------------------------------------------------------------------ R-weighting

// R-weighting (ApplyFilterInplace mode 2, filter_prj.cpp:652-736) as the
// circular convolution it is: the row is extended to N = X + X/10 with
// buf2fft_padding_1D's ramp, and y[j] = sum_d hk[d] e[(j - d) mod N] for
// j < X, where hk is the inverse DFT of the weights, already divided by X.
// A work-group filters 4 rows held interleaved in LDS as float4.
// Row k of the launch starts at data + (k / rpg) * gstride + (k % rpg) * rstride
// and its elements are xstride apart.
__global__ __launch_bounds__(WG) void k_rweight(float *__restrict__ data, const float *__restrict__ hk, int N,
                                                int X, long long nrows, long long rstride, int xstride, int rpg,
                                                long long gstride)
{
    extern __shared__ float smem[];
    float *h = smem;                                   // [N]
    float4 *e = reinterpret_cast<float4 *>(smem + ((N + 3) & ~3));  // [N]
    const long long row0 = (long long)blockIdx.x * 4;
    long long base[4];
    bool live[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const long long k = row0 + q;
        live[q] = k < nrows;
        const long long kk = live[q] ? k : row0;
        base[q] = (kk / rpg) * gstride + (kk % rpg) * rstride;
    }
    for (int m = threadIdx.x; m < N; m += WG) h[m] = hk[m];
    const int nxp = N - X + 1;
    for (int jx = threadIdx.x; jx < N; jx += WG) {
        float v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float o;
            if (jx < X) {
                o = data[base[q] + (long long)jx * xstride];
            } else {
                const float a = data[base[q]];
                const float bb = data[base[q] + (long long)(X - 1) * xstride];
                o = (float(jx - X + 1) / float(nxp)) * a + (float(N - jx) / float(nxp)) * bb;
            }
            v[q] = o;
        }
        e[jx] = make_float4(v[0], v[1], v[2], v[3]);
    }
    __syncthreads();
    for (int jx = threadIdx.x; jx < X; jx += WG) {
        float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f), a1 = a0;
        // d <= jx: e[jx - d]; d > jx: e[jx - d + N].  Two accumulators per
        // row halve the dependent-add chain.
        int d = 0;
        for (; d + 1 <= jx; d += 2) {
            const float h0 = h[d], h1 = h[d + 1];
            const float4 e0 = e[jx - d], e1 = e[jx - d - 1];
            a0.x = fmaf(h0, e0.x, a0.x); a0.y = fmaf(h0, e0.y, a0.y);
            a0.z = fmaf(h0, e0.z, a0.z); a0.w = fmaf(h0, e0.w, a0.w);
            a1.x = fmaf(h1, e1.x, a1.x); a1.y = fmaf(h1, e1.y, a1.y);
            a1.z = fmaf(h1, e1.z, a1.z); a1.w = fmaf(h1, e1.w, a1.w);
        }
        for (; d <= jx; ++d) {
            const float h0 = h[d];
            const float4 e0 = e[jx - d];
            a0.x = fmaf(h0, e0.x, a0.x); a0.y = fmaf(h0, e0.y, a0.y);
            a0.z = fmaf(h0, e0.z, a0.z); a0.w = fmaf(h0, e0.w, a0.w);
        }
        for (; d + 1 < N; d += 2) {
            const float h0 = h[d], h1 = h[d + 1];
            const float4 e0 = e[jx - d + N], e1 = e[jx - d - 1 + N];
            a0.x = fmaf(h0, e0.x, a0.x); a0.y = fmaf(h0, e0.y, a0.y);
            a0.z = fmaf(h0, e0.z, a0.z); a0.w = fmaf(h0, e0.w, a0.w);
            a1.x = fmaf(h1, e1.x, a1.x); a1.y = fmaf(h1, e1.y, a1.y);
            a1.z = fmaf(h1, e1.z, a1.z); a1.w = fmaf(h1, e1.w, a1.w);
        }
        for (; d < N; ++d) {
            const float h0 = h[d];
            const float4 e0 = e[jx - d + N];
            a0.x = fmaf(h0, e0.x, a0.x); a0.y = fmaf(h0, e0.y, a0.y);
            a0.z = fmaf(h0, e0.z, a0.z); a0.w = fmaf(h0, e0.w, a0.w);
        }
        const float out[4] = {a0.x + a1.x, a0.y + a1.y, a0.z + a1.z, a0.w + a1.w};
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (live[q]) data[base[q] + (long long)jx * xstride] = out[q];
    }
}

// Sum of squares in double: the stack's scale for the 16-bit forms and the
// residual norms (one atomic per work-group).
template <typename T>
__global__ __launch_bounds__(WG) void k_sumsq(const T *__restrict__ a, size_t n, double *__restrict__ out)
{
    __shared__ double red[WG];
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * WG + threadIdx.x; i < n; i += (size_t)gridDim.x * WG) {
        const double v = (double)(float)a[i];
        s += v * v;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = WG / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

// ------------------------------------------------------------------ general path

// Exact ValCoef for any pitch (main-mpi.cpp:204-226).
__device__ __forceinline__ void general_map(const TiltGeneral &t, int x, int y, int z, int &xm, int &ym, float &xd,
                                            float &yd)
{
    const double Xc = (double)((float)x - (float)t.ox);
    const double Yc = (double)((float)y - (float)t.oy);
    const double Zc = (double)((float)z - (float)t.oz);
    const double n0 = __dadd_rn(__dadd_rn(__dadd_rn(t.a[0], __dmul_rn(t.a[1], Xc)), __dmul_rn(t.a[2], Yc)),
                                __dmul_rn(t.a[3], Zc));
    const double n1 = __dadd_rn(__dadd_rn(__dadd_rn(t.b[0], __dmul_rn(t.b[1], Xc)), __dmul_rn(t.b[2], Yc)),
                                __dmul_rn(t.b[3], Zc));
    const double u = __dadd_rn(n0, t.ox), v = __dadd_rn(n1, t.oy);
    const double fu = floor(u), fv = floor(v);
    xm = (int)fu;
    ym = (int)fv;
    xd = (float)(u - fu);
    yd = (float)(v - fv);
}

// Splat of one voxel column (x, z) over all rows into rval (or rwt with ones),
// CuReprojectKernel / CuReWeightKernel semantics with float atomics.
__global__ __launch_bounds__(WG) void k_general_forward(const GeneralJob g, int theta0, int ones)
{
    const int idx = blockIdx.x * WG + threadIdx.x;
    if (idx >= g.X * g.T) return;
    const int th = theta0 + blockIdx.y;
    const int z = idx / g.X, x = idx % g.X;
    const TiltGeneral &t = g.tg[th];
    float *out = (ones ? g.rwt : g.rval) + (size_t)th * g.Y * g.X;
    for (int y = 0; y < g.Y; ++y) {
        const size_t vi = g.layout == TREC_LAYOUT_Y ? ((size_t)y * g.T + z) * g.X + x : ((size_t)z * g.Y + y) * g.X + x;
        const float v = ones ? 1.f : g.vol[vi];
        int xm, ym;
        float xd, yd;
        general_map(t, x, y, z, xm, ym, xd, yd);
        if (xm >= 0 && xm < g.X && ym >= 0 && ym < g.Y) atomicAdd(out + (size_t)ym * g.X + xm, (1 - xd) * (1 - yd) * v);
        if (xm + 1 >= 0 && xm + 1 < g.X && ym >= 0 && ym < g.Y) atomicAdd(out + (size_t)ym * g.X + xm + 1, xd * (1 - yd) * v);
        if (xm >= 0 && xm < g.X && ym + 1 >= 0 && ym + 1 < g.Y) atomicAdd(out + (size_t)(ym + 1) * g.X + xm, (1 - xd) * yd * v);
        if (xm + 1 >= 0 && xm + 1 < g.X && ym + 1 >= 0 && ym + 1 < g.Y)
            atomicAdd(out + (size_t)(ym + 1) * g.X + xm + 1, xd * yd * v);
    }
}

// rval <- p - rval / rwt (rwt above the threshold), in place.
__global__ __launch_bounds__(WG) void k_general_residual(const GeneralJob g, int theta0, size_t n)
{
    const size_t i = (size_t)blockIdx.x * WG + threadIdx.x;
    if (i >= n) return;
    const size_t k = (size_t)theta0 * g.Y * g.X + i;
    float r = g.rval[k];
    const float w = g.rwt[k];
    if (w > g.thr_res) r /= w;  // weights are >= 0: thr_res = 0 is the reference's w != 0
    g.rval[k] = g.stack[k] - r;
}

__device__ __forceinline__ void general_bilinear(const float *img, int X, int Y, int xm, int ym, float xd, float yd,
                                                 float &s, float &c)
{
    if (xm >= 0 && xm < X && ym >= 0 && ym < Y) {
        s += (1 - xd) * (1 - yd) * img[(size_t)ym * X + xm];
        c += (1 - xd) * (1 - yd);
    }
    if (xm + 1 >= 0 && xm + 1 < X && ym >= 0 && ym < Y) {
        s += xd * (1 - yd) * img[(size_t)ym * X + xm + 1];
        c += xd * (1 - yd);
    }
    if (xm >= 0 && xm < X && ym + 1 >= 0 && ym + 1 < Y) {
        s += (1 - xd) * yd * img[(size_t)(ym + 1) * X + xm];
        c += (1 - xd) * yd;
    }
    if (xm + 1 >= 0 && xm + 1 < X && ym + 1 >= 0 && ym + 1 < Y) {
        s += xd * yd * img[(size_t)(ym + 1) * X + xm + 1];
        c += xd * yd;
    }
}

// Voxel-driven gather, BackMode semantics, natural layouts.
__global__ __launch_bounds__(WG) void k_general_back(const GeneralJob g, int mode, int theta0, int ntheta,
                                                     const float *__restrict__ img)
{
    const size_t idx0 = (size_t)blockIdx.x * WG + threadIdx.x;
    if (idx0 >= (size_t)g.X * g.T * g.Y) return;
    const int x = idx0 % g.X;
    const int z = (idx0 / g.X) % g.T;
    const int y = idx0 / ((size_t)g.X * g.T);
    const size_t idx = g.layout == TREC_LAYOUT_Y ? idx0 : ((size_t)z * g.Y + y) * g.X + x;
    float val = 0.f, wt = 0.f;
    float out = g.vol[idx];
    for (int th = theta0; th < theta0 + ntheta; ++th) {
        int xm, ym;
        float xd, yd;
        general_map(g.tg[th], x, y, z, xm, ym, xd, yd);
        float s = 0.f, c = 0.f;
        general_bilinear(img + (size_t)th * g.Y * g.X, g.X, g.Y, xm, ym, xd, yd, s, c);
        if (mode == BACK_SIRT) {
            val += s;
            wt += c;
        } else if (c > g.thr_upd) {
            out += (mode == BACK_SART) ? (s / c) * g.gamma : s / c;
        }
    }
    if (mode == BACK_SIRT) {
        if (wt > g.thr_upd) out += val / wt * g.gamma;
    }
    g.vol[idx] = out;
}

// ------------------------------------------------------------------ small helpers

__global__ void k_set_scale(JobDesc *d, const double *sumsq, double n)
{
    const double rms = sqrt(*sumsq / n);
    d->scale = (rms > 0.0 && rms < 3.0e38) ? (float)rms : 1.0f;
}

// blocked [nyb][P][X][RB] float -> [P][Y][X]
__global__ __launch_bounds__(WG) void k_unblock_proj(const float *__restrict__ src, int P, int X, int Y, int nyb,
                                                     float *__restrict__ dst)
{
    const size_t k = (size_t)blockIdx.x * WG + threadIdx.x;
    const size_t n = (size_t)P * Y * X;
    if (k >= n) return;
    const int x = k % X;
    const int y = (k / X) % Y;
    const int th = k / ((size_t)X * Y);
    dst[k] = src[(((size_t)(y / RB) * P + th) * X + x) * RB + (y % RB)];
}

// [P][X] -> [P][Y][X] (the same row for every y)
__global__ __launch_bounds__(WG) void k_broadcast_rows(const float *__restrict__ src, int P, int X, int Y,
                                                       float *__restrict__ dst)
{
    const size_t k = (size_t)blockIdx.x * WG + threadIdx.x;
    const size_t n = (size_t)P * Y * X;
    if (k >= n) return;
    const int x = k % X;
    const int th = k / ((size_t)X * Y);
    dst[k] = src[(size_t)th * X + x];
}

__global__ void k_copy(float *__restrict__ dst, const float *__restrict__ src, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

// ------------------------------------------------------------------ launchers

hipError_t launch_set_scale(JobDesc *d, const double *sumsq, double n, hipStream_t st)
{
    hipLaunchKernelGGL(k_set_scale, dim3(1), dim3(1), 0, st, d, sumsq, n);
    return hipGetLastError();
}

hipError_t launch_unblock_proj(const float *src, int P, int X, int Y, int nyb, float *dst, hipStream_t st)
{
    const size_t n = (size_t)P * Y * X;
    hipLaunchKernelGGL(k_unblock_proj, dim3((unsigned)((n + WG - 1) / WG)), dim3(WG), 0, st, src, P, X, Y, nyb, dst);
    return hipGetLastError();
}

hipError_t launch_broadcast_rows(const float *src, int P, int X, int Y, float *dst, hipStream_t st)
{
    const size_t n = (size_t)P * Y * X;
    hipLaunchKernelGGL(k_broadcast_rows, dim3((unsigned)((n + WG - 1) / WG)), dim3(WG), 0, st, src, P, X, Y, dst);
    return hipGetLastError();
}

hipError_t launch_copy(float *dst, const float *src, size_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<size_t>((n + WG - 1) / WG, 65536);
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(WG), 0, st, dst, src, n);
    return hipGetLastError();
}

static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

hipError_t launch_block_stack(const JobDesc &h, const JobDesc *d, const float *stack, hipStream_t st)
{
    const unsigned n = cdiv(h.X, 64) * (unsigned)h.P * (unsigned)h.nyb;
    switch (h.storage) {
    case TREC_STORE_F16: hipLaunchKernelGGL(k_block_stack<_Float16>, dim3(n), dim3(WG), 0, st, d, stack); break;
    case TREC_STORE_BF16: hipLaunchKernelGGL(k_block_stack<__bf16>, dim3(n), dim3(WG), 0, st, d, stack); break;
    default: hipLaunchKernelGGL(k_block_stack<float>, dim3(n), dim3(WG), 0, st, d, stack); break;
    }
    return hipGetLastError();
}

hipError_t launch_block_volume(const JobDesc &h, const JobDesc *d, const float *src, int layout, hipStream_t st)
{
    const unsigned n = cdiv(h.X, 64) * (unsigned)h.T * (unsigned)h.nyb;
    hipLaunchKernelGGL(k_block_volume, dim3(n), dim3(WG), 0, st, d, src, layout);
    return hipGetLastError();
}

hipError_t launch_unblock_volume(const JobDesc &h, const JobDesc *d, float *dst, int layout, hipStream_t st)
{
    const unsigned n = cdiv(h.X, 64) * (unsigned)h.T * (unsigned)h.nyb;
    hipLaunchKernelGGL(k_unblock_volume, dim3(n), dim3(WG), 0, st, d, dst, layout);
    return hipGetLastError();
}

hipError_t launch_zero(float *p, size_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<size_t>(cdiv(n, WG), 65536);
    hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(WG), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_weights(const JobDesc &h, const JobDesc *d, hipStream_t st)
{
    hipLaunchKernelGGL(k_voxel_weights, dim3(cdiv((size_t)h.X * h.T, WG)), dim3(WG), 0, st, d);
    return hipGetLastError();
}

hipError_t launch_forward(const LaunchList &L, unsigned nblk, int storage, bool fused, hipStream_t st)
{
    // storage < 0 selects the splat of ones (Rwt)
    if (storage < 0) {
        hipLaunchKernelGGL((k_forward<float, false, true>), dim3(nblk), dim3(WG), 0, st, L, nblk);
    } else if (fused) {
        switch (storage) {
        case TREC_STORE_F16: hipLaunchKernelGGL((k_forward<_Float16, true, false>), dim3(nblk), dim3(WG), 0, st, L, nblk); break;
        case TREC_STORE_BF16: hipLaunchKernelGGL((k_forward<__bf16, true, false>), dim3(nblk), dim3(WG), 0, st, L, nblk); break;
        default: hipLaunchKernelGGL((k_forward<float, true, false>), dim3(nblk), dim3(WG), 0, st, L, nblk); break;
        }
    } else {
        hipLaunchKernelGGL((k_forward<float, false, false>), dim3(nblk), dim3(WG), 0, st, L, nblk);
    }
    return hipGetLastError();
}

hipError_t launch_combine(const LaunchList &L, unsigned nblk, int storage, hipStream_t st)
{
    switch (storage) {
    case TREC_STORE_F16: hipLaunchKernelGGL(k_combine<_Float16>, dim3(nblk), dim3(WG), 0, st, L, nblk); break;
    case TREC_STORE_BF16: hipLaunchKernelGGL(k_combine<__bf16>, dim3(nblk), dim3(WG), 0, st, L, nblk); break;
    default: hipLaunchKernelGGL(k_combine<float>, dim3(nblk), dim3(WG), 0, st, L, nblk); break;
    }
    return hipGetLastError();
}

template <int MODE>
static void launch_back_mode(const LaunchList &L, unsigned nblk, int storage, hipStream_t st)
{
    switch (storage) {
    case TREC_STORE_F16: hipLaunchKernelGGL((k_back<_Float16, MODE>), dim3(nblk), dim3(WG), 0, st, L, nblk); break;
    case TREC_STORE_BF16: hipLaunchKernelGGL((k_back<__bf16, MODE>), dim3(nblk), dim3(WG), 0, st, L, nblk); break;
    default: hipLaunchKernelGGL((k_back<float, MODE>), dim3(nblk), dim3(WG), 0, st, L, nblk); break;
    }
}

hipError_t launch_back(const LaunchList &L, unsigned nblk, int storage, int mode, hipStream_t st)
{
    switch (mode) {
    case BACK_SIRT: launch_back_mode<BACK_SIRT>(L, nblk, storage, st); break;
    case BACK_BPT: launch_back_mode<BACK_BPT>(L, nblk, storage, st); break;
    default: launch_back_mode<BACK_SART>(L, nblk, storage, st); break;
    }
    return hipGetLastError();
}

hipError_t launch_rweight(float *data, const float *hk, int N, int X, long long nrows, long long row_stride,
                          int x_stride, int rows_per_group, long long group_stride, hipStream_t st)
{
    const size_t shmem = sizeof(float) * (size_t)((N + 3) & ~3) + sizeof(float4) * (size_t)N;
    if (shmem > 160 * 1024) return hipErrorInvalidValue;  // rows beyond X ~ 7400 do not fit one CU's LDS
    if (shmem > 64 * 1024) {
        static hipError_t attr = hipFuncSetAttribute((const void *)k_rweight,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (attr != hipSuccess) return attr;
    }
    const unsigned blocks = (unsigned)((nrows + 3) / 4);
    hipLaunchKernelGGL(k_rweight, dim3(blocks), dim3(WG), shmem, st, data, hk, N, X, nrows, row_stride, x_stride,
                       rows_per_group, group_stride);
    return hipGetLastError();
}

hipError_t launch_sumsq(const void *a, int storage, size_t n, double *out, hipStream_t st)
{
    const unsigned blocks = (unsigned)std::min<size_t>(cdiv(n, WG), 4096);
    switch (storage) {
    case TREC_STORE_F16:
        hipLaunchKernelGGL(k_sumsq<_Float16>, dim3(blocks), dim3(WG), 0, st, (const _Float16 *)a, n, out);
        break;
    case TREC_STORE_BF16:
        hipLaunchKernelGGL(k_sumsq<__bf16>, dim3(blocks), dim3(WG), 0, st, (const __bf16 *)a, n, out);
        break;
    default: hipLaunchKernelGGL(k_sumsq<float>, dim3(blocks), dim3(WG), 0, st, (const float *)a, n, out); break;
    }
    return hipGetLastError();
}

hipError_t launch_general_forward(const GeneralJob &g, int theta0, int ntheta, bool ones, hipStream_t st)
{
    hipLaunchKernelGGL(k_general_forward, dim3(cdiv((size_t)g.X * g.T, WG), ntheta), dim3(WG), 0, st, g, theta0,
                       ones ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_general_residual(const GeneralJob &g, int theta0, int ntheta, hipStream_t st)
{
    const size_t n = (size_t)ntheta * g.Y * g.X;
    hipLaunchKernelGGL(k_general_residual, dim3(cdiv(n, WG)), dim3(WG), 0, st, g, theta0, n);
    return hipGetLastError();
}

hipError_t launch_general_back(const GeneralJob &g, int mode, int theta0, int ntheta, const float *img,
                               hipStream_t st)
{
    const size_t n = (size_t)g.X * g.T * g.Y;
    hipLaunchKernelGGL(k_general_back, dim3(cdiv(n, WG)), dim3(WG), 0, st, g, mode, theta0, ntheta, img);
    return hipGetLastError();
}

}  // namespace trec
