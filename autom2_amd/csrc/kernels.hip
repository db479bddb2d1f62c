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
// The general path (any pitch) is in general.hip.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstdlib>
#include <type_traits>
#include <utility>
#include <vector>

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

// Job and tilt tables are read-only while a kernel runs.  Reading them
// through the constant address space turns uniform fields into scalar loads:
// they do not wait on the vector-memory counter (so they do not serialise a
// wave's prefetches) and live in SGPRs.
#define TREC_KCONST __attribute__((address_space(4)))
using KJob = const TREC_KCONST JobDesc;
using KTilt = const TREC_KCONST TiltConst;
template <typename T>
__device__ __forceinline__ const TREC_KCONST T *kconst(const T *p)
{
    return (const TREC_KCONST T *)p;
}

// Locates the job that owns work-group `b` of a multi-job launch.
__device__ __forceinline__ int find_job(const LaunchList &L, unsigned b)
{
    int j = 0;
    while (j + 1 < L.n && b >= L.off[j + 1]) ++j;
    return j;
}

// Row-block slices.  A slice holds the RB rows of NC cells (pixels of one
// tilt, or voxels of one z layer) as RB / PR planes; plane q holds rows
// q*PR .. q*PR+PR-1 of every cell as one 16-byte vector per cell, PR = 4
// for float and 8 for 16-bit storage:
//     element (cell c, row r) = slice[((r / PR) * NC + c) * PR + r % PR]
// A wave that reads one plane for 64 neighbouring cells touches one
// contiguous run of memory, so each cache line is looked up once.
template <typename T>
struct Slice {
    static constexpr int PR = 16 / (int)sizeof(T);
    static constexpr int NP = RB / PR;
    __device__ static __forceinline__ size_t index(int NC, int c, int r)
    {
        return ((size_t)(r / PR) * NC + c) * PR + (r % PR);
    }
    __device__ static __forceinline__ void unpack(const uint4 raw, float *o)
    {
        if constexpr (sizeof(T) == 4) {
            o[0] = __uint_as_float(raw.x);
            o[1] = __uint_as_float(raw.y);
            o[2] = __uint_as_float(raw.z);
            o[3] = __uint_as_float(raw.w);
        } else {
            const T *h = reinterpret_cast<const T *>(&raw);
#pragma unroll
            for (int e = 0; e < PR; ++e) o[e] = (float)h[e];
        }
    }
    __device__ static __forceinline__ uint4 pack(const float *v)
    {
        uint4 raw;
        if constexpr (sizeof(T) == 4) {
            raw = make_uint4(__float_as_uint(v[0]), __float_as_uint(v[1]), __float_as_uint(v[2]), __float_as_uint(v[3]));
        } else {
            T *h = reinterpret_cast<T *>(&raw);
#pragma unroll
            for (int e = 0; e < PR; ++e) h[e] = (T)v[e];
        }
        return raw;
    }
    __device__ static __forceinline__ void load(const void *s, int NC, int c, float (&v)[RB])
    {
        const T *b = static_cast<const T *>(s);
#pragma unroll
        for (int q = 0; q < NP; ++q) unpack(*reinterpret_cast<const uint4 *>(b + ((size_t)q * NC + c) * PR), v + q * PR);
    }
    __device__ static __forceinline__ void store(void *s, int NC, int c, const float (&v)[RB])
    {
        T *b = static_cast<T *>(s);
#pragma unroll
        for (int q = 0; q < NP; ++q) *reinterpret_cast<uint4 *>(b + ((size_t)q * NC + c) * PR) = pack(v + q * PR);
    }
};

// Buffer-resource view of a slice, one resource per plane: a cell outside
// [0, NC) (including negative cells) reads as zero, so taps that fall off
// the detector need no clamp or select.  The slice base must be uniform.
constexpr int BUF_DWORD3 = 0x00020000;  // raw buffer, gfx9 family

template <typename T>
struct BufSlice {
    __amdgpu_buffer_rsrc_t r[Slice<T>::NP];
    __device__ __forceinline__ BufSlice(const void *s, int NC)
    {
        const T *b = static_cast<const T *>(s);
#pragma unroll
        for (int q = 0; q < Slice<T>::NP; ++q)
            r[q] = __builtin_amdgcn_make_buffer_rsrc((void *)(b + (size_t)q * NC * Slice<T>::PR), (short)0, NC * 16,
                                                     BUF_DWORD3);
    }
    __device__ __forceinline__ void load(int c, float (&v)[RB]) const
    {
#pragma unroll
        for (int q = 0; q < Slice<T>::NP; ++q) {
            const auto raw = __builtin_amdgcn_raw_buffer_load_b128(r[q], c * 16, 0, 0);
            Slice<T>::unpack(make_uint4(raw[0], raw[1], raw[2], raw[3]), v + q * Slice<T>::PR);
        }
    }
};

// TiltRec-mpi geometry (ValCoef, main-mpi.cpp:204-226) at pitch 0:
// u = ((a0 + a1 * X) + a3 * Z) + ox in double, taps floor(u) and floor(u)+1,
// fraction rounded to float as Weight::x_min_del is.  The compiler fuses the
// products into the sums, so u may differ from the reference's by an ulp; a
// tap index can then change only where u is within an ulp of an integer,
// where both choices put the same weights (to 1e-16) on the same cells.
template <typename TC>
__device__ __forceinline__ void exact_tap(const TC &t, double Xc, double Zc, int &g, float &dx)
{
    double u = __dadd_rn(__dadd_rn(__dadd_rn(t.a0, __dmul_rn(t.a1, Xc)), __dmul_rn(t.a3, Zc)), t.ox);
    double f = floor(u);
    g = (int)f;
    dx = (float)__dsub_rn(u, f);
}

// ------------------------------------------------------------------ layouts

// stack [P][Y][X] float -> proj slices (yb, t), storage type, / scale.
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
        const int y = J.y0 + yb * RB + r, x = x0 + xx;
        float v = 0.f;
        if (y < J.Y && x < J.X) v = stack[((size_t)th * J.Y + y) * J.X + x] * inv;
        tile[r][xx] = v;
    }
    __syncthreads();
    using S = Slice<PT>;
    PT *slice = static_cast<PT *>(J.proj) + ((size_t)yb * J.P + th) * J.X * RB;
    for (int k = threadIdx.x; k < 64 * S::NP; k += WG) {
        const int xx = k % 64, q = k / 64;
        const int x = x0 + xx;
        if (x < J.X) {
            float v[S::PR];
#pragma unroll
            for (int e = 0; e < S::PR; ++e) v[e] = tile[q * S::PR + e][xx];
            *reinterpret_cast<uint4 *>(slice + ((size_t)q * J.X + x) * S::PR) = S::pack(v);
        }
    }
}

// volume (layout Y [Y][T][X] or Z [T][Y][X]) -> vol slices (yb, z), / scale.
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
        const int y = J.y0 + yb * RB + r, x = x0 + xx;
        float v = 0.f;
        if (y < J.Y && x < J.X) {
            const size_t o = layout == TREC_LAYOUT_Y ? ((size_t)y * J.T + z) * J.X + x : ((size_t)z * J.Y + y) * J.X + x;
            v = src[o] * inv;
        }
        tile[r][xx] = v;
    }
    __syncthreads();
    using S = Slice<float>;
    float *slice = J.vol + ((size_t)yb * J.T + z) * J.X * RB;
    for (int k = threadIdx.x; k < 64 * S::NP; k += WG) {
        const int xx = k % 64, q = k / 64;
        const int x = x0 + xx;
        if (x < J.X)
            *reinterpret_cast<float4 *>(slice + ((size_t)q * J.X + x) * 4) =
                make_float4(tile[q * 4][xx], tile[q * 4 + 1][xx], tile[q * 4 + 2][xx], tile[q * 4 + 3][xx]);
    }
}

// vol slices (yb, z) -> volume (layout Y or Z), * scale.
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
    const float *slice = J.vol + ((size_t)yb * J.T + z) * J.X * RB;
    for (int k = threadIdx.x; k < 64 * 4; k += WG) {
        const int xx = k % 64, q = k / 64;
        const int x = x0 + xx;
        if (x < J.X) {
            const float4 v = *reinterpret_cast<const float4 *>(slice + ((size_t)q * J.X + x) * 4);
            tile[q * 4][xx] = v.x;
            tile[q * 4 + 1][xx] = v.y;
            tile[q * 4 + 2][xx] = v.z;
            tile[q * 4 + 3][xx] = v.w;
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < RB * 64; k += WG) {
        const int r = k / 64, xx = k % 64;
        const int y = J.y0 + yb * RB + r, x = x0 + xx;
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

// ------------------------------------------------------------------ half volume copy

typedef _Float16 half2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ unsigned pack_half2(float a, float b)
{
    const half2v h = {(_Float16)a, (_Float16)b};
    return __builtin_bit_cast(unsigned, h);
}

// hvol entry c + 1 (c in [-1, X)) of every quad from vol cells c and c+1
// (zero outside [0, X)).
__global__ __launch_bounds__(WG) void k_shadow(const JobDesc *__restrict__ jd)
{
    const JobDesc &J = *jd;
    const int X = J.X;
    const int ne = X + 1;
    const unsigned nb = (unsigned)(ne + WG - 1) / WG;
    unsigned b = blockIdx.x;
    const int e = (int)(b % nb) * WG + threadIdx.x;
    b /= nb;
    const int z = b % J.T;
    const int yb = b / J.T;
    if (e >= ne) return;
    const int c = e - 1;
    const float *slice = J.vol + ((size_t)yb * J.T + z) * X * RB;
    float v0[RB], v1[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) v0[r] = v1[r] = 0.f;
    if (c >= 0) Slice<float>::load(slice, X, c, v0);
    if (c + 1 < X) Slice<float>::load(slice, X, c + 1, v1);
    uint4 *hs = static_cast<uint4 *>(J.hvol) + ((size_t)yb * J.T + z) * 4 * ne;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        hs[(size_t)q * ne + e] = make_uint4(pack_half2(v0[4 * q], v1[4 * q]), pack_half2(v0[4 * q + 1], v1[4 * q + 1]),
                                            pack_half2(v0[4 * q + 2], v1[4 * q + 2]),
                                            pack_half2(v0[4 * q + 3], v1[4 * q + 3]));
}

// ------------------------------------------------------------------ forward

unsigned forward_blocks(const JobDesc &h, int ntheta)
{
    const unsigned npb = (h.X + FP_PIX - 1) / FP_PIX;
    return npb * (unsigned)ntheta * (unsigned)h.nzc * (unsigned)h.nyb;
}

// Writes a forward-projection result: the ray weights (ONES), the residual
// p - R / Rwt (FUSED), or the z-chunk partial.
template <typename PT, bool FUSED, bool ONES, int NA, typename JD>
__device__ __forceinline__ void fp_epilogue(const JD &J, int th, int zc, int yb, int i, float (&acc)[NA])
{
    const int X = J.X;
    if constexpr (ONES) {
        if (yb == 0) {
            J.inv_rwt[th * X + i] = (acc[0] > J.thr_res) ? 1.0f / acc[0] : 0.f;
            if (J.rwt) J.rwt[th * X + i] = acc[0];
        }
    } else if constexpr (FUSED) {
        const size_t slice = ((size_t)yb * J.P + th) * X * RB;
        const float inv = J.inv_rwt[th * X + i];
        float p[RB];
        Slice<PT>::load(static_cast<const PT *>(J.proj) + slice, X, i, p);
#pragma unroll
        for (int r = 0; r < RB; ++r) acc[r] = fmaf(-acc[r], inv, p[r]);
        Slice<PT>::store(static_cast<PT *>(J.res) + slice, X, i, acc);
    } else {
        const size_t slice = (((size_t)zc * J.nyb + yb) * J.P + th) * X * RB;
        Slice<float>::store(J.part + slice, X, i, acc);
    }
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
    KJob &J = *kconst(L.job[j]);
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
    KTilt &t = kconst(J.tc)[th];
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
        // no voxel of this layer reaches pixel i
        if (xf >= X || xf + ncand <= 0) continue;
        const float d0 = fmaf(-a1f, frac, a1m1);
        if constexpr (ONES) {
            for (int c = 0; c < ncand; ++c) {
                const int x = xf + c;
                const float d = fmaf(a1f, (float)c, d0);
                acc[0] += ((unsigned)x < (unsigned)X) ? fmaxf(0.f, 1.f - fabsf(d)) : 0.f;
            }
        } else {
            const BufSlice<float> vz(vbase + (size_t)z * X * RB, X);
            for (int c = 0; c < ncand; ++c) {
                const float d = fmaf(a1f, (float)c, d0);
                const float w = fmaxf(0.f, 1.f - fabsf(d));
                float v[RB];
                vz.load(xf + c, v);  // voxels outside [0, X) read as 0
#pragma unroll
                for (int r = 0; r < RB; ++r) acc[r] = fmaf(w, v[r], acc[r]);
            }
        }
    }
    fp_epilogue<PT, FUSED, ONES>(J, th, zc, yb, i, acc);
}

// Candidate gathers of the forward projection.  For pixel i of a tilt, the
// voxels of layer z with weight > 0 are cells xf .. xf + ncand - 1 of the
// layer's LDS band (cl = xf - first cell of the band); candidate c sits at
// distance d0 + a1 c from the pixel and weighs max(0, 1 - |d|).  NC is
// ncand as a compile-time count (0 = run-time loop).
template <int NC, int S>
__device__ __forceinline__ void fw_gather(const float4 (*band)[S], int cl, int ncand, float d0, float a1f,
                                          float (&acc)[RB])
{
    const int n = NC ? NC : ncand;
#pragma unroll
    for (int c = 0; c < n; ++c) {
        // candidates are read two at a time: the compiler would otherwise
        // hoist every read of the layer and run out of registers
        if (c > 0 && c % 2 == 0) __builtin_amdgcn_sched_barrier(0);
        const float wt = fmaxf(0.f, 1.f - fabsf(fmaf(a1f, (float)c, d0)));
        // the last candidate has weight 0 for most pixels unless 2 / a1 is
        // near a whole number: those lanes skip its reads (adding 0 * v
        // changes nothing), which relieves the LDS banks
        if (NC && c == n - 1 && !(wt > 0.f)) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 v = band[q][cl + c];
            acc[4 * q] = fmaf(wt, v.x, acc[4 * q]);
            acc[4 * q + 1] = fmaf(wt, v.y, acc[4 * q + 1]);
            acc[4 * q + 2] = fmaf(wt, v.z, acc[4 * q + 2]);
            acc[4 * q + 3] = fmaf(wt, v.w, acc[4 * q + 3]);
        }
    }
}

// The same sums from the half copy of the volume: candidates c and c+1
// share one (c, c+1) pair entry, applied by one v_dot2_f32_f16 per row with
// half weights and float accumulation.  A missing odd candidate gets weight 0.
template <int NC, int S>
__device__ __forceinline__ void fw_gather_h(const float4 (*band)[S], int cl, int ncand, float d0, float a1f,
                                            float (&acc)[RB])
{
    const int n = NC ? NC : ncand;
#pragma unroll
    for (int c = 0; c < n; c += 2) {
        const float wa = fmaxf(0.f, 1.f - fabsf(fmaf(a1f, (float)c, d0)));
        const float wb = c + 1 < n ? fmaxf(0.f, 1.f - fabsf(fmaf(a1f, (float)(c + 1), d0))) : 0.f;
        // as in fw_gather: lanes whose last pair weighs nothing skip it
        if (NC && c + 2 >= n && !(wa > 0.f) && !(wb > 0.f)) continue;
        const half2v w = {(_Float16)wa, (_Float16)wb};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint4 e = *reinterpret_cast<const uint4 *>(&band[q][cl + c]);
            acc[4 * q] = __builtin_amdgcn_fdot2(__builtin_bit_cast(half2v, e.x), w, acc[4 * q], false);
            acc[4 * q + 1] = __builtin_amdgcn_fdot2(__builtin_bit_cast(half2v, e.y), w, acc[4 * q + 1], false);
            acc[4 * q + 2] = __builtin_amdgcn_fdot2(__builtin_bit_cast(half2v, e.z), w, acc[4 * q + 2], false);
            acc[4 * q + 3] = __builtin_amdgcn_fdot2(__builtin_bit_cast(half2v, e.w), w, acc[4 * q + 3], false);
        }
    }
}

// ------------------------------------------------------------------ forward, tilt groups

// Uniform double in scalar registers.
__device__ __forceinline__ double uniform_d(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

unsigned forward_group_blocks(const JobDesc &h, int ntheta)
{
    const unsigned npb = (h.X + FW_PIX - 1) / FW_PIX;
    const unsigned ntg = ((unsigned)ntheta + FG - 1) / FG;
    return npb * ntg * (unsigned)h.nzc * (unsigned)h.nyb;
}

// Forward projection of FG neighbouring tilts per work-group, one wave per
// tilt, for the same 64 pixels, row block and z chunk.  Neighbouring tilts
// see nearly the same voxels of a layer, so the group stages the union of
// their bands in LDS once (loaded by all four waves, 16 cells a time) and
// each wave gathers its own candidates from it; this divides the loads and
// LDS stores per tilt by up to FG.  Unions wider than the LDS band (tilts
// past about 60 degrees far from the tilt axis) are gathered straight from
// L2 by every wave instead.  Per pixel, the sums and their order are those
// of k_forward.  HV (half storage) stages the half copy of
// the volume, hvol, whose pair entries the waves apply with fw_gather_h.
//
// The union of layer z is [ulo, ulo + uw) with ulo the least first cell and
// ulo + uw the greatest end over the group's tilts, each computed with the
// expressions the pixels use, so every lane's candidates lie inside it.
// S is the band's length in cells: FGS (six groups per CU), or FGS_SMALL
// (seven) when the host has found every union of the launch to fit it.
template <typename PT, bool FUSED, bool HV, int S = FGS>
__global__ __launch_bounds__(FW_PIX *FG) __attribute__((amdgpu_waves_per_eu(S <= FGS_SMALL ? 7 : 6))) void k_forward_g(
    const LaunchList L, unsigned nblk)
{
    __shared__ float4 band[4][S];
    const unsigned b0 = xcd_remap(blockIdx.x, nblk);
    const int j = find_job(L, b0);
    KJob &J = *kconst(L.job[j]);
    unsigned lb = b0 - L.off[j];
    const int ntheta = L.ntheta[j] ? L.ntheta[j] : J.P;
    const int ntg = (ntheta + FG - 1) / FG;
    const unsigned npb = (J.X + FW_PIX - 1) / FW_PIX;
    const int pb = lb % npb;
    lb /= npb;
    const int tg = lb % ntg;
    lb /= ntg;
    const int zc = lb % J.nzc;
    const int yb = lb / J.nzc;
    const int th0 = L.theta0[j] + tg * FG;
    const int nt = min(FG, L.theta0[j] + ntheta - th0);  // tilts of this group
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / FW_PIX));
    const int lane = threadIdx.x % FW_PIX;
    const bool mine = wv < nt;  // wave wv owns tilt th0 + wv
    const int th = th0 + min(wv, nt - 1);
    const int i0 = pb * FW_PIX;
    const int i = i0 + lane;
    const int X = J.X;
    const int z0 = zc * J.zc_len;
    const int z1 = min(J.T, z0 + J.zc_len);
    const float *vbase = J.vol + (size_t)yb * J.T * X * RB;

    // band ends of every tilt of the group (tilts past nt repeat the last)
    double sl[FG], slo[FG], shi[FG];
    int nca[FG];
#pragma unroll
    for (int k = 0; k < FG; ++k) {
        KTilt &tk = kconst(J.tc)[th0 + min(k, nt - 1)];
        sl[k] = tk.slope;
        slo[k] = uniform_d(((double)(i0 - 1) - tk.c0) * tk.inv_a1);
        shi[k] = uniform_d(((double)(i0 + FW_PIX - 2) - tk.c0) * tk.inv_a1);
        nca[k] = tk.ncand;
    }
    // Union ends of 64 layers at once: lane l holds layer zb + l.
    auto union_block = [&](int zb, int &lo_v, int &hi_v) {
        const double zl = (double)(zb + lane);
        int lo = 0x7fffffff, hi = -0x7fffffff;
#pragma unroll
        for (int k = 0; k < FG; ++k) {
            lo = min(lo, (int)floor(fma(sl[k], zl, slo[k])) + 1);
            hi = max(hi, (int)floor(fma(sl[k], zl, shi[k])) + 1 + nca[k]);
        }
        lo_v = lo;
        hi_v = hi;
    };
    // widest union of the chunk (identical in every wave of the group)
    int umax = 0;
    for (int zb = z0; zb < z1; zb += FW_PIX) {
        int lo_v, hi_v;
        union_block(zb, lo_v, hi_v);
        int wmax = zb + lane < z1 ? hi_v - lo_v : 0;
#pragma unroll
        for (int o = FW_PIX / 2; o > 0; o >>= 1) wmax = max(wmax, __shfl_xor(wmax, o, FW_PIX));
        umax = max(umax, __builtin_amdgcn_readfirstlane(wmax));
    }

    KTilt &t = kconst(J.tc)[th];
    const double slope = t.slope;
    const double sbase = ((double)(i - 1) - t.c0) * t.inv_a1;
    const float a1f = t.a1f, a1m1 = t.a1f - 1.0f;
    const int ncand = t.ncand;
    float acc[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) acc[r] = 0.f;

    if (umax + FW_G > S) {
        if (mine) {
            for (int z = z0; z < z1; ++z) {
                const double s = fma(slope, (double)z, sbase);
                const double fl = floor(s);
                const float frac = (float)(s - fl);
                const int xf = (int)fl + 1;
                const float d0 = fmaf(-a1f, frac, a1m1);
                const BufSlice<float> vz(vbase + (size_t)z * X * RB, X);
                for (int c = 0; c < ncand; ++c) {
                    const float wt = fmaxf(0.f, 1.f - fabsf(fmaf(a1f, (float)c, d0)));
                    float v[RB];
                    vz.load(xf + c, v);
#pragma unroll
                    for (int r = 0; r < RB; ++r) acc[r] = fmaf(wt, v[r], acc[r]);
                }
            }
        }
    } else {
        // Chunk c of the union (16 cells) is loaded by wave c % FG; lane l
        // carries plane l / 16 of cell 16 c + l % 16.
        constexpr int NLW = (S / FW_G + FG - 1) / FG;
        float4 pre[NLW];
#pragma unroll
        for (int m = 0; m < NLW; ++m) pre[m] = make_float4(0.f, 0.f, 0.f, 0.f);
        const int lq = lane >> 4, lc = lane & 15;
        // a slice is 4 planes (float rows 4q .. 4q+3) or 4 quads of pair
        // entries (half), PS entries of 16 bytes each; cell x is entry x + E0
        const int PS = HV ? X + 1 : X;
        constexpr int E0 = HV ? 1 : 0;
        const unsigned slice_bytes = 4u * (unsigned)PS * 16u;
        const char *bsrc = HV ? static_cast<const char *>(J.hvol) + (size_t)yb * J.T * slice_bytes
                              : reinterpret_cast<const char *>(vbase);
        auto chunks = [](int wl) { return (wl + FW_G - 1) / FW_G; };
        // Lane offsets differ by 1 KB per chunk.  A cell outside [-E0, X) may
        // load a neighbouring plane's data (or zero past the slice), so at
        // the detector's edges the stash stores zero for it instead.
        int x0 = 0;
        auto fetch = [&](int z, int xl, int nch) {
            const auto rs = __builtin_amdgcn_make_buffer_rsrc((void *)(bsrc + (size_t)z * slice_bytes), (short)0,
                                                              (int)slice_bytes, BUF_DWORD3);
            x0 = xl + wv * FW_G + lc;
            const int voff = (lq * PS + x0 + E0) * 16;
#pragma unroll
            for (int m = 0; m < NLW; ++m) {
                if (wv + FG * m < nch) {
                    const auto raw = __builtin_amdgcn_raw_buffer_load_b128(rs, voff + m * (FG * FW_G * 16), 0, 0);
                    pre[m] = make_float4(__uint_as_float(raw[0]), __uint_as_float(raw[1]), __uint_as_float(raw[2]),
                                         __uint_as_float(raw[3]));
                }
            }
        };
        auto stash = [&](int xl, int nch) {
            if (xl < -E0 || xl + nch * FW_G > X) {
#pragma unroll
                for (int m = 0; m < NLW; ++m)
                    if ((unsigned)(x0 + m * (FG * FW_G) + E0) >= (unsigned)(X + E0))
                        pre[m] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
            // one lane address, the chunks at constant offsets from it
            float4 *bp = &band[lq][wv * FW_G + lc];
#pragma unroll
            for (int m = 0; m < NLW; ++m) {
                if (wv + FG * m < nch) bp[m * FG * FW_G] = pre[m];
            }
            // Every path now holds all of pre[] (a load skipped above was
            // skipped when fetching too): saying so keeps the next fetch from
            // waiting on each of its own loads.
#pragma unroll
            for (int m = 0; m < NLW; ++m)
                asm volatile("" ::"v"(pre[m].x), "v"(pre[m].y), "v"(pre[m].z), "v"(pre[m].w));
        };
        // The layer loop, compiled once per candidate count: every variant
        // passes the same barriers, so the waves of a group may run
        // different ones.
        auto layers = [&](auto nc_const) {
            constexpr int NC = decltype(nc_const)::value;
            int zb = z0, lo_v, hi_v;
            union_block(zb, lo_v, hi_v);
            int ulo = __builtin_amdgcn_readfirstlane(lo_v);
            const int uw = __builtin_amdgcn_readfirstlane(hi_v) - ulo;
            fetch(z0, ulo, chunks(uw));
            stash(ulo, chunks(uw));
            __syncthreads();
            for (int z = z0; z < z1; ++z) {
                int ulon = 0, uwn = 0;
                if (z + 1 < z1) {
                    if (z + 1 - zb == FW_PIX) {
                        zb += FW_PIX;
                        union_block(zb, lo_v, hi_v);
                    }
                    ulon = __builtin_amdgcn_readlane(lo_v, z + 1 - zb);
                    uwn = __builtin_amdgcn_readlane(hi_v, z + 1 - zb) - ulon;
                    fetch(z + 1, ulon, chunks(uwn));
                }
                if (mine) {
                    const double s = fma(slope, (double)z, sbase);
                    const double fl = floor(s);
                    const float frac = (float)(s - fl);
                    const int cl = (int)fl + 1 - ulo;
                    const float d0 = fmaf(-a1f, frac, a1m1);
                    if constexpr (HV)
                        fw_gather_h<NC, S>(band, cl, ncand, d0, a1f, acc);
                    else
                        fw_gather<NC, S>(band, cl, ncand, d0, a1f, acc);
                }
#pragma unroll
                for (int r = 0; r < RB; ++r) asm volatile("" : "+v"(acc[r]));
                __syncthreads();  // every wave is done with layer z
                if (z + 1 < z1) stash(ulon, chunks(uwn));
                __syncthreads();  // the band holds layer z + 1
                ulo = ulon;
            }
        };
        switch (ncand) {
        case 2: layers(std::integral_constant<int, 2>{}); break;
        case 3: layers(std::integral_constant<int, 3>{}); break;
        case 4: layers(std::integral_constant<int, 4>{}); break;
        default: layers(std::integral_constant<int, 0>{}); break;
        }
    }
    if (!mine || i >= X) return;
    fp_epilogue<PT, FUSED, false>(J, th, zc, yb, i, acc);
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
    KJob &J = *kconst(L.job[j]);
    const unsigned lb = b0 - L.off[j];
    const int ntheta = L.ntheta[j] ? L.ntheta[j] : J.P;
    const size_t k = (size_t)lb * WG + threadIdx.x;
    const size_t ncell = (size_t)J.nyb * ntheta * J.X;
    if (k >= ncell) return;
    const int i = k % J.X;
    const int th = L.theta0[j] + (int)((k / J.X) % ntheta);
    const int yb = k / ((size_t)J.X * ntheta);
    const size_t slice = ((size_t)yb * J.P + th) * J.X * RB;
    const size_t zstride = (size_t)J.nyb * J.P * J.X * RB;
    float acc[RB];
    Slice<float>::load(J.part + slice, J.X, i, acc);
    for (int zc = 1; zc < J.nzc; ++zc) {
        float v[RB];
        Slice<float>::load(J.part + slice + zc * zstride, J.X, i, v);
#pragma unroll
        for (int r = 0; r < RB; ++r) acc[r] += v[r];
    }
    const float inv = J.inv_rwt[th * J.X + i];
    float p[RB];
    Slice<PT>::load(static_cast<const PT *>(J.proj) + slice, J.X, i, p);
#pragma unroll
    for (int r = 0; r < RB; ++r) acc[r] = fmaf(-acc[r], inv, p[r]);
    Slice<PT>::store(static_cast<PT *>(J.res) + slice, J.X, i, acc);
}

// ------------------------------------------------------------------ back

unsigned back_blocks(const JobDesc &h)
{
    const unsigned bz = (WG / BP_X) * (unsigned)(h.bp_zt ? h.bp_zt : BP_ZT);
    const unsigned nxb = (h.X + BP_X - 1) / BP_X;
    const unsigned nzb = (h.T + bz - 1) / bz;
    return nxb * nzb * (unsigned)h.nyb;
}

// Voxel-driven back-projection of one 64 x BZ voxel tile of one row block,
// BZ = 4 ZT (ZT voxels per thread, z, z + 4, ...).  Per pass, the residual
// cells the tile can touch for G tilts are staged in LDS as float planes
// (half pairs for half storage); every thread then gathers its voxels' two
// taps per tilt from LDS with the exact double geometry.  MODE selects the
// update (BackMode).  For BACK_SIRT and BACK_SART the image is the residual,
// for BACK_BPT the (filtered) stack.  G and ZT are the shapes the autotuner
// chooses among; every job of a launch has the same.

template <typename PT, int MODE, int G, int ZT>
__global__ __launch_bounds__(WG) void k_back(const LaunchList L, unsigned nblk)
{
    constexpr int BZ = (WG / BP_X) * ZT;
    constexpr int BW = BP_X + BZ + 4;  // taps of a tile span at most 64 + BZ cells
    // half storage keeps the residual in half in LDS: entry (t, q, c) holds
    // rows 4q .. 4q+3 of cells c and c+1 as (c, c+1) pairs
    constexpr bool HALF = std::is_same<PT, _Float16>::value;
    __shared__ float4 win[G][4][BW];
    __shared__ int gmin_s[G];
    const unsigned b0 = xcd_remap(blockIdx.x, nblk);
    const int j = find_job(L, b0);
    KJob &J = *kconst(L.job[j]);
    unsigned lb = b0 - L.off[j];
    const unsigned nzb = (J.T + BZ - 1) / BZ;
    const unsigned nxb = (J.X + BP_X - 1) / BP_X;
    const int zb = lb % nzb;
    lb /= nzb;
    const int xb = lb % nxb;
    const int yb = lb / nxb;
    const int X = J.X;
    const int xl = threadIdx.x % BP_X;
    const int x = xb * BP_X + xl;
    const int zt0 = zb * BZ + (int)(threadIdx.x / BP_X);  // thread's voxels: zt0 + 4k
    const double ox = kconst(J.tc)[0].ox, oz = kconst(J.tc)[0].oz;
    const double Xc = (double)((float)x - (float)ox);
    double Zc[ZT];
#pragma unroll
    for (int k = 0; k < ZT; ++k) Zc[k] = (double)((float)(zt0 + 4 * k) - (float)oz);
    const void *img = (MODE == BACK_BPT) ? J.proj : J.res;
    const PT *imgb = static_cast<const PT *>(img) + (size_t)yb * J.P * X * RB;
    const int th0 = L.theta0[j];
    const int th1 = L.ntheta[j] ? th0 + L.ntheta[j] : J.P;

    float val[ZT][RB];
    float cs[ZT];
#pragma unroll
    for (int k = 0; k < ZT; ++k) {
        cs[k] = 0.f;
#pragma unroll
        for (int r = 0; r < RB; ++r) val[k][r] = 0.f;
    }

    for (int tg = th0; tg < th1; tg += G) {
        const int ng = min(G, th1 - tg);
        __syncthreads();  // the previous pass is done with win / gmin_s
        if ((int)threadIdx.x < ng) {
            // lowest tap of the tile: u is monotone in x and z, so a corner
            const double X0 = (double)((float)(xb * BP_X) - (float)ox);
            const double Z0 = (double)((float)(zb * BZ) - (float)oz);
            const double Z1 = (double)((float)(zb * BZ + BZ - 1) - (float)oz);
            int ga, gb;
            float d;
            exact_tap(J.tc[tg + threadIdx.x], X0, Z0, ga, d);  // per-lane tilt: vector loads
            exact_tap(J.tc[tg + threadIdx.x], X0, Z1, gb, d);
            gmin_s[threadIdx.x] = min(ga, gb);
        }
        __syncthreads();
        if constexpr (HALF) {
            // plane p (rows 8p .. 8p+7) of cells g and g+1 -> entries q = 2p
            // and 2p + 1 as (g, g+1) pairs
            for (int v = threadIdx.x; v < ng * 2 * BW; v += WG) {
                const int t = v / (2 * BW);
                const int rem = v % (2 * BW);
                const int pl = rem / BW, c = rem % BW;
                const int g = gmin_s[t] + c;
                const PT *rows = imgb + (size_t)(tg + t) * X * RB + (size_t)pl * X * 8;
                uint4 a = make_uint4(0u, 0u, 0u, 0u), b = a;
                if ((unsigned)g < (unsigned)X) a = *reinterpret_cast<const uint4 *>(rows + (size_t)g * 8);
                if ((unsigned)(g + 1) < (unsigned)X) b = *reinterpret_cast<const uint4 *>(rows + (size_t)(g + 1) * 8);
                *reinterpret_cast<uint4 *>(&win[t][2 * pl][c]) =
                    make_uint4((a.x & 0xffffu) | (b.x << 16), (a.x >> 16) | (b.x & 0xffff0000u),
                               (a.y & 0xffffu) | (b.y << 16), (a.y >> 16) | (b.y & 0xffff0000u));
                *reinterpret_cast<uint4 *>(&win[t][2 * pl + 1][c]) =
                    make_uint4((a.z & 0xffffu) | (b.z << 16), (a.z >> 16) | (b.z & 0xffff0000u),
                               (a.w & 0xffffu) | (b.w << 16), (a.w >> 16) | (b.w & 0xffff0000u));
            }
        }
        if constexpr (!HALF && sizeof(PT) == 2) {
            // bfloat16: plane p (rows 8p .. 8p+7) of cell g -> entries 2p and
            // 2p + 1 as floats
            for (int v = threadIdx.x; v < ng * 2 * BW; v += WG) {
                const int t = v / (2 * BW);
                const int rem = v % (2 * BW);
                const int pl = rem / BW, c = rem % BW;
                const int g = gmin_s[t] + c;
                float e[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) e[k] = 0.f;
                if ((unsigned)g < (unsigned)X)
                    Slice<PT>::unpack(*reinterpret_cast<const uint4 *>(imgb + (size_t)(tg + t) * X * RB +
                                                                       ((size_t)pl * X + g) * 8),
                                      e);
                win[t][2 * pl][c] = make_float4(e[0], e[1], e[2], e[3]);
                win[t][2 * pl + 1][c] = make_float4(e[4], e[5], e[6], e[7]);
            }
        }
        // float: rows 4q .. 4q+3 of cell g
        for (int v = threadIdx.x; sizeof(PT) == 4 && v < ng * 4 * BW; v += WG) {
            const int t = v / (4 * BW);
            const int rem = v % (4 * BW);
            const int q = rem / BW, c = rem % BW;
            const int g = gmin_s[t] + c;
            float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
            if ((unsigned)g < (unsigned)X)
                o = *reinterpret_cast<const float4 *>(imgb + (size_t)(tg + t) * X * RB + ((size_t)q * X + g) * 4);
            win[t][q][c] = o;
        }
        __syncthreads();
        for (int t = 0; t < ng; ++t) {
            KTilt &tc = kconst(J.tc)[tg + t];
            const int gm = gmin_s[t];
#pragma unroll
            for (int k = 0; k < ZT; ++k) {
                int g;
                float dx;
                exact_tap(tc, Xc, Zc[k], g, dx);
                const int c = min(max(g - gm, 0), BW - 2);
                const float w0 = 1.f - dx, w1 = dx;
                if constexpr (HALF) {
                    // both taps of a row in one v_dot2_f32_f16: weights and
                    // residuals in half, products and sums in float
                    // dx arrives as a float rounded from double: the empty asm
                    // keeps the compiler from fusing the two roundings into a
                    // (software) double-to-half conversion
                    float w1f = w1;
                    asm volatile("" : "+v"(w1f));
                    const half2v w = {(_Float16)w0, (_Float16)w1f};
                    float ic = 1.f;
                    bool upd = true;
                    if constexpr (MODE != BACK_SIRT && MODE != BACK_LIN) {
                        float cw = 0.f;
                        cw += ((unsigned)g < (unsigned)X) ? w0 : 0.f;
                        cw += ((unsigned)(g + 1) < (unsigned)X) ? w1 : 0.f;
                        if constexpr (MODE == BACK_SART) cs[k] = cw;
                        upd = cw > J.thr_upd;
                        ic = upd ? 1.0f / cw : 0.f;
                    }
                    if (upd) {
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const uint4 e = *reinterpret_cast<const uint4 *>(&win[t][q][c]);
                            const unsigned ev[4] = {e.x, e.y, e.z, e.w};
#pragma unroll
                            for (int h = 0; h < 4; ++h) {
                                const half2v pr = __builtin_bit_cast(half2v, ev[h]);
                                if constexpr (MODE == BACK_SIRT || MODE == BACK_LIN)
                                    val[k][4 * q + h] = __builtin_amdgcn_fdot2(pr, w, val[k][4 * q + h], false);
                                else
                                    val[k][4 * q + h] =
                                        fmaf(__builtin_amdgcn_fdot2(pr, w, 0.f, false), ic, val[k][4 * q + h]);
                            }
                        }
                    }
                    continue;
                }
                float r0[RB], r1[RB];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float4 a = win[t][q][c], bq = win[t][q][c + 1];
                    r0[4 * q] = a.x; r0[4 * q + 1] = a.y; r0[4 * q + 2] = a.z; r0[4 * q + 3] = a.w;
                    r1[4 * q] = bq.x; r1[4 * q + 1] = bq.y; r1[4 * q + 2] = bq.z; r1[4 * q + 3] = bq.w;
                }
                if constexpr (MODE == BACK_SIRT || MODE == BACK_LIN) {
#pragma unroll
                    for (int r = 0; r < RB; ++r) val[k][r] = fmaf(w1, r1[r], fmaf(w0, r0[r], val[k][r]));
                } else {
                    // c = 0; c += 1 - dx; c += dx over the in-bounds taps (BilinearValue order)
                    float cw = 0.f;
                    cw += ((unsigned)g < (unsigned)X) ? w0 : 0.f;
                    cw += ((unsigned)(g + 1) < (unsigned)X) ? w1 : 0.f;
                    if constexpr (MODE == BACK_SART) cs[k] = cw;
                    if (cw > J.thr_upd) {
                        const float ic = 1.0f / cw;
#pragma unroll
                        for (int r = 0; r < RB; ++r) val[k][r] = fmaf(fmaf(w1, r1[r], w0 * r0[r]), ic, val[k][r]);
                    }
                }
            }
        }
    }

    // a wave holds 64 consecutive x of one z row (xl is the lane)
    const bool live = x < X;
#pragma unroll
    for (int k = 0; k < ZT; ++k) {
        const int z = zt0 + 4 * k;
        if (z >= J.T) continue;  // uniform in the wave
        float *vslice = J.vol + ((size_t)yb * J.T + z) * X * RB;
        float v[RB];
#pragma unroll
        for (int r = 0; r < RB; ++r) v[r] = 0.f;
        if constexpr (MODE == BACK_LIN) {
            // out = a * sum s + b * x, each product rounded on its own
            if (live) {
                Slice<float>::load(vslice, X, x, v);
                float o[RB];
#pragma unroll
                for (int r = 0; r < RB; ++r) {
                    float pa = J.lin_a * val[k][r], pb = J.lin_b * v[r];
                    asm volatile("" : "+v"(pa), "+v"(pb));
                    o[r] = pa + pb;
                }
                Slice<float>::store(J.out + ((size_t)yb * J.T + z) * X * RB, X, x, o);
            }
            continue;
        }
        if (live) {
            Slice<float>::load(vslice, X, x, v);
            bool upd = true;
            if constexpr (MODE == BACK_SIRT) {
                const float f = J.gamma * J.inv_c[z * X + x];
#pragma unroll
                for (int r = 0; r < RB; ++r) v[r] = fmaf(val[k][r], f, v[r]);
            } else if constexpr (MODE == BACK_SART) {
                upd = cs[k] > J.thr_upd;
                if (upd) {
#pragma unroll
                    for (int r = 0; r < RB; ++r) v[r] = fmaf(val[k][r], J.gamma, v[r]);
                }
            } else {
#pragma unroll
                for (int r = 0; r < RB; ++r) v[r] += val[k][r];
            }
            if (upd) Slice<float>::store(vslice, X, x, v);
        }
        if constexpr (HALF && MODE != BACK_BPT && MODE != BACK_LIN) {
            // Keep the half copy in step: lane xl writes pair entry x (cells
            // x and x+1, the latter from lane xl+1); the entry straddling two
            // tiles is written in halves, its first cell by this tile's last
            // lane and its second by the next tile's first lane.
            float vn[RB];
#pragma unroll
            for (int r = 0; r < RB; ++r) vn[r] = __shfl_down(v[r], 1, BP_X);
            const int ne = X + 1;
            char *hs = static_cast<char *>(J.hvol) + ((size_t)yb * J.T + z) * 4 * (size_t)ne * 16;
            if (live && xl < BP_X - 1) {
                const bool next = x + 1 < X;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    *reinterpret_cast<uint4 *>(hs + ((size_t)q * ne + x + 1) * 16) =
                        make_uint4(pack_half2(v[4 * q], next ? vn[4 * q] : 0.f),
                                   pack_half2(v[4 * q + 1], next ? vn[4 * q + 1] : 0.f),
                                   pack_half2(v[4 * q + 2], next ? vn[4 * q + 2] : 0.f),
                                   pack_half2(v[4 * q + 3], next ? vn[4 * q + 3] : 0.f));
            } else if (live) {
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int h = 0; h < 4; ++h)
                        *reinterpret_cast<_Float16 *>(hs + ((size_t)q * ne + x + 1) * 16 + 4 * h) =
                            (_Float16)v[4 * q + h];
            }
            if (live && xl == 0) {
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int h = 0; h < 4; ++h)
                        *reinterpret_cast<_Float16 *>(hs + ((size_t)q * ne + x) * 16 + 4 * h + 2) =
                            (_Float16)v[4 * q + h];
            }
        }
    }
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

// The same convolution ordered by source cell: y[j] = sum_m h[(j - m) mod N]
// e[m].  Every lane of the work-group reads the same e[m] (the 4 rows as a
// float4, one LDS broadcast) and applies it to RW_J outputs j = j0 + 256 k
// with taps from h2, the kernel written out twice so that (j - m) mod N is
// j - m + N without a branch: 16 multiply-adds per broadcast read and 4
// conflict-free tap reads.  LDS: 16 N bytes of rows + 8 N bytes of taps.
constexpr int RW_J = 4;
__global__ __launch_bounds__(WG) void k_rweight_m(float *__restrict__ data, const float *__restrict__ hk, int N,
                                                  int X, long long nrows, long long rstride, int xstride, int rpg,
                                                  long long gstride)
{
    extern __shared__ float smem[];
    float4 *e = reinterpret_cast<float4 *>(smem);  // [N]
    float *h2 = smem + 4 * (size_t)N;               // [2 N]
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
    for (int m = threadIdx.x; m < 2 * N; m += WG) h2[m] = hk[m < N ? m : m - N];
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
    // rows are read into LDS before the barrier above, so writing them in
    // place below races with nothing (each work-group owns its 4 rows)
    for (int jb = 0; jb < X; jb += WG * RW_J) {
        const int j0 = jb + (int)threadIdx.x;
        float4 acc[RW_J];
        int hb[RW_J];  // h2 index of tap (j - 0): j + N, clamped for outputs past X
#pragma unroll
        for (int k = 0; k < RW_J; ++k) {
            acc[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            hb[k] = min(j0 + k * WG, X - 1) + N;
        }
        for (int m = 0; m < N; ++m) {
            const float4 em = e[m];
#pragma unroll
            for (int k = 0; k < RW_J; ++k) {
                const float t = h2[hb[k] - m];
                acc[k].x = fmaf(t, em.x, acc[k].x);
                acc[k].y = fmaf(t, em.y, acc[k].y);
                acc[k].z = fmaf(t, em.z, acc[k].z);
                acc[k].w = fmaf(t, em.w, acc[k].w);
            }
        }
#pragma unroll
        for (int k = 0; k < RW_J; ++k) {
            const int j = j0 + k * WG;
            if (j >= X) continue;
            const float out[4] = {acc[k].x, acc[k].y, acc[k].z, acc[k].w};
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (live[q]) data[base[q] + (long long)j * xstride] = out[q];
        }
    }
}

// FBP spatial filter (ApplyFilterInplace modes 0 / 1, filter_prj.cpp:266-446),
// on four rows per work-group, in place, with the addressing of k_rweight.
// The reference convolves through OpenCV's DFT; the same circular
// convolution is summed here directly:
//   pad[p]  = row[clamp(p - 1, 0, X - 1)]          p in [0, X + 2)
//   ext[j]  = pad[j] (j < X + 2), pad[2 (X+2) - 1 - j] (j < 2 (X+2)), else 0,
//             for j in [0, N), N = the reference's optimal DFT size
//   out[i]  = sum_k h[k] ext[(i + L/2 + 1 - k) mod N],  L = FBP_TAPS
// The taps are uniform (scalar loads); each thread produces two neighbouring
// outputs, which share all but one of their ext reads.
__global__ __launch_bounds__(WG) void k_fbp_filter(float *__restrict__ data, const float *__restrict__ taps, int N,
                                                   int X, long long nrows, long long rstride, int xstride, int rpg,
                                                   long long gstride)
{
    extern __shared__ float smem[];
    float4 *e = reinterpret_cast<float4 *>(smem);  // [N]
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
    const int npad = X + 2;
    for (int j = threadIdx.x; j < N; j += WG) {
        const int p = j < npad ? j : (j < 2 * npad ? 2 * npad - 1 - j : -1);
        float v[4];
        if (p < 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = 0.f;
        } else {
            const int x = min(max(p - 1, 0), X - 1);
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = data[base[q] + (long long)x * xstride];
        }
        e[j] = make_float4(v[0], v[1], v[2], v[3]);
    }
    __syncthreads();
    const TREC_KCONST float *h = kconst(taps);
    for (int i0 = 2 * (int)threadIdx.x; i0 < X; i0 += 2 * WG) {
        // output i0 reads ext[i0 + L/2 + 1 - k], output i0 + 1 the cell one
        // further on, which output i0 read at tap k - 1
        float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f), a1 = a0;
        const int top = i0 + FBP_TAPS / 2 + 1;
        float4 prev = e[top + 1 < N ? top + 1 : top + 1 - N];
        for (int k = 0; k < FBP_TAPS; ++k) {
            int j = top - k;
            j += j < 0 ? N : 0;
            const float4 cur = e[j];
            const float hk = h[k];
            a0.x = fmaf(hk, cur.x, a0.x); a0.y = fmaf(hk, cur.y, a0.y);
            a0.z = fmaf(hk, cur.z, a0.z); a0.w = fmaf(hk, cur.w, a0.w);
            a1.x = fmaf(hk, prev.x, a1.x); a1.y = fmaf(hk, prev.y, a1.y);
            a1.z = fmaf(hk, prev.z, a1.z); a1.w = fmaf(hk, prev.w, a1.w);
            prev = cur;
        }
        const float o0[4] = {a0.x, a0.y, a0.z, a0.w}, o1[4] = {a1.x, a1.y, a1.z, a1.w};
        // every read of this work-group's rows happened before the barrier
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (!live[q]) continue;
            data[base[q] + (long long)i0 * xstride] = o0[q];
            if (i0 + 1 < X) data[base[q] + (long long)(i0 + 1) * xstride] = o1[q];
        }
    }
}

// --f2b byte output (recon_z.cu:253-262, CufloatToByteKernel basic.cu:442-474).
// Partial sums in double of x (pass 0) or of (x - mean)^2 with the
// difference taken in float (pass 1), one per work-group, in a fixed order.
__global__ __launch_bounds__(WG) void k_moment_part(const float *__restrict__ a, size_t n, int pass,
                                                    const float *__restrict__ stats, double *__restrict__ part)
{
    const float mean = pass ? stats[0] : 0.f;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * WG + threadIdx.x; i < n; i += (size_t)gridDim.x * WG) {
        const float v = a[i];
        if (pass) {
            const float d = v - mean;
            acc += (double)d * (double)d;
        } else {
            acc += (double)v;
        }
    }
    __shared__ double red[WG];
    red[threadIdx.x] = acc;
    __syncthreads();
#pragma unroll
    for (int w = WG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// stats[0] = mean (pass 0) or stats[1] = standard deviation (pass 1), float.
__global__ __launch_bounds__(64) void k_moment_finish(const double *__restrict__ part, int nb, size_t n, int pass,
                                                      float *__restrict__ stats)
{
    double acc = 0.0;
    for (int k = threadIdx.x; k < nb; k += 64) acc += part[k];
#pragma unroll
    for (int w = 32; w > 0; w >>= 1) acc += __shfl_down(acc, w, 64);
    if (threadIdx.x == 0) {
        if (pass) stats[1] = (float)sqrt(acc / (double)n);
        else stats[0] = (float)(acc / (double)n);
    }
}

// byte = (unsigned char)(((clamp((v - mean) / sd, -10, 10) + 10) / 20) * 255)
__global__ __launch_bounds__(WG) void k_to_byte(const float *__restrict__ a, size_t n,
                                                const float *__restrict__ stats, unsigned char *__restrict__ out)
{
#pragma clang fp contract(off)
    const float mean = stats[0], sd = stats[1];
    const float scale = 10.0f;
    for (size_t i = (size_t)blockIdx.x * WG + threadIdx.x; i < n; i += (size_t)gridDim.x * WG) {
        float v = (a[i] - mean) / sd;
        v = fmaxf(-scale, fminf(scale, v));
        v = ((v + scale) / (2.0f * scale)) * 255.0f;
        out[i] = (unsigned char)v;
    }
}

// Sum of squares in double: the stack's scale for the 16-bit forms.  Each
// work-group writes its own partial, part[blockIdx.x]; k_set_scale adds the
// partials in a fixed order, so the scale does not depend on the order in
// which work-groups finish.
template <typename T>
__global__ __launch_bounds__(WG) void k_sumsq(const T *__restrict__ a, size_t n, double *__restrict__ part)
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
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// ------------------------------------------------------------------ small helpers

// scale = RMS of the stack from the nb partials of k_sumsq: lane l adds
// partials l, l + 64, ... in order, then a fixed shuffle tree adds the lanes.
__global__ __launch_bounds__(64) void k_set_scale(JobDesc *d, const double *__restrict__ part, int nb, double n)
{
    double acc = 0.0;
    for (int k = threadIdx.x; k < nb; k += 64) acc += part[k];
#pragma unroll
    for (int w = 32; w > 0; w >>= 1) acc += __shfl_down(acc, w, 64);
    if (threadIdx.x == 0) {
        const double rms = sqrt(acc / n);
        d->scale = (rms > 0.0 && rms < 3.0e38) ? (float)rms : 1.0f;
    }
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
    dst[k] = src[((size_t)(y / RB) * P + th) * X * RB + Slice<float>::index(X, x, y % RB)];
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

hipError_t launch_set_scale(JobDesc *d, const double *part, size_t n, hipStream_t st)
{
    const int nb = (int)sumsq_blocks(n);
    hipLaunchKernelGGL(k_set_scale, dim3(1), dim3(64), 0, st, d, part, nb, (double)n);
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

hipError_t launch_shadow(const JobDesc &h, const JobDesc *d, hipStream_t st)
{
    const size_t blocks = cdiv((size_t)h.X + 1, WG) * h.T * h.nyb;
    hipLaunchKernelGGL(k_shadow, dim3((unsigned)blocks), dim3(WG), 0, st, d);
    return hipGetLastError();
}

// Forward-projection kernel for value projections: tilt groups sharing an
// LDS band, or (TREC_FP_DIRECT set) the direct gather, kept as a reference
// path for comparisons.
static bool forward_direct()
{
    static const bool direct = getenv("TREC_FP_DIRECT") != nullptr;
    return direct;
}

unsigned fwd_value_blocks(const JobDesc &h, int ntheta)
{
    return forward_direct() ? forward_blocks(h, ntheta) : forward_group_blocks(h, ntheta);
}

hipError_t launch_forward(const LaunchList &L, unsigned nblk, int storage, bool fused, hipStream_t st, bool small)
{
    // storage < 0 selects the splat of ones (Rwt)
    if (storage < 0) {
        hipLaunchKernelGGL((k_forward<float, false, true>), dim3(nblk), dim3(WG), 0, st, L, nblk);
    } else if (!forward_direct()) {
        // half storage reads the half copy of the volume
        const dim3 b(FW_PIX * FG);
        const bool hv = storage == TREC_STORE_F16;
        if (!fused) {
            if (hv) hipLaunchKernelGGL((k_forward_g<float, false, true>), dim3(nblk), b, 0, st, L, nblk);
            else hipLaunchKernelGGL((k_forward_g<float, false, false>), dim3(nblk), b, 0, st, L, nblk);
        } else {
            switch (storage) {
            case TREC_STORE_F16:
                if (small) hipLaunchKernelGGL((k_forward_g<_Float16, true, true, FGS_SMALL>), dim3(nblk), b, 0, st, L, nblk);
                else hipLaunchKernelGGL((k_forward_g<_Float16, true, true>), dim3(nblk), b, 0, st, L, nblk);
                break;
            case TREC_STORE_BF16:
                if (small) hipLaunchKernelGGL((k_forward_g<__bf16, true, false, FGS_SMALL>), dim3(nblk), b, 0, st, L, nblk);
                else hipLaunchKernelGGL((k_forward_g<__bf16, true, false>), dim3(nblk), b, 0, st, L, nblk);
                break;
            default:
                if (small) hipLaunchKernelGGL((k_forward_g<float, true, false, FGS_SMALL>), dim3(nblk), b, 0, st, L, nblk);
                else hipLaunchKernelGGL((k_forward_g<float, true, false>), dim3(nblk), b, 0, st, L, nblk);
                break;
            }
        }
    } else {
        switch (fused ? storage : -1) {
        case TREC_STORE_F16: hipLaunchKernelGGL((k_forward<_Float16, true, false>), dim3(nblk), dim3(WG), 0, st, L, nblk); break;
        case TREC_STORE_BF16: hipLaunchKernelGGL((k_forward<__bf16, true, false>), dim3(nblk), dim3(WG), 0, st, L, nblk); break;
        case TREC_STORE_F32: hipLaunchKernelGGL((k_forward<float, true, false>), dim3(nblk), dim3(WG), 0, st, L, nblk); break;
        default: hipLaunchKernelGGL((k_forward<float, false, false>), dim3(nblk), dim3(WG), 0, st, L, nblk); break;
        }
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

// The back-projection shapes (tilts per pass, voxels per thread along z)
// instantiated for SIRT, the autotuner's candidates; the first is the
// default.  The other modes use the default only.
const std::vector<std::pair<int, int>> &back_shapes(int storage)
{
    static const std::vector<std::pair<int, int>> f32 = {{BP_G32, BP_ZT}, {4, 4}, {8, 4}, {6, 2}, {8, 2}};
    static const std::vector<std::pair<int, int>> f16 = {{BP_G, BP_ZT}, {6, 4}, {12, 4}, {8, 2}};
    static const std::vector<std::pair<int, int>> bf16 = {{BP_G, BP_ZT}};
    return storage == TREC_STORE_F16 ? f16 : (storage == TREC_STORE_BF16 ? bf16 : f32);
}

template <typename PT, int MODE, int G, int ZT>
static void back_one(const LaunchList &L, unsigned nblk, hipStream_t st)
{
    hipLaunchKernelGGL((k_back<PT, MODE, G, ZT>), dim3(nblk), dim3(WG), 0, st, L, nblk);
}

template <int MODE>
static hipError_t launch_back_mode(const LaunchList &L, unsigned nblk, int storage, hipStream_t st)
{
    switch (storage) {
    case TREC_STORE_F16: back_one<_Float16, MODE, BP_G, BP_ZT>(L, nblk, st); break;
    case TREC_STORE_BF16: back_one<__bf16, MODE, BP_G, BP_ZT>(L, nblk, st); break;
    default: back_one<float, MODE, BP_G32, BP_ZT>(L, nblk, st); break;
    }
    return hipSuccess;
}

static hipError_t launch_back_sirt(const LaunchList &L, unsigned nblk, int storage, int g, int zt, hipStream_t st)
{
    if (storage == TREC_STORE_F16) {
        if (g == BP_G && zt == BP_ZT) back_one<_Float16, BACK_SIRT, BP_G, BP_ZT>(L, nblk, st);
        else if (g == 6 && zt == 4) back_one<_Float16, BACK_SIRT, 6, 4>(L, nblk, st);
        else if (g == 12 && zt == 4) back_one<_Float16, BACK_SIRT, 12, 4>(L, nblk, st);
        else if (g == 8 && zt == 2) back_one<_Float16, BACK_SIRT, 8, 2>(L, nblk, st);
        else return hipErrorInvalidValue;
    } else if (storage == TREC_STORE_BF16) {
        if (g == BP_G && zt == BP_ZT) back_one<__bf16, BACK_SIRT, BP_G, BP_ZT>(L, nblk, st);
        else return hipErrorInvalidValue;
    } else {
        if (g == BP_G32 && zt == BP_ZT) back_one<float, BACK_SIRT, BP_G32, BP_ZT>(L, nblk, st);
        else if (g == 4 && zt == 4) back_one<float, BACK_SIRT, 4, 4>(L, nblk, st);
        else if (g == 8 && zt == 4) back_one<float, BACK_SIRT, 8, 4>(L, nblk, st);
        else if (g == 6 && zt == 2) back_one<float, BACK_SIRT, 6, 2>(L, nblk, st);
        else if (g == 8 && zt == 2) back_one<float, BACK_SIRT, 8, 2>(L, nblk, st);
        else return hipErrorInvalidValue;
    }
    return hipSuccess;
}

hipError_t launch_back(const LaunchList &L, unsigned nblk, int storage, int mode, hipStream_t st, int g, int zt)
{
    const auto &def = back_shapes(storage)[0];
    if (g == 0) g = def.first;
    if (zt == 0) zt = def.second;
    hipError_t e = hipSuccess;
    if (mode == BACK_SIRT) {
        e = launch_back_sirt(L, nblk, storage, g, zt, st);
    } else {
        if (g != def.first || zt != def.second) return hipErrorInvalidValue;
        switch (mode) {
        case BACK_BPT: e = launch_back_mode<BACK_BPT>(L, nblk, storage, st); break;
        case BACK_LIN: back_one<float, BACK_LIN, BP_G32, BP_ZT>(L, nblk, st); break;
        default: e = launch_back_mode<BACK_SART>(L, nblk, storage, st); break;
        }
    }
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t launch_rweight(float *data, const float *hk, int N, int X, long long nrows, long long row_stride,
                          int x_stride, int rows_per_group, long long group_stride, hipStream_t st)
{
    const unsigned blocks = (unsigned)((nrows + 3) / 4);
    const size_t shm = 16 * (size_t)N + 8 * (size_t)N;
    if (shm <= 160 * 1024) {
        if (shm > 64 * 1024) {
            static hipError_t attr = hipFuncSetAttribute((const void *)k_rweight_m,
                                                         hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            if (attr != hipSuccess) return attr;
        }
        hipLaunchKernelGGL(k_rweight_m, dim3(blocks), dim3(WG), shm, st, data, hk, N, X, nrows, row_stride, x_stride,
                           rows_per_group, group_stride);
        return hipGetLastError();
    }
    // rows of about 6200 pixels and more: the tap-ordered kernel needs less LDS
    const size_t shmem = sizeof(float) * (size_t)((N + 3) & ~3) + sizeof(float4) * (size_t)N;
    if (shmem > 160 * 1024) return hipErrorInvalidValue;  // rows beyond X ~ 7400 do not fit one CU's LDS
    if (shmem > 64 * 1024) {
        static hipError_t attr = hipFuncSetAttribute((const void *)k_rweight,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (attr != hipSuccess) return attr;
    }
    hipLaunchKernelGGL(k_rweight, dim3(blocks), dim3(WG), shmem, st, data, hk, N, X, nrows, row_stride, x_stride,
                       rows_per_group, group_stride);
    return hipGetLastError();
}

hipError_t launch_fbp_filter(float *data, const float *taps, int N, int X, long long nrows, long long row_stride,
                             int x_stride, int rows_per_group, long long group_stride, hipStream_t st)
{
    const size_t shmem = sizeof(float4) * (size_t)N;
    if (shmem > 160 * 1024) return hipErrorInvalidValue;  // rows beyond X ~ 9900 do not fit one CU's LDS
    if (shmem > 64 * 1024) {
        static hipError_t attr = hipFuncSetAttribute((const void *)k_fbp_filter,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (attr != hipSuccess) return attr;
    }
    const unsigned blocks = (unsigned)((nrows + 3) / 4);
    hipLaunchKernelGGL(k_fbp_filter, dim3(blocks), dim3(WG), shmem, st, data, taps, N, X, nrows, row_stride,
                       x_stride, rows_per_group, group_stride);
    return hipGetLastError();
}

// Element conversion for MRC data (trec_convert): every type to float
// exactly (bytes over 255 as mrctools' reader divides them), float to half
// rounded to nearest even.  Grid-stride over n elements.
template <typename Tin, typename Tout, bool DIV255>
__global__ __launch_bounds__(WG) void k_convert(const Tin *__restrict__ in, Tout *__restrict__ out, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * WG + threadIdx.x; i < n; i += (size_t)gridDim.x * WG) {
        float v = (float)in[i];
        if (DIV255) v = v / 255.0f;
        out[i] = (Tout)v;
    }
}

template <typename Tout>
static hipError_t convert_to(const void *in, int in_type, Tout *out, size_t n, hipStream_t st)
{
    const unsigned blocks = (unsigned)std::min<size_t>(std::max<size_t>(cdiv(n, WG), 1), 16384);
    switch (in_type) {
    case TREC_DT_I8:
        hipLaunchKernelGGL((k_convert<signed char, Tout, false>), dim3(blocks), dim3(WG), 0, st,
                           (const signed char *)in, out, n);
        break;
    case TREC_DT_U8:
        hipLaunchKernelGGL((k_convert<unsigned char, Tout, false>), dim3(blocks), dim3(WG), 0, st,
                           (const unsigned char *)in, out, n);
        break;
    case TREC_DT_U8_DIV255:
        hipLaunchKernelGGL((k_convert<unsigned char, Tout, true>), dim3(blocks), dim3(WG), 0, st,
                           (const unsigned char *)in, out, n);
        break;
    case TREC_DT_I16:
        hipLaunchKernelGGL((k_convert<short, Tout, false>), dim3(blocks), dim3(WG), 0, st, (const short *)in, out, n);
        break;
    case TREC_DT_U16:
        hipLaunchKernelGGL((k_convert<unsigned short, Tout, false>), dim3(blocks), dim3(WG), 0, st,
                           (const unsigned short *)in, out, n);
        break;
    case TREC_DT_F16:
        hipLaunchKernelGGL((k_convert<_Float16, Tout, false>), dim3(blocks), dim3(WG), 0, st, (const _Float16 *)in,
                           out, n);
        break;
    default:
        hipLaunchKernelGGL((k_convert<float, Tout, false>), dim3(blocks), dim3(WG), 0, st, (const float *)in, out, n);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_convert(const void *in, int in_type, void *out, int out_type, size_t n, hipStream_t st)
{
    if (out_type == TREC_DT_F16) return convert_to(in, in_type, (_Float16 *)out, n, st);
    return convert_to(in, in_type, (float *)out, n, st);
}

hipError_t launch_to_byte(const float *a, size_t n, double *part, float *stats, unsigned char *out, hipStream_t st)
{
    const int nb = (int)std::min<size_t>(std::max<size_t>(cdiv(n, WG), 1), 2048);
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_moment_part, dim3(nb), dim3(WG), 0, st, a, n, pass, stats, part);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_moment_finish, dim3(1), dim3(64), 0, st, part, nb, n, pass, stats);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const unsigned blocks = (unsigned)std::min<size_t>(std::max<size_t>(cdiv(n, WG), 1), 8192);
    hipLaunchKernelGGL(k_to_byte, dim3(blocks), dim3(WG), 0, st, a, n, stats, out);
    return hipGetLastError();
}

unsigned sumsq_blocks(size_t n) { return (unsigned)std::min<size_t>(std::max<size_t>(cdiv(n, WG), 1), SUMSQ_BLOCKS); }

hipError_t launch_sumsq(const void *a, int storage, size_t n, double *part, hipStream_t st)
{
    const unsigned blocks = sumsq_blocks(n);
    switch (storage) {
    case TREC_STORE_F16:
        hipLaunchKernelGGL(k_sumsq<_Float16>, dim3(blocks), dim3(WG), 0, st, (const _Float16 *)a, n, part);
        break;
    case TREC_STORE_BF16:
        hipLaunchKernelGGL(k_sumsq<__bf16>, dim3(blocks), dim3(WG), 0, st, (const __bf16 *)a, n, part);
        break;
    default: hipLaunchKernelGGL(k_sumsq<float>, dim3(blocks), dim3(WG), 0, st, (const float *)a, n, part); break;
    }
    return hipGetLastError();
}

}  // namespace trec
