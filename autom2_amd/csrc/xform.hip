// xform.hip -- the .xf application of mrcstack on gfx950: every section of a
// tilt stack warped by its own 2x3 affine map, bicubic or Lanczos-4, with
// the section's mean outside (src/mrctools/src/mrcstack/main.cpp:40-80,
// cv::warpAffine with BORDER_CONSTANT).
//
// The arithmetic is OpenCV's warpAffine for one float channel, bit for bit:
//   - the source coordinate of output pixel (x, y) is fixed point: a row
//     term round((m1 y + m2) 1024) + 16 and a column term round(m0 x 1024),
//     each rounded to nearest even in double; their sum >> 5 gives 1/32
//     pixel steps, whose integer part (saturated to short) is the tap origin
//     and whose fraction indexes the 32-entry 1-D kernel tables;
//   - a tap weight is the float product of the row and column kernels;
//   - a pixel whose K x K taps lie inside the section is the row-by-row
//     float sum of tap * weight; one whose taps all lie outside is the
//     border value; the rest are border + sum of (tap - border) * weight
//     over the taps inside.
// No operation may be contracted into an FMA: every kernel here switches
// contraction off.
//
// Memory plan (HBM-bound: 4 bytes read and 4 written per pixel):
//   k_xf_terms    row and column terms of every section, [P][H] and [P][W]
//                 int2, as warpAffine's own per-row / per-column tables;
//   k_xf_sum      per-section partial sums in double (fixed order), and
//   k_xf_mean     the border value (float)(sum * (1 / n));
//   k_xf_warp     32 x 32 output pixels per one-wave work-group.  The source window
//                 of a tile is exact from its corners (both terms are
//                 monotone in x and y), is staged once in LDS with
//                 coalesced loads, and every tap is then an LDS read.  A
//                 window beyond the LDS budget (magnification above about 2
//                 along the axes, 1.2 at 45 degrees) reads its taps from
//                 global memory instead.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>

#include "trec_internal.h"

namespace trec {

#define TREC_KCONST __attribute__((address_space(4)))

namespace {

template <typename T>
__device__ __forceinline__ const TREC_KCONST T *kc(const T *p)
{
    return (const TREC_KCONST T *)p;
}

// cvRound of a double: nearest, ties to even; values outside int give
// INT_MIN, as x86-64's cvtsd2si does.
__device__ __forceinline__ int cv_round(double v)
{
    const double r = __builtin_rint(v);
    return (r >= -2147483648.0 && r <= 2147483647.0) ? (int)r : INT_MIN;
}

__device__ __forceinline__ int sat_short(int v) { return min(max(v, -32768), 32767); }

// An int2 of a read-only table through the scalar path when the address is
// uniform.
__device__ __forceinline__ int2 ld_term(const int2 *p)
{
    const TREC_KCONST int *q = kc((const int *)p);
    return make_int2(q[0], q[1]);
}

}  // namespace

__global__ __launch_bounds__(WG) void k_xf_terms(const XfConst *__restrict__ xc, int W, int H,
                                                 int2 *__restrict__ rows, int2 *__restrict__ cols)
{
#pragma clang fp contract(off)
    const int i = blockIdx.y;
    const int t = blockIdx.x * WG + threadIdx.x;
    const TREC_KCONST XfConst &c = *kc(&xc[i]);
    const double m0 = c.m[0], m1 = c.m[1], m2 = c.m[2], m3 = c.m[3], m4 = c.m[4], m5 = c.m[5];
    if (t < H) {
        const double y = t;
        rows[(size_t)i * H + t] = make_int2(cv_round((m1 * y + m2) * XF_AB_SCALE) + XF_ROUND_DELTA,
                                            cv_round((m4 * y + m5) * XF_AB_SCALE) + XF_ROUND_DELTA);
    }
    if (t < W) {
        const double x = t;
        cols[(size_t)i * W + t] = make_int2(cv_round(m0 * x * XF_AB_SCALE), cv_round(m3 * x * XF_AB_SCALE));
    }
}

// Partial sums of section i: block b sums elements [b * chunk, (b + 1) * chunk)
// in double: four running sums per thread (one per float of a 16-byte load
// when the section is 16-byte aligned, else element k to sum k % 4), then a
// fixed tree over threads.  The order is fixed, so the sum is reproducible.
__global__ __launch_bounds__(WG) void k_xf_sum(const float *__restrict__ src, size_t n, size_t chunk, int nb,
                                               int vec, double *__restrict__ part)
{
    const int i = blockIdx.y, b = blockIdx.x;
    const float *s = src + (size_t)i * n;
    const size_t lo = (size_t)b * chunk, hi = min(n, lo + chunk);
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    if (vec) {
        const float4 *s4 = reinterpret_cast<const float4 *>(s + lo);
        const size_t m = hi > lo ? (hi - lo) / 4 : 0;
#pragma unroll 4
        for (size_t k = threadIdx.x; k < m; k += WG) {
            const float4 v = s4[k];
            a0 += (double)v.x;
            a1 += (double)v.y;
            a2 += (double)v.z;
            a3 += (double)v.w;
        }
        for (size_t k = lo + 4 * m + threadIdx.x; k < hi; k += WG) a0 += (double)s[k];
    } else {
#pragma unroll 4
        for (size_t k = lo + threadIdx.x; k < hi; k += WG) {
            const double v = s[k];
            switch ((k - lo) / WG & 3) {
            case 0: a0 += v; break;
            case 1: a1 += v; break;
            case 2: a2 += v; break;
            default: a3 += v; break;
            }
        }
    }
    __shared__ double red[WG];
    red[threadIdx.x] = (a0 + a1) + (a2 + a3);
    __syncthreads();
#pragma unroll
    for (int w = WG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)i * nb + b] = red[0];
}

// Border value of each section from its nb partial sums (one wave each).
__global__ __launch_bounds__(64) void k_xf_mean(const double *__restrict__ part, int nb, size_t n,
                                                float *__restrict__ cval)
{
    const int i = blockIdx.x, lane = threadIdx.x;
    double acc = 0;
    for (int k = lane; k < nb; k += 64) acc += part[(size_t)i * nb + k];
#pragma unroll
    for (int w = 32; w > 0; w >>= 1) acc += __shfl_down(acc, w, 64);
    if (lane == 0) cval[i] = (float)(acc * (n ? 1. / (double)n : 0.));
}

// ------------------------------------------------------------------ warp

namespace {

// One output pixel.  Taps come from the staged window (LDS) or from global
// memory; `sbase` is the window offset of the pixel's first tap (LDS) or
// its section offset (global), `stride` the row pitch of that source.
template <int K>
__device__ __forceinline__ float warp_pixel(const float *__restrict__ srcp, int sbase, int stride, int sx, int sy,
                                            int W, int H, const float *ty, const float *tx, float cval)
{
#pragma clang fp contract(off)
    float w[K][K];
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
        for (int b = 0; b < K; ++b) w[a][b] = ty[a] * tx[b];
    constexpr int width_pad = K - 1;
    const bool inside = (unsigned)sx < (unsigned)max(W - width_pad, 0) && (unsigned)sy < (unsigned)max(H - width_pad, 0);
    if (inside) {
        float row[K];
        float sum;
        if constexpr (K == 4) {
#pragma unroll
            for (int a = 0; a < 4; ++a) {
#pragma unroll
                for (int b = 0; b < 4; ++b) row[b] = srcp[sbase + a * stride + b];
                const float r = row[0] * w[a][0] + row[1] * w[a][1] + row[2] * w[a][2] + row[3] * w[a][3];
                sum = a == 0 ? r : sum + r;
            }
        } else {
            sum = 0.f;
#pragma unroll
            for (int a = 0; a < 8; ++a) {
#pragma unroll
                for (int b = 0; b < 8; ++b) row[b] = srcp[sbase + a * stride + b];
                sum += row[0] * w[a][0] + row[1] * w[a][1] + row[2] * w[a][2] + row[3] * w[a][3] +
                       row[4] * w[a][4] + row[5] * w[a][5] + row[6] * w[a][6] + row[7] * w[a][7];
            }
        }
        return sum;
    }
    if (sx >= W || sx + K <= 0 || sy >= H || sy + K <= 0) return cval;
    float sum = cval;
#pragma unroll
    for (int a = 0; a < K; ++a) {
        if ((unsigned)(sy + a) >= (unsigned)H) continue;
#pragma unroll
        for (int b = 0; b < K; ++b) {
            if ((unsigned)(sx + b) >= (unsigned)W) continue;
            sum += (srcp[sbase + a * stride + b] - cval) * w[a][b];
        }
    }
    return sum;
}

}  // namespace

// Grid: x = tile (tiles_x * tiles_y), y = section.  Wave v owns rows
// y0 + v, y0 + v + 4, ... of the tile; lane l owns column x0 + l.
constexpr int XF_SB = 12;                  // window cells per lane per staging round
constexpr int XF_BUF_DWORD3 = 0x00020000;  // raw buffer, gfx9 family

// The pixel loop of one tile: lane l owns column x0 + (l & 31) of rows
// y0 + 2 j + (l >> 5).  Row terms are uniform per half-wave (two scalar
// loads per row pair), the column term is loaded once.
template <int K, bool STAGED>
__device__ __forceinline__ void warp_rows(const float *__restrict__ taps, int stride, int rx0, int ry0,
                                          const float *__restrict__ tab, const int2 *rw, int2 cx, int x, int ty0,
                                          int y1, int W, int H, float cval, float *__restrict__ d)
{
    constexpr int o = K / 2 - 1;
    const int rh = (threadIdx.x >> 5) & 1;
    for (int j = 0; j < XF_T / 2; ++j) {
        const int ya = ty0 + 2 * j;
        if (ya > y1) break;
        const int2 r0 = ld_term(&rw[ya]), r1 = ld_term(&rw[min(ya + 1, y1)]);
        const int y = ya + rh;
        if (x >= W || y > y1) continue;
        const int2 ry = rh ? r1 : r0;
        // two's-complement sums, as warpAffine's int arithmetic wraps
        const int X = (int)((unsigned)ry.x + (unsigned)cx.x) >> 5;
        const int Y = (int)((unsigned)ry.y + (unsigned)cx.y) >> 5;
        const int ix = sat_short(X >> 5), iy = sat_short(Y >> 5);
        const float *tyk = &tab[(Y & (XF_TAB - 1)) * K];
        const float *txk = &tab[(X & (XF_TAB - 1)) * K];
        float ty[K], txv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            ty[k] = tyk[k];
            txv[k] = txk[k];
        }
        const int sx = ix - o, sy = iy - o;
        float v;
        if constexpr (STAGED)
            v = warp_pixel<K>(taps, (sy - ry0) * stride + (sx - rx0), stride, sx, sy, W, H, ty, txv, cval);
        else
            // wraps only for pixels whose taps all lie outside, which read nothing
            v = warp_pixel<K>(taps, (int)((unsigned)sy * (unsigned)W + (unsigned)sx), stride, sx, sy, W, H, ty,
                              txv, cval);
        d[(size_t)y * W + x] = v;
    }
}

// Grid: x = tile (tiles_x * tiles_y of XF_T x XF_T pixels), y = section; one
// wave per work-group.  win_cap is the dynamic LDS window in floats.
template <int K>
__global__ __launch_bounds__(64) void k_xf_warp(const float *__restrict__ src, float *__restrict__ dst, int W, int H,
                                                const int2 *__restrict__ rows, const int2 *__restrict__ cols,
                                                const float *__restrict__ cvals, const float *__restrict__ tab1,
                                                int tiles_x, int win_cap)
{
    constexpr int o = K / 2 - 1;
    __shared__ __align__(16) float tab[XF_TAB * K];
    extern __shared__ float win[];
    const int i = blockIdx.y;
    const int tx0 = (blockIdx.x % tiles_x) * XF_T, ty0 = (blockIdx.x / tiles_x) * XF_T;
    const int x1 = min(tx0 + XF_T, W) - 1, y1 = min(ty0 + XF_T, H) - 1;
    const int lane = threadIdx.x;
    const size_t sec = (size_t)W * H;
    const float *s = src + (size_t)i * sec;
    const int2 *rw = rows + (size_t)i * H;
    const int2 *cl = cols + (size_t)i * W;

    for (int k = lane; k < XF_TAB * K; k += 64) tab[k] = tab1[k];

    // Source window of the tile from its corners: the row term is monotone
    // in y and the column term in x, so the extremes sit at the corners.
    const int2 ra = ld_term(&rw[ty0]), rb = ld_term(&rw[y1]), ca = ld_term(&cl[tx0]), cb = ld_term(&cl[x1]);
    const int ixlo = sat_short((min(ra.x, rb.x) + min(ca.x, cb.x)) >> 10);
    const int ixhi = sat_short((max(ra.x, rb.x) + max(ca.x, cb.x)) >> 10);
    const int iylo = sat_short((min(ra.y, rb.y) + min(ca.y, cb.y)) >> 10);
    const int iyhi = sat_short((max(ra.y, rb.y) + max(ca.y, cb.y)) >> 10);
    const int rx0 = ixlo - o, ry0 = iylo - o;
    const int RW = ixhi - ixlo + K, RH = iyhi - iylo + K;
    const bool staged = RW <= win_cap && RH <= win_cap && RW * RH <= win_cap;
    if (staged) {
        // The window in linear order, XF_SB cells per lane per round: all
        // loads of a round are in flight together.  Cell idx = lane + 64 m
        // sits at row r, column c, stepped without division.  Cells outside
        // the section read as 0 through the buffer's range check.
        const int n = RW * RH;
        const int q64 = 64 / RW, r64 = 64 % RW;
        int r = lane / RW, c = lane % RW;
        const auto rs = __builtin_amdgcn_make_buffer_rsrc((void *)s, (short)0, (int)(sec * 4), XF_BUF_DWORD3);
        for (int base = 0; base < n; base += 64 * XF_SB) {
            float v[XF_SB];
#pragma unroll
            for (int u = 0; u < XF_SB; ++u) {
                const int gy = ry0 + r, gx = rx0 + c;
                const bool in = base + u * 64 + lane < n && (unsigned)gy < (unsigned)H && (unsigned)gx < (unsigned)W;
                const unsigned off = in ? 4u * ((unsigned)gy * (unsigned)W + (unsigned)gx) : 0x80000000u;
                v[u] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
                r += q64;
                c += r64;
                if (c >= RW) {
                    c -= RW;
                    ++r;
                }
            }
#pragma unroll
            for (int u = 0; u < XF_SB; ++u) {
                const int idx = base + u * 64 + lane;
                if (idx < n) win[idx] = v[u];
            }
        }
    }
    __syncthreads();

    const float cval = *kc(&cvals[i]);
    const int x = tx0 + (lane & 31);
    const int2 cx = cl[min(x, x1)];
    float *d = dst + (size_t)i * sec;
    if (staged)
        warp_rows<K, true>(win, RW, rx0, ry0, tab, rw, cx, x, ty0, y1, W, H, cval, d);
    else
        warp_rows<K, false>(s, W, 0, 0, tab, rw, cx, x, ty0, y1, W, H, cval, d);
}

// ------------------------------------------------------------------ launchers

hipError_t launch_xf_terms(const XfLaunch &a, hipStream_t st)
{
    const unsigned gx = (unsigned)((max(a.W, a.H) + WG - 1) / WG);
    hipLaunchKernelGGL(k_xf_terms, dim3(gx, a.P), dim3(WG), 0, st, a.xc, a.W, a.H, a.rows, a.cols);
    return hipGetLastError();
}

hipError_t launch_xf_mean(const XfLaunch &a, hipStream_t st)
{
    const size_t n = (size_t)a.W * a.H;
    const size_t chunk = ((n + a.nb - 1) / a.nb + 3) & ~(size_t)3;
    const int vec = n % 4 == 0 && reinterpret_cast<uintptr_t>(a.src) % 16 == 0;
    hipLaunchKernelGGL(k_xf_sum, dim3(a.nb, a.P), dim3(WG), 0, st, a.src, n, chunk, a.nb, vec, a.part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_xf_mean, dim3(a.P), dim3(64), 0, st, a.part, a.nb, n, a.cval);
    return hipGetLastError();
}

hipError_t launch_xf_warp(const XfLaunch &a, hipStream_t st)
{
    const int tiles_x = (a.W + XF_T - 1) / XF_T, tiles_y = (a.H + XF_T - 1) / XF_T;
    const dim3 grid((unsigned)(tiles_x * tiles_y), (unsigned)a.P);
    const size_t lds = sizeof(float) * (size_t)a.win_cap;
    if (a.lanczos)
        hipLaunchKernelGGL(k_xf_warp<8>, grid, dim3(64), lds, st, a.src, a.dst, a.W, a.H, a.rows, a.cols, a.cval,
                           a.tab, tiles_x, a.win_cap);
    else
        hipLaunchKernelGGL(k_xf_warp<4>, grid, dim3(64), lds, st, a.src, a.dst, a.W, a.H, a.rows, a.cols, a.cval,
                           a.tab, tiles_x, a.win_cap);
    return hipGetLastError();
}

}  // namespace trec
