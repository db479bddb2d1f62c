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
//   k_xf_warp     64 x 16 output pixels per one-wave work-group.  The source
//                 window of a tile is exact from its corners (both terms
//                 are monotone in x and y).  Inside the section it is loaded
//                 by the buffer unit straight into LDS with a fixed row
//                 pitch, so every tap of a pixel is an immediate offset from
//                 one LDS address; at the section's edges it is staged
//                 through registers; a window beyond the LDS budget reads
//                 its taps from global memory.
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

// Weights of one pixel: the float products of the row and column kernels.
template <int K>
__device__ __forceinline__ void tap_weights(const float *ty, const float *tx, float (&w)[K][K])
{
#pragma clang fp contract(off)
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
        for (int b = 0; b < K; ++b) w[a][b] = ty[a] * tx[b];
}

// A pixel whose K x K taps all lie inside the section: the row-by-row float
// sums of tap * weight.  `base` indexes the first tap in `srcp`, `stride`
// is the row pitch.
template <int K>
__device__ __forceinline__ float interior_sum(const float *__restrict__ srcp, int base, int stride,
                                              const float (&w)[K][K])
{
#pragma clang fp contract(off)
    float row[K];
    float sum;
    if constexpr (K == 4) {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
#pragma unroll
            for (int b = 0; b < 4; ++b) row[b] = srcp[base + a * stride + b];
            const float r = row[0] * w[a][0] + row[1] * w[a][1] + row[2] * w[a][2] + row[3] * w[a][3];
            sum = a == 0 ? r : sum + r;
        }
    } else {
        sum = 0.f;
#pragma unroll
        for (int a = 0; a < 8; ++a) {
#pragma unroll
            for (int b = 0; b < 8; ++b) row[b] = srcp[base + a * stride + b];
            sum += row[0] * w[a][0] + row[1] * w[a][1] + row[2] * w[a][2] + row[3] * w[a][3] + row[4] * w[a][4] +
                   row[5] * w[a][5] + row[6] * w[a][6] + row[7] * w[a][7];
        }
    }
    return sum;
}

// Any pixel: interior, entirely outside (the border value), or touching the
// border (border + sum of (tap - border) * weight over the taps inside).
// `base` indexes the pixel's first tap, inside the section or not.
template <int K>
__device__ __forceinline__ float warp_pixel(const float *__restrict__ srcp, int base, int stride, int sx, int sy,
                                            int W, int H, const float *ty, const float *tx, float cval)
{
#pragma clang fp contract(off)
    float w[K][K];
    tap_weights<K>(ty, tx, w);
    const bool inside = (unsigned)sx < (unsigned)max(W - (K - 1), 0) && (unsigned)sy < (unsigned)max(H - (K - 1), 0);
    if (inside) return interior_sum<K>(srcp, base, stride, w);
    if (sx >= W || sx + K <= 0 || sy >= H || sy + K <= 0) return cval;
    float sum = cval;
#pragma unroll
    for (int a = 0; a < K; ++a) {
        if ((unsigned)(sy + a) >= (unsigned)H) continue;
#pragma unroll
        for (int b = 0; b < K; ++b) {
            if ((unsigned)(sx + b) >= (unsigned)W) continue;
            sum += (srcp[base + a * stride + b] - cval) * w[a][b];
        }
    }
    return sum;
}

}  // namespace

constexpr int XF_SB = 12;                  // window cells per lane per staging round
constexpr int XF_BUF_DWORD3 = 0x00020000;  // raw buffer, gfx9 family

// How a tile's window reaches its taps.  XF_BY_DMA: the window lies inside
// the section (so every pixel is interior) and is loaded straight into LDS
// with a fixed row pitch.  XF_BY_LOADS: loaded through registers with its
// own pitch, pixels classified one by one.  XF_BY_GLOBAL: taps read from
// global memory (windows beyond the LDS budget).
enum { XF_BY_DMA = 0, XF_BY_LOADS = 1, XF_BY_GLOBAL = 2 };

// One tile: section, pixels, source window and path.  All fields uniform.
struct XfTile {
    int i, tx0, ty0, x1, y1;
    int rx0, ry0, RW, RH;
    int how;
};

// S is the LDS row pitch of DMA windows (floats, a multiple of 4).
template <int K, int S>
__device__ __forceinline__ XfTile xf_tile(unsigned t_all, unsigned ntiles, int tiles_x, int W, int H,
                                          const int2 *__restrict__ rows, const int2 *__restrict__ cols, int win_cap)
{
    constexpr int o = K / 2 - 1;
    XfTile T;
    T.i = (int)(t_all / ntiles);
    const int t = (int)(t_all % ntiles);
    T.tx0 = (t % tiles_x) * XF_TW;
    T.ty0 = (t / tiles_x) * XF_TH;
    T.x1 = min(T.tx0 + XF_TW, W) - 1;
    T.y1 = min(T.ty0 + XF_TH, H) - 1;
    const int2 *rw = rows + (size_t)T.i * H;
    const int2 *cl = cols + (size_t)T.i * W;
    // The row term is monotone in y and the column term in x, so the
    // extremes of the source coordinate sit at the tile's corners.
    const int2 ra = ld_term(&rw[T.ty0]), rb = ld_term(&rw[T.y1]), ca = ld_term(&cl[T.tx0]), cb = ld_term(&cl[T.x1]);
    const int ixlo = sat_short((min(ra.x, rb.x) + min(ca.x, cb.x)) >> 10);
    const int ixhi = sat_short((max(ra.x, rb.x) + max(ca.x, cb.x)) >> 10);
    const int iylo = sat_short((min(ra.y, rb.y) + min(ca.y, cb.y)) >> 10);
    const int iyhi = sat_short((max(ra.y, rb.y) + max(ca.y, cb.y)) >> 10);
    T.rx0 = ixlo - o;
    T.ry0 = iylo - o;
    T.RW = ixhi - ixlo + K;
    T.RH = iyhi - iylo + K;
    const bool inner =
        T.rx0 >= 0 && T.ry0 >= 0 && T.rx0 + T.RW <= W && T.ry0 + T.RH <= H && ixhi < 32767 && iyhi < 32767;
    if (inner && T.RW <= S && T.RH <= win_cap / S && ((S * T.RH + 255) & ~255) <= win_cap)
        T.how = XF_BY_DMA;
    else if (T.RW <= win_cap && T.RH <= win_cap && T.RW * T.RH <= win_cap)
        T.how = XF_BY_LOADS;
    else
        T.how = XF_BY_GLOBAL;
    return T;
}

// Inner window into LDS by the buffer unit: rows of S cells (the window
// and whatever follows it), 16 bytes per lane, 64 quads per instruction;
// the last instruction's quads past the window are out of range and fill
// padding.  Completes on the vector-memory count.
template <int S>
__device__ __forceinline__ void xf_window_dma(const XfTile &T, const float *s, int W, size_t sec, float *buf)
{
    constexpr int q4 = S / 4, step_r = 64 / q4, step_c = 64 % q4;
    const int lane = threadIdx.x;
    const auto rs = __builtin_amdgcn_make_buffer_rsrc((void *)s, (short)0, (int)(sec * 4), XF_BUF_DWORD3);
    const int n4 = q4 * T.RH;
    int r = lane / q4, cq = lane % q4;
    for (int q0 = 0; q0 < n4; q0 += 64) {
        const unsigned off = q0 + lane < n4
                                 ? 4u * ((unsigned)(T.ry0 + r) * (unsigned)W + (unsigned)(T.rx0 + 4 * cq))
                                 : 0x80000000u;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void *)(buf + 4 * q0), 16,
                                                 off, 0, 0, 0);
        r += step_r;
        cq += step_c;
        if (cq >= q4) {
            cq -= q4;
            ++r;
        }
    }
}

// Any window into LDS through registers, XF_SB cells per lane per round:
// cells outside the section read as 0 through the buffer's range check.
// Cell idx = lane + 64 m sits at row r, column c, stepped without division.
__device__ __forceinline__ void xf_window_loads(const XfTile &T, const float *s, int W, int H, size_t sec,
                                                float *buf)
{
    const int lane = threadIdx.x;
    const auto rs = __builtin_amdgcn_make_buffer_rsrc((void *)s, (short)0, (int)(sec * 4), XF_BUF_DWORD3);
    const int RW = T.RW, n = T.RW * T.RH;
    const int q64 = 64 / RW, r64 = 64 % RW;
    int r = lane / RW, c = lane % RW;
    for (int base = 0; base < n; base += 64 * XF_SB) {
        float v[XF_SB];
#pragma unroll
        for (int u = 0; u < XF_SB; ++u) {
            const int gy = T.ry0 + r, gx = T.rx0 + c;
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
            if (idx < n) buf[idx] = v[u];
        }
    }
}

// The rows of an inner tile: lane l owns column x0 + l, the row term of
// each row is uniform.  With the fixed pitch S every tap of a pixel is an
// immediate offset from one LDS address.
template <int K, int S>
__device__ __forceinline__ void warp_inner(const XfTile &T, const float *buf, const float *__restrict__ tab,
                                           const int2 *rw, int2 cx, float *__restrict__ d, int W)
{
    constexpr int o = K / 2 - 1;
    const int x = T.tx0 + (int)threadIdx.x;
    // window origin folded in: cell (iy, ix) of the first tap sits at
    // wb[iy * S + ix]
    const float *wb = buf - ((T.ry0 + o) * S + T.rx0 + o);
    if (x > T.x1) return;
#pragma unroll
    for (int j = 0; j < XF_TH; ++j) {
        const int y = T.ty0 + j;
        if (y > T.y1) break;
        const int2 r = ld_term(&rw[y]);
        // two's-complement sums, as warpAffine's int arithmetic wraps; bits
        // 5..9 are the 1/32 fraction, the rest the integer part
        const int Sx = (int)((unsigned)r.x + (unsigned)cx.x), Sy = (int)((unsigned)r.y + (unsigned)cx.y);
        const float *txk = &tab[(Sx & 0x3e0) >> 2];
        const float *tyk = &tab[(Sy & 0x3e0) >> 2];
        float ty[K], txv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            ty[k] = tyk[k];
            txv[k] = txk[k];
        }
        float w[K][K];
        tap_weights<K>(ty, txv, w);
        // inner: both coordinates are small and non-negative
        const float v = interior_sum<K>(wb, __mul24(Sy >> 10, S) + (Sx >> 10), S, w);
        d[(size_t)y * W + x] = v;
    }
}

// The rows of any other tile, each pixel classified.  BY_LOADS: taps in
// the window `buf` of pitch RW; BY_GLOBAL: taps in the section `s`.
template <int K, bool STAGED>
__device__ __forceinline__ void warp_general(const XfTile &T, const float *taps, const float *__restrict__ tab,
                                             const int2 *rw, int2 cx, float cval, float *__restrict__ d, int W,
                                             int H)
{
    constexpr int o = K / 2 - 1;
    const int x = T.tx0 + (int)threadIdx.x;
    if (x > T.x1) return;
    for (int y = T.ty0; y <= T.y1; ++y) {
        const int2 r = ld_term(&rw[y]);
        const int Sx = (int)((unsigned)r.x + (unsigned)cx.x), Sy = (int)((unsigned)r.y + (unsigned)cx.y);
        const float *txk = &tab[(Sx & 0x3e0) >> 2];
        const float *tyk = &tab[(Sy & 0x3e0) >> 2];
        float ty[K], txv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            ty[k] = tyk[k];
            txv[k] = txk[k];
        }
        const int sx = sat_short(Sx >> 10) - o, sy = sat_short(Sy >> 10) - o;
        float v;
        if constexpr (STAGED)
            v = warp_pixel<K>(taps, (sy - T.ry0) * T.RW + (sx - T.rx0), T.RW, sx, sy, W, H, ty, txv, cval);
        else
            // wraps only for pixels whose taps all lie outside, which read nothing
            v = warp_pixel<K>(taps, (int)((unsigned)sy * (unsigned)W + (unsigned)sx), W, sx, sy, W, H, ty, txv,
                              cval);
        d[(size_t)y * W + x] = v;
    }
}

// Grid: one wave per work-group and per tile of XF_TW x XF_TH pixels, over
// all tiles of every section (raster order within a section).  Work-groups
// reach the 8 XCDs round-robin; each XCD takes one contiguous run of the
// list, so the windows of neighbouring tiles, which overlap, meet in one
// L2.  tab1 is the kernel table, [32][XF_TP] floats; win_cap the dynamic LDS
// window in floats.
template <int K, int S>
__global__ __launch_bounds__(64) void k_xf_warp(const float *__restrict__ src, float *__restrict__ dst, int W, int H,
                                                const int2 *__restrict__ rows, const int2 *__restrict__ cols,
                                                const float *__restrict__ cvals, const float *__restrict__ tab1,
                                                int tiles_x, unsigned ntiles, int win_cap)
{
    __shared__ __align__(16) float tab[XF_TAB * XF_TP];
    extern __shared__ float win[];
    const unsigned nb = gridDim.x, q = nb >> 3, rem = nb & 7u;
    const unsigned xcd = blockIdx.x & 7u;
    const unsigned b = xcd * q + (xcd < rem ? xcd : rem) + (blockIdx.x >> 3);
    const int lane = threadIdx.x;
    const size_t sec = (size_t)W * H;

    for (int k = lane; k < XF_TAB * XF_TP; k += 64) tab[k] = tab1[k];
    const XfTile T = xf_tile<K, S>(b, ntiles, tiles_x, W, H, rows, cols, win_cap);
    const float *s = src + (size_t)T.i * sec;
    if (T.how == XF_BY_DMA)
        xf_window_dma<S>(T, s, W, sec, win);
    else if (T.how == XF_BY_LOADS)
        xf_window_loads(T, s, W, H, sec, win);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    const int2 *rw = rows + (size_t)T.i * H;
    const int2 cx = cols[(size_t)T.i * W + min(T.tx0 + lane, T.x1)];
    float *d = dst + (size_t)T.i * sec;
    if (T.how == XF_BY_DMA) {
        warp_inner<K, S>(T, win, tab, rw, cx, d, W);
    } else {
        const float cval = *kc(&cvals[T.i]);
        if (T.how == XF_BY_LOADS)
            warp_general<K, true>(T, win, tab, rw, cx, cval, d, W, H);
        else
            warp_general<K, false>(T, s, tab, rw, cx, cval, d, W, H);
    }
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
    const int tiles_x = (a.W + XF_TW - 1) / XF_TW, tiles_y = (a.H + XF_TH - 1) / XF_TH;
    const unsigned ntiles = (unsigned)tiles_x * (unsigned)tiles_y;
    const dim3 grid(ntiles * (unsigned)a.P);
    const size_t lds = sizeof(float) * (size_t)a.win_cap;
#define XF_LAUNCH(K, S)                                                                                        \
    hipLaunchKernelGGL((k_xf_warp<K, S>), grid, dim3(64), lds, st, a.src, a.dst, a.W, a.H, a.rows, a.cols, a.cval, \
                       a.tab, tiles_x, ntiles, a.win_cap)
    if (a.lanczos) {
        if (a.pitch <= 72) XF_LAUNCH(8, 72);
        else XF_LAUNCH(8, 80);
    } else {
        if (a.pitch <= 72) XF_LAUNCH(4, 72);
        else XF_LAUNCH(4, 80);
    }
#undef XF_LAUNCH
    return hipGetLastError();
}

}  // namespace trec
