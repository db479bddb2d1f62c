// general.hip -- the general path of libtiltrec_hip.so (any pitch, offset or
// steep tilt), on the natural layouts: stack [P][Y][X], volume [Y][T][X] or
// [T][Y][X].
//
// TiltRec maps voxel (x, y, z) of tilt th to u = a0 + a1 X + a2 Y + a3 Z + ox,
// v = b0 + b1 X + b2 Y + b3 Z + oy, X = x - ox etc. (ValCoef,
// main-mpi.cpp:204-226), splats it bilinearly in Reproject (main-mpi.cpp:
// 286-344) and gathers it bilinearly in the back-projection (main-mpi.cpp:
// 378-450).  Its b is (0, 0, 1, 0) turned by the pitch and offset only
// (DecorateCoefficients, main-mpi.cpp:85-183), so every tilt maps rows the
// same way; the fast kernels below rest on that:
//
//   k_general_rows        forward projection when v does not depend on x
//                         (offset 0 or pitch 0): the splat factors into an
//                         x pass per volume row and a y pass over rows
//                         (k_general_rows3: three rows per slot).
//   k_general_rows3b      the same row windows when v moves slowly along x
//                         (a pitch with an offset), every voxel weighing its
//                         rows by its own v.
//   k_general_splat       forward projection of any pitch and offset (v may
//                         depend on x): the reference's splat into a tile of
//                         detector cells in LDS, owned by one wave.
//   k_general_back_rows   back-projection: v and its two detector rows are
//                         evaluated once per voxel, u once per voxel and tilt.
//   k_general_back_lds    the same with the taps staged in LDS: a wave's row
//                         of voxels reaches two detector rows (v independent
//                         of x) or a few more (a pitch with an offset), which
//                         the wave stages per tilt.
//
// Geometries outside their conditions use the reference-order kernels
// (k_general_gather, k_general_back).  Every kernel is deterministic.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "trec_internal.h"

namespace trec {

static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

__device__ __forceinline__ float hat(float d) { return fmaxf(0.f, 1.f - fabsf(d)); }

// Work-group b of n as dispatched (round robin over the 8 XCDs) -> its
// place in the work list, so that each XCD walks one contiguous range of it
// and the work-groups in flight on one XCD share their volume rows in its L2.
__device__ __forceinline__ unsigned xcd_order(unsigned b, unsigned n)
{
    const unsigned q = n >> 3, r = n & 7u;
    const unsigned xcd = b & 7u, k = b >> 3;
    return xcd * q + (xcd < r ? xcd : r) + k;
}

// ------------------------------------------------------------------ reference order

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

// Integer range [lo, hi] of k with (k - o) in [a, b) (rounding slack eps),
// clamped to [0, n).
__device__ __forceinline__ void gp_range(double a, double b, double o, int n, int &lo, int &hi)
{
    const double eps = 1e-6 * (1.0 + fabs(a) + fabs(b));
    const double l = ceil(a + o - eps), h = ceil(b + o + eps) - 1.0;
    lo = l < 0.0 ? 0 : (l > n ? n : (int)l);
    hi = h >= n ? n - 1 : (h < -1.0 ? -1 : (int)h);
}

// Reproject's splat (main-mpi.cpp:286-344) as a gather, for any geometry:
// pixel (i, j) of tilt th receives from voxel (x, y, z) the weight
// max(0, 1 - |u - i|) max(0, 1 - |v - j|), which is the splat's (1 - xd) or
// xd times (1 - yd) or yd, up to float rounding.  Sums run over z, then y,
// then x ascending with no atomics, so the result is deterministic.  Per
// layer the rows come from the v condition |v - j| < 1 (exactly when v does
// not depend on x, else from the inverse of the map's (x, y) part on the
// square's corners); per row, as in the row-blocked k_forward, the first
// column with u > i - 1 and its distance d0 from the pixel follow from one
// double evaluation, and candidate c sits at d0 + a1 c.
__global__ __launch_bounds__(WG) void k_general_gather(const GeneralJob g, int theta0, int ones)
{
    const unsigned nib = (g.X + WG - 1) / WG;
    const int i = (int)(blockIdx.x % nib) * WG + threadIdx.x;
    const int j = blockIdx.x / nib;
    if (i >= g.X) return;
    const int th = theta0 + blockIdx.y;
    const TiltGeneral &t = g.tg[th];
    const double a1 = t.a[1], a2 = t.a[2], b1 = t.b[1], b2 = t.b[2];
    const double det = a1 * b2 - a2 * b1;
    const double rb2 = b2 != 0.0 ? 1.0 / b2 : 0.0, rdet = det != 0.0 ? 1.0 / det : 0.0;
    const double ox = (float)t.ox, oy = (float)t.oy;
    const bool v_of_row = b1 == 0.0;
    const double aa1 = fabs(a1);
    // columns a pixel can see in one row: |a1| (x' - x) < 2
    const int ncand = aa1 > 1e-9 ? (int)ceil(2.0 / aa1 - 1e-6) : g.X;
    const double ra1 = aa1 > 1e-9 ? 1.0 / a1 : 0.0;
    const float a1f = (float)a1, b1f = (float)b1;
    float acc = 0.f;
    for (int z = 0; z < g.T; ++z) {
        const double Zc = (double)((float)z - (float)t.oz);
        const double cu = t.a[0] + t.a[3] * Zc + t.ox;  // u = cu + a1 Xc + a2 Yc
        const double cv = t.b[0] + t.b[3] * Zc + t.oy;  // v = cv + b1 Xc + b2 Yc
        const double du = (double)(i - 1) - cu, dv = (double)(j - 1) - cv;
        int y0, y1;
        if (v_of_row && b2 != 0.0) {
            const double p = dv * rb2, q = (dv + 2.0) * rb2;
            gp_range(fmin(p, q), fmax(p, q), oy, g.Y, y0, y1);
        } else if (det != 0.0) {
            const double c = -b1 * rdet, d = a1 * rdet;
            const double e0 = c * du + d * dv, e1 = c * (du + 2.0) + d * dv, e2 = c * du + d * (dv + 2.0),
                         e3 = c * (du + 2.0) + d * (dv + 2.0);
            gp_range(fmin(fmin(e0, e1), fmin(e2, e3)), fmax(fmax(e0, e1), fmax(e2, e3)), oy, g.Y, y0, y1);
        } else {
            y0 = 0;
            y1 = g.Y - 1;
        }
        for (int y = y0; y <= y1; ++y) {
            const double Yc = (double)((float)y - (float)t.oy);
            const double V0 = cv + b2 * Yc - (double)j;  // v - j at Xc = 0
            float wyr = 1.f;
            if (v_of_row) {
                wyr = fmaxf(0.f, 1.f - fabsf((float)V0));
                if (wyr == 0.f) continue;  // the row adds nothing to pixel row j
            }
            const double U0 = cu + a2 * Yc - (double)i;   // u - i at Xc = 0
            const float *vrow = g.vol + (g.layout == TREC_LAYOUT_Y ? ((size_t)y * g.T + z) * g.X
                                                                   : ((size_t)z * g.Y + y) * g.X);
            // first column xf whose u - i exceeds -1 (a1 > 0) or lies below
            // +1 (a1 < 0), walking towards growing |d|; d0 = u(xf) - i
            int xf, step;
            float d0;
            if (aa1 > 1e-9) {
                const double sx = (a1 > 0 ? (-1.0 - U0) : (1.0 - U0)) * ra1 + ox;
                const double fl = floor(sx);
                xf = (int)fl + 1;
                d0 = (float)(U0 + a1 * ((double)xf - ox));
                step = 1;
            } else {
                xf = 0;
                d0 = (float)U0;
                step = 1;
            }
            const float a1s = aa1 > 1e-9 ? a1f : 0.f;
            for (int c = 0; c < ncand; c += step) {
                const int x = xf + c;
                if ((unsigned)x >= (unsigned)g.X) continue;
                const float d = fmaf(a1s, (float)c, d0);
                const float wx = fmaxf(0.f, 1.f - fabsf(d));
                float wy = wyr;
                if (!v_of_row) wy = fmaxf(0.f, 1.f - fabsf((float)V0 + b1f * ((float)x - (float)ox)));
                const float val = ones ? 1.f : vrow[x];
                acc = fmaf(wx * wy, val, acc);
            }
        }
    }
    float *out = (ones ? g.rwt : g.rval) + (size_t)th * g.Y * g.X;
    out[(size_t)j * g.X + i] = acc;
}

// rval <- p - rval / rwt (rwt above the threshold), in place.
// QUOT (ADMM's operator, main-mpiz.cpp:700-760): rval <- -rval / rwt where
// rwt exceeds the threshold and 0 elsewhere, a tap the reference skips.
// With nzc > 1 the forward projection is the sum of nzc partials in
// g.part, added in chunk order.
template <bool QUOT>
__global__ __launch_bounds__(WG) void k_general_residual(const GeneralJob g, int theta0, size_t n, int nzc)
{
    const size_t i = (size_t)blockIdx.x * WG + threadIdx.x;
    if (i >= n) return;
    const size_t k = (size_t)theta0 * g.Y * g.X + i;
    float r;
    if (nzc > 1) {
        r = g.part[i];
        for (int c = 1; c < nzc; ++c) r += g.part[(size_t)c * n + i];
    } else {
        r = g.rval[k];
    }
    const float w = g.rwt[k];
    if (QUOT) {
        g.rval[k] = w > g.thr_res ? -(r / w) : 0.f;
        return;
    }
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
        if (mode == BACK_SIRT || mode == BACK_LIN) {
            val += s;
            wt += c;
        } else if (c > g.thr_upd) {
            out += (mode == BACK_SART) ? (s / c) * g.gamma : s / c;
        }
    }
    if (mode == BACK_LIN) {
        g.out[idx] = g.lin_a * val + g.lin_b * out;
        return;
    }
    if (mode == BACK_SIRT) {
        if (wt > g.thr_upd) out += val / wt * g.gamma;
    }
    g.vol[idx] = out;
}


// ------------------------------------------------------------------ row windows

// Forward projection for tilts whose v = cv(z) + b2 Y does not depend on x
// (b1 = 0).  The splat's weight hat(u - i) hat(v - j), hat(d) = max(0,
// 1 - |d|), then factors:
//
//   R[j][i] = sum_z sum_y hat(v(y, z) - j) S(y, z, i),
//   S(y, z, i) = sum_x hat(u(x, y, z) - i) V[x, y, z],
//
// an x pass per volume row (u moves with y through a2) and a y pass whose
// weights do not depend on i.  A thread owns detector rows j0 .. j0 + RJ - 1
// of pixel i of one tilt.  With ya(j) = floor(y*), v(y*) = j, the rows of a
// layer that reach row j are ya(j) - 1 .. ya(j) + 2 (b2 >= 1/2), and for
// (RJ - 1) / RJ <= b2 <= 1, ya(j0 + r) - ya(j0) is r or r + 1: the rows
// ya(j0) - 1 + k, k < RJ + 4 ("slots"), cover every detector row, and slot
// k reaches only rows r in [k - 4, k].
//
// Each slot's S is formed from NC candidate columns starting at the first
// with |u - i| < 1 (clamped into the row, so every load is in bounds; a
// column that is no candidate gets hat = 0).  The y weights hat(v - j0 - r)
// are the same in every lane: per layer the wave evaluates the 5 (RJ + 4)
// slot/row pairs once, one per lane, stores them in its LDS, and each slot
// reads its five from there as uniform operands (zero for rows it does not
// reach and for slots outside the volume).  Sums run over z, then slots in y
// order, then columns in x order.  Work-groups are ordered tilt first, so
// those in flight together read the same rows of the volume, and each XCD
// walks a contiguous range of that order.
template <int RJ, int NC, bool ONES>
__global__ __launch_bounds__(WG) void k_general_rows(const GeneralJob g, int theta0, int ntheta, int nzc)
{
    constexpr int NS = RJ + 4;
    constexpr int NE = 5 * NS;             // slot/row pairs: slot q / 5, row q / 5 - 4 + q % 5
    constexpr int NH = (NE + 63) / 64;     // lane registers holding them
    constexpr int SB = 5;                  // slots whose loads may be in flight together
    __shared__ float ywl[WG / 64][64 * NH];
    float *yw = ywl[threadIdx.x / 64];
    const unsigned nib = (g.X + WG - 1) / WG;
    const unsigned lin = xcd_order(blockIdx.x, gridDim.x);
    const int th = theta0 + (int)(lin % (unsigned)ntheta);
    const unsigned rest = lin / (unsigned)ntheta;
    // every lane stays to the end: the wave's lanes hold the y weights of all
    // of them; pixels beyond the row repeat its last one and are not stored
    const int i = (int)(rest % nib) * WG + threadIdx.x;
    const int ic = min(i, g.X - 1);
    const unsigned njb = (g.Y + RJ - 1) / RJ;
    const int j0 = (int)((rest / nib) % njb) * RJ;
    // with nzc > 1, layers zc * zlen .. of a chunk, summed into g.part
    const int zc = (int)(rest / nib / njb);
    const int zlen = (g.T + nzc - 1) / nzc;
    const int z0 = zc * zlen, z1 = min(g.T, z0 + zlen);
    const TiltGeneral &t = g.tg[th];
    const double a1 = t.a[1], a2 = t.a[2], b2 = t.b[2];
    const double ra1 = 1.0 / a1, rb2 = 1.0 / b2;
    const double ox = (float)t.ox, oy = (float)t.oy;
    // u - i = W + a1 x with W = cu - i + a2 Y - a1 ox; the first candidate is
    // the least x with (W + a1 x - s1) / a1 > 0
    const double s1ra1 = (a1 > 0.0 ? -1.0 : 1.0) * ra1;
    const float a1f = (float)a1, a2f = (float)a2, rkf = (float)(a2 * ra1);
    const int xmax = g.X - NC;
    // volume strides of a row and a layer
    const size_t sy = g.layout == TREC_LAYOUT_Y ? (size_t)g.T * g.X : (size_t)g.X;
    const size_t sz = g.layout == TREC_LAYOUT_Y ? (size_t)g.X : (size_t)g.Y * g.X;
    // this lane's slot/row pairs
    const int lane = (int)(threadIdx.x & 63);
    int pk[NH];
    double pkd[NH], pj[NH];
    float pm[NH];
    bool prow[NH];
#pragma unroll
    for (int h = 0; h < NH; ++h) {
        const int q = 64 * h + lane, k = q / 5, r = k - 4 + q % 5;
        pk[h] = k;
        pkd[h] = (double)k;
        pj[h] = (double)(j0 + k - 2);
        pm[h] = (float)(r - k + 2);
        prow[h] = q < NE && r >= 0 && r < RJ;
    }
    float acc[RJ];
#pragma unroll
    for (int r = 0; r < RJ; ++r) acc[r] = 0.f;
    for (int z = z0; z < z1; ++z) {
        const double Zc = (double)((float)z - (float)t.oz);
        const double cu = t.a[0] + t.a[3] * Zc + t.ox;  // u = cu + a1 X + a2 Y
        const double cv = t.b[0] + t.b[3] * Zc + t.oy;  // v = cv + b2 Y
        // uniform (the layer's first slot row): the slots' row bases then live
        // in scalar registers and a load adds only the lane's column
        const int ylo = __builtin_amdgcn_readfirstlane((int)floor(((double)j0 - cv) * rb2 + oy) - 1);
        const double Yc0 = (double)((float)ylo - (float)t.oy);
        // y weights: v - (j0 + k - 2) of slot k, then hat of its distance to row r
        float pw[NH];
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            const float e = (float)((b2 * (Yc0 + pkd[h]) + cv) - pj[h]);
            pw[h] = prow[h] && (unsigned)(ylo + pk[h]) < (unsigned)g.Y ? hat(e - pm[h]) : 0.f;
        }
        // the slot/row weights go to the wave's LDS, from where every slot
        // reads its five as uniform operands (broadcast reads, no VALU)
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int h = 0; h < NH; ++h) yw[64 * h + lane] = pw[h];
        __builtin_amdgcn_wave_barrier();
        const double W0 = cu - (double)ic + a2 * Yc0 - a1 * ox;
        // slot k's first candidate is floor(F0 - rk k) + 1 with F0 = (s1 -
        // W0) / a1, and u - i at column x is E0 + a1 (x - xf0) + a2 k with
        // E0 = a1 xf0 + W0 (|E0| <= 1 + |a1|): both in float from the layer's
        // double F0 and E0.  Where float rounding moves a slot's start by a
        // column, the column gained or lost lies within rounding of
        // |u - i| = 1, where its weight is of that order.
        const double F0 = s1ra1 - W0 * ra1;
        const double xf0d = floor(F0);
        const int xf0 = (int)xf0d;
        const float f0 = (float)(F0 - xf0d);
        const float E0 = (float)(a1 * xf0d + W0);
        const float *vlayer = g.vol + (size_t)z * sz;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int xf = xf0 + (int)floorf(fmaf(-rkf, (float)k, f0)) + 1;
            const int xs = min(max(xf, 0), xmax);
            const float d0 = fmaf(a1f, (float)(xs - xf0), fmaf(a2f, (float)k, E0));  // u - i at column xs
            float s = 0.f;
            if (ONES) {
#pragma unroll
                for (int c = 0; c < NC; ++c) s += hat(fmaf(a1f, (float)c, d0));
            } else {
                const int yc = min(max(ylo + k, 0), g.Y - 1);
                const float *p = reinterpret_cast<const float *>(
                    reinterpret_cast<const char *>(vlayer + (size_t)yc * sy) + (unsigned)xs * 4u);
                float v[NC];
#pragma unroll
                for (int c = 0; c < NC; ++c) v[c] = p[c];
#pragma unroll
                for (int c = 0; c < NC; ++c) s = fmaf(hat(fmaf(a1f, (float)c, d0)), v[c], s);
            }
#pragma unroll
            for (int r = (k > 4 ? k - 4 : 0); r <= (k < RJ - 1 ? k : RJ - 1); ++r) {
                const int q = 5 * k + r - k + 4;
                acc[r] = fmaf(yw[q], s, acc[r]);
            }
            if ((k + 1) % SB == 0) __builtin_amdgcn_sched_barrier(0);
        }
    }
    if (i >= g.X) return;
    float *out = nzc > 1 ? g.part + ((size_t)zc * ntheta + (th - theta0)) * g.Y * g.X
                         : (ONES ? g.rwt : g.rval) + (size_t)th * g.Y * g.X;
#pragma unroll
    for (int r = 0; r < RJ; ++r)
        if (j0 + r < g.Y) out[(size_t)(j0 + r) * g.X + i] = acc[r];
}

// The row-window sums of k_general_rows for a1 > 0, arranged so that a slot
// costs about half the instructions:
//
//  - With e = 1 - b2, ylo is the least volume row with v(ylo) >= j0 - e.
//    Then g_k = v(ylo + k) - j0 - k = g0 - e k with g0 in [-e, 1 - 2e), so
//    g_k lies in [-1, 1] for k = -2 .. RJ + 1 while e (RJ + 2) <= 1.  Slot k
//    then reaches rows k - 1, k, k + 1 only, with weights max(0, -g_k),
//    1 - |g_k| and max(0, g_k), and only slots k = -1 .. RJ reach rows 0 ..
//    RJ - 1 (RJ + 2 slots, three weights each, instead of RJ + 4 and five).
//  - The first candidate column of slot k is floor(F0 - (a2 / a1) k) + 1;
//    with fr the fraction of F0 - (a2 / a1) k, candidate c lies at
//    a1 (1 + c - fr) - 1 from the pixel.  Its weight is a1 (1 - fr) for c =
//    0 and 1 - |a1 (2 - fr) - 1| for c = 1, both known to lie in [0, 1], and
//    for the later ones whose distance is known to be positive (2 c >= NC)
//    2 - a1 (1 + c - fr) clamped into [0, 1] (the clamp is an instruction
//    modifier): one or two instructions each instead of three.
//  - The per-layer constants (ylo, g0, the u offset, whether the wave's
//    windows reach the row at all, and whether they all lie inside it) are
//    evaluated for 64 layers at once, one per lane, and read per layer with
//    v_readlane.  A layer whose windows all miss the row is skipped.
//  - A wave whose windows all lie inside the row (decided from its first and
//    last pixel, with a column of margin) loads them as they are; the others
//    clamp each window into the row and weigh every column by the hat of its
//    distance, as k_general_rows does.
//
// Sums run over z, then slots in y order, then columns in x order.
__device__ __forceinline__ float clamp01(float x) { return __builtin_amdgcn_fmed3f(x, 0.f, 1.f); }

template <int RJ, int NC, bool ONES>
__global__ __launch_bounds__(WG) void k_general_rows3(const GeneralJob g, int theta0, int ntheta, int nzc)
{
    constexpr int NS = RJ + 2;  // slots k = -1 .. RJ at s = k + 1
    // slots whose loads may be in flight together: the 18 slots of 16 rows
    // in two groups of 9 (2.4 % faster than groups of 6, at the same 4 waves
    // per SIMD), else 6
    constexpr int SB = RJ == 16 ? 9 : 6;
    static_assert(NS <= 64, "a lane per slot");
    __shared__ float4 ywl[WG / 64][NS];
    float4 *yw = ywl[threadIdx.x / 64];
    const unsigned nib = (g.X + WG - 1) / WG;
    const unsigned lin = xcd_order(blockIdx.x, gridDim.x);
    const int th = theta0 + (int)(lin % (unsigned)ntheta);
    const unsigned rest = lin / (unsigned)ntheta;
    // pixels beyond the row repeat its last one and are not stored
    const int i = (int)(rest % nib) * WG + threadIdx.x;
    const int ic = min(i, g.X - 1);
    const int lane = (int)(threadIdx.x & 63);
    const int iw = __builtin_amdgcn_readfirstlane(i - lane);  // the wave's first pixel
    const int ia = min(iw, g.X - 1), ib = min(iw + 63, g.X - 1);
    const unsigned njb = (g.Y + RJ - 1) / RJ;
    const int j0 = (int)((rest / nib) % njb) * RJ;
    const int zc = (int)(rest / nib / njb);
    const int zlen = (g.T + nzc - 1) / nzc;
    const int z0 = zc * zlen, z1 = min(g.T, z0 + zlen);
    const TiltGeneral &t = g.tg[th];
    const double a1 = t.a[1], a2 = t.a[2], b2 = t.b[2];
    const double ra1 = 1.0 / a1, rb2 = 1.0 / b2;
    const double ox = (float)t.ox, oy = (float)t.oy;
    const float a1f = (float)a1, rkf = (float)(a2 * ra1), ef = (float)(1.0 - b2);
    float kc[NC];  // c = 1: 2 a1 - 1; 2 c >= NC: 2 - a1 (1 + c); else a1 (1 + c) - 1
#pragma unroll
    for (int c = 0; c < NC; ++c)
        kc[c] = c == 1 ? (float)(2.0 * a1 - 1.0) : (2 * c >= NC ? (float)(2.0 - a1 * (1.0 + c)) : (float)(a1 * (1.0 + c) - 1.0));
    // how far the first candidate moves over the slots, plus rounding slack
    const double drift = fabs(a2 * ra1) * (double)RJ + 0.01;
    const size_t sy = g.layout == TREC_LAYOUT_Y ? (size_t)g.T * g.X : (size_t)g.X;
    const size_t sz = g.layout == TREC_LAYOUT_Y ? (size_t)g.X : (size_t)g.Y * g.X;
    // constants of layers zb .. zb + 63, lane l holding layer zb + l: u - i
    // at column x of slot k is W - i + a1 x + a2 k; mode 0 skips the layer,
    // 1 takes the clamped windows, 2 the inner ones
    int ylo_v = 0, mode_v = 0;
    float g0_v = 0.f;
    double w_v = 0.0;
    auto layer_block = [&](int zb) {
        const int z = min(zb + lane, z1 - 1);
        const double Zc = (double)((float)z - (float)t.oz);
        const double cu = t.a[0] + t.a[3] * Zc + t.ox;  // u = cu + a1 X + a2 Y
        const double cv = t.b[0] + t.b[3] * Zc + t.oy;  // v = cv + b2 Y
        ylo_v = (int)ceil(((double)j0 - (1.0 - b2) - cv) * rb2 + oy);
        const double Yc0 = (double)((float)ylo_v - (float)t.oy);
        g0_v = (float)(cv + b2 * Yc0 - (double)j0);
        w_v = cu + a2 * Yc0 - a1 * ox;
        const double la = floor(((double)ia - 1.0 - w_v) * ra1 - drift) + 1.0;  // least first column
        const double lb = floor(((double)ib - 1.0 - w_v) * ra1 + drift) + 1.0;  // greatest first column
        const bool reach = lb + (double)NC >= 1.0 && la <= (double)(g.X - 1) && ylo_v + RJ >= 0 && ylo_v - 1 < g.Y;
        const bool inner = la >= 0.0 && lb + (double)NC <= (double)g.X;
        mode_v = reach ? (inner ? 2 : 1) : 0;
    };
    float acc[RJ];
#pragma unroll
    for (int r = 0; r < RJ; ++r) acc[r] = 0.f;
    int zb = z0;
    layer_block(zb);
    for (int z = z0; z < z1; ++z) {
        if (z - zb == 64) {
            zb = z;
            layer_block(zb);
        }
        const int zl = z - zb;
        const int mode = __builtin_amdgcn_readlane(mode_v, zl);
        if (mode == 0) continue;
        const int ylo = __builtin_amdgcn_readlane(ylo_v, zl);
        const float g0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(g0_v), zl));
        const unsigned long long wb = (unsigned long long)__double_as_longlong(w_v);
        const double W = __longlong_as_double(
            (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(wb >> 32), zl) << 32) |
                        (unsigned)__builtin_amdgcn_readlane((int)(unsigned)wb, zl)));
        {
            // lane s takes slot s; rows outside the volume weigh nothing
            const int s = min(lane, NS - 1);
            const float gk = __builtin_amdgcn_fmed3f(fmaf(-ef, (float)(s - 1), g0), -1.f, 1.f);
            const bool in = (unsigned)(ylo - 1 + s) < (unsigned)g.Y;
            const float4 w = in ? make_float4(fmaxf(0.f, -gk), 1.f - fabsf(gk), fmaxf(0.f, gk), 0.f)
                                : make_float4(0.f, 0.f, 0.f, 0.f);
            __builtin_amdgcn_wave_barrier();
            if (lane < NS) yw[s] = w;
            __builtin_amdgcn_wave_barrier();
        }
        const double F0 = ((double)ic - 1.0 - W) * ra1;
        const double X0d = floor(F0);
        const int X1 = (int)X0d + 1;
        const float f0 = (float)(F0 - X0d);
        const float *vlayer = g.vol + (size_t)z * sz;
        auto slots = [&](auto edge_tag) {
            constexpr bool EDGE = decltype(edge_tag)::value;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const int k = s - 1;
                const float tk = fmaf(-rkf, (float)k, f0);
                const int xf = X1 + (int)floorf(tk);
                const float fr = __builtin_amdgcn_fractf(tk);
                const int xs = EDGE ? min(max(xf, 0), g.X - NC) : xf;
                float w[NC];
                if constexpr (EDGE) {
                    const float d0 = fmaf(a1f, (float)(xs - xf), fmaf(-a1f, fr, a1f - 1.f));
#pragma unroll
                    for (int c = 0; c < NC; ++c) w[c] = clamp01(1.f - fabsf(fmaf(a1f, (float)c, d0)));
                } else {
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        if (c == 0) w[c] = fmaf(-a1f, fr, a1f);
                        else if (c == 1) w[c] = 1.f - fabsf(fmaf(-a1f, fr, kc[1]));
                        else if (2 * c >= NC) w[c] = clamp01(fmaf(a1f, fr, kc[c]));
                        else w[c] = clamp01(1.f - fabsf(fmaf(-a1f, fr, kc[c])));
                    }
                }
                float sx = 0.f;
                if constexpr (ONES) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) sx += w[c];
                } else {
                    const int yc = min(max(ylo + k, 0), g.Y - 1);
                    const float *p = reinterpret_cast<const float *>(
                        reinterpret_cast<const char *>(vlayer + (size_t)yc * sy) + (unsigned)xs * 4u);
                    float v[NC];
#pragma unroll
                    for (int c = 0; c < NC; ++c) v[c] = p[c];
#pragma unroll
                    for (int c = 0; c < NC; ++c) sx = fmaf(w[c], v[c], sx);
                }
                const float4 wy = yw[s];
                if (s >= 2 && s - 2 < RJ) acc[s - 2] = fmaf(wy.x, sx, acc[s - 2]);
                if (s >= 1 && s - 1 < RJ) acc[s - 1] = fmaf(wy.y, sx, acc[s - 1]);
                if (s < RJ) acc[s] = fmaf(wy.z, sx, acc[s]);
                if ((s + 1) % SB == 0) __builtin_amdgcn_sched_barrier(0);
            }
        };
        if (mode == 2) slots(std::false_type{});
        else slots(std::true_type{});
    }
    if (i >= g.X) return;
    float *out = nzc > 1 ? g.part + ((size_t)zc * ntheta + (th - theta0)) * g.Y * g.X
                         : (ONES ? g.rwt : g.rval) + (size_t)th * g.Y * g.X;
#pragma unroll
    for (int r = 0; r < RJ; ++r)
        if (j0 + r < g.Y) out[(size_t)(j0 + r) * g.X + i] = acc[r];
}

// The three-row windows of k_general_rows3 for a pitch with an offset.  Then
// v = cv + b1 X + b2 Y moves along x (b1 = sin(pitch) sin(offset), the same
// in every tilt), and each voxel weighs its detector rows by its own v:
//
//  - The wave's ylo is chosen at its reference column xr (near the middle
//    of its candidate columns), with g0 = v(xr, ylo) - j0 in [lo, lo + b2),
//    lo = -1 + e RJ + D (the launcher's dlo).  Without LANE every lane takes
//    those rows, and D bounds |b1 (x - xr)| over the wave's candidates x.
//    With LANE a lane whose first candidate X1 has g = g0 + b1 (X1 - xr)
//    outside [lo, lo + b2) shifts its rows by dl = ceil((lo - g) / b2),
//    |dl| <= dmax (a pixel's u then moves by a2 dl, and X1 with it), and D
//    bounds |b1 (x - X1)| over a lane's candidates and the shift of X1:
//    some instructions more per layer, but |b1| up to about 0.05 instead of
//    about half a row over the wave's columns.  Either way the voxel at
//    column x of slot k has g in [-1, 1] for k = -1 .. RJ while
//    e RJ + 2 D <= 1, and reaches rows k - 1, k, k + 1 with weights
//    max(0, -g), 1 - |g| and max(0, g).  The per-layer tests (whether the
//    wave's windows reach the row and the volume, and whether they lie
//    inside them) allow for dmax rows and a2 / a1 dmax columns of shift.
//  - With p_c the column-weighted voxel of candidate c and g_c = g_k + b1 c,
//    rows k - 1, k, k + 1 take (A - G) / 2, S - A and (A + G) / 2 for S =
//    sum p_c, A = sum p_c |g_c| and G = sum p_c g_c = g_k S + b1 S1,
//    S1 = sum c p_c.  The candidates are summed last to first, so the partial
//    sums of that chain add up to S1.  When g keeps its sign over a lane's
//    candidates, A = +-G; a slot where some lane's candidates straddle g = 0
//    forms A and G voxel by voxel in every lane of the wave.
//  - A tilt takes 2 / a1 candidate columns rounded up (at least 3), chosen
//    per work-group, so the tilts near 0 degrees do not pay for the steep
//    ones.
//  - Slots go in groups of 6, 5, 4 or 3 for 3, 4, 5 or 6 candidates (at
//    most 20 loaded values, which keeps four waves per SIMD), each group's
//    loads issued before any of its branches.  A slot none of whose rows
//    lies inside the volume is skipped.  Without LANE a load is the slot's
//    row (clamped) plus the lane's column.  With LANE, inside the volume it
//    is the layer's row ylo - dmax - 1 plus the lane's 32-bit offset (dl +
//    dmax + k + 1) sy + xs; in a wave at the volume's or the row's edges a
//    lane whose row lies outside the volume loads a clamped row and adds
//    nothing.
//
// Candidate columns, their weights, the edge handling and the per-layer
// constants are those of k_general_rows3.  Sums run over z, then slots in y
// order.
template <int RJ, bool ONES, bool LANE>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(4))) void k_general_rows3b(const GeneralJob g, int theta0, int ntheta, int nzc,
                                                       double dlo, int dmax_arg)
{
    const int dmax = LANE ? dmax_arg : 0;  // without LANE every lane takes the wave's rows
    constexpr int NS = RJ + 2;  // slots k = -1 .. RJ at s = k + 1
    constexpr int NCM = 6;      // most candidate columns (a1 >= 1/3)
    const unsigned nib = (g.X + WG - 1) / WG;
    const unsigned lin = xcd_order(blockIdx.x, gridDim.x);
    const int th = theta0 + (int)(lin % (unsigned)ntheta);
    const unsigned rest = lin / (unsigned)ntheta;
    // pixels beyond the row repeat its last one and are not stored
    const int i = (int)(rest % nib) * WG + threadIdx.x;
    const int ic = min(i, g.X - 1);
    const int lane = (int)(threadIdx.x & 63);
    const int iw = __builtin_amdgcn_readfirstlane(i - lane);  // the wave's first pixel
    const int ia = min(iw, g.X - 1), ib = min(iw + 63, g.X - 1);
    const unsigned njb = (g.Y + RJ - 1) / RJ;
    const int j0 = (int)((rest / nib) % njb) * RJ;
    const int zc = (int)(rest / nib / njb);
    const int zlen = (g.T + nzc - 1) / nzc;
    const int z0 = zc * zlen, z1 = min(g.T, z0 + zlen);
    const TiltGeneral &t = g.tg[th];
    const double a1 = t.a[1], a2 = t.a[2], b1 = t.b[1], b2 = t.b[2];
    const double ra1 = 1.0 / a1, rb2 = 1.0 / b2;
    const double ox = (float)t.ox, oy = (float)t.oy;
    const float a1f = (float)a1, rkf = (float)(a2 * ra1), ef = (float)(1.0 - b2), b1f = (float)b1;
    const float b2f = (float)b2, rb2f = (float)rb2, lof = (float)dlo;
    const double rk = a2 * ra1;
    // the tilt's candidate count
    const int nct = __builtin_amdgcn_readfirstlane(min(NCM, max(3, (int)ceil(2.0 / a1 - 1e-6))));
    // how far the first candidate moves over the slots and a lane's row
    // shift, plus rounding slack
    const double drift = fabs(rk) * (double)(RJ + dmax) + 0.01;
    const size_t sy = g.layout == TREC_LAYOUT_Y ? (size_t)g.T * g.X : (size_t)g.X;
    const size_t sz = g.layout == TREC_LAYOUT_Y ? (size_t)g.X : (size_t)g.Y * g.X;
    // constants of layers zb .. zb + 63, lane l holding layer zb + l: u - i
    // at column x of slot k is W - i + a1 x + a2 k; mode 0 skips the layer,
    // 1 takes the clamped windows, 2 the inner ones
    int ylo_v = 0, mode_v = 0, xr_v = 0;
    float g0_v = 0.f;
    double w_v = 0.0;
    auto layer_block = [&](int zb) {
        const int z = min(zb + lane, z1 - 1);
        const double Zc = (double)((float)z - (float)t.oz);
        const double cu = t.a[0] + t.a[3] * Zc + t.ox;  // u = cu + a1 X + a2 Y
        const double cv = t.b[0] + t.b[3] * Zc + t.oy;  // v = cv + b1 X + b2 Y
        // the reference column: the wave's middle pixel on the rows near j0
        const double xm = (0.5 * (double)(ia + ib) - cu - a2 * (((double)j0 - cv) * rb2)) * ra1 + ox;
        xr_v = (int)fmin(fmax(floor(xm), -1e9), 1e9);
        const double Xr = (double)((float)xr_v - (float)t.ox);
        ylo_v = (int)fmin(fmax(ceil(((double)j0 + dlo - cv - b1 * Xr) * rb2 + oy), -1e9), 1e9);
        const double Yc0 = (double)((float)ylo_v - (float)t.oy);
        g0_v = (float)(cv + b1 * Xr + b2 * Yc0 - (double)j0);
        w_v = cu + a2 * Yc0 - a1 * ox;
        const double la = floor(((double)ia - 1.0 - w_v) * ra1 - drift) + 1.0;  // least first column
        const double lb = floor(((double)ib - 1.0 - w_v) * ra1 + drift) + 1.0;  // greatest first column
        const bool reach =
            lb + (double)nct >= 1.0 && la <= (double)(g.X - 1) && ylo_v + dmax + RJ >= 0 && ylo_v - dmax - 1 < g.Y;
        // with each lane's rows, also every lane's rows inside the volume
        // (the wave's own rows outside it are skipped slot by slot)
        const bool inner = la >= 0.0 && lb + (double)nct <= (double)g.X &&
                           (!LANE || (ylo_v - dmax - 1 >= 0 && ylo_v + dmax + RJ <= g.Y - 1));
        mode_v = reach ? (inner ? 2 : 1) : 0;
    };
    const unsigned sy4 = (unsigned)(sy * 4u);  // the launcher keeps (2 dmax + RJ + 2) sy + X floats under 4 GiB
    float acc[RJ];
#pragma unroll
    for (int r = 0; r < RJ; ++r) acc[r] = 0.f;
    int zb = z0;
    layer_block(zb);
    for (int z = z0; z < z1; ++z) {
        if (z - zb == 64) {
            zb = z;
            layer_block(zb);
        }
        const int zl = z - zb;
        const int mode = __builtin_amdgcn_readlane(mode_v, zl);
        if (mode == 0) continue;
        const int ylo = __builtin_amdgcn_readlane(ylo_v, zl);
        const int xr = __builtin_amdgcn_readlane(xr_v, zl);
        const float g0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(g0_v), zl));
        const unsigned long long wb = (unsigned long long)__double_as_longlong(w_v);
        const double W = __longlong_as_double(
            (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(wb >> 32), zl) << 32) |
                        (unsigned)__builtin_amdgcn_readlane((int)(unsigned)wb, zl)));
        // the lane's row shift dl from g at its first candidate, then that
        // candidate with the rows shifted (u moves by a2 dl) and g there
        const double F0w = ((double)ic - 1.0 - W) * ra1;
        // (the launcher's dmax bounds |dl|; the clamp keeps every load inside
        // the volume even so)
        const int dl =
            LANE ? min(max((int)ceilf((lof - fmaf(b1f, (float)((int)floor(F0w) + 1 - xr), g0)) * rb2f), -dmax), dmax)
                 : 0;
        const double F0 = F0w - rk * (double)dl;
        const double X0d = floor(F0);
        const int X1 = (int)X0d + 1;
        const float f0 = (float)(F0 - X0d);
        // g at column X1 of slot k = 0
        const float gl = fmaf(b1f, (float)(X1 - xr), fmaf(b2f, (float)dl, g0));
        const unsigned roff = (unsigned)(dl + dmax) * sy4;  // the lane's rows past ylo - dmax - 1
        const float *vlayer = g.vol + (size_t)z * sz;
        const char *vbase =
            reinterpret_cast<const char *>(vlayer + (size_t)max(ylo - dmax - 1, 0) * sy);  // used when inside
        auto slots = [&](auto edge_tag, auto nc_tag) {
            constexpr bool EDGE = decltype(edge_tag)::value;
            constexpr int NC = decltype(nc_tag)::value;
            // a1 (1 + c) - 1, 2 - a1 (1 + c), b1 c (formed here, so that only
            // the variant running holds them)
            float kp[NC], kn[NC], bc[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                kp[c] = fmaf(a1f, (float)(1 + c), -1.f);
                kn[c] = fmaf(-a1f, (float)(1 + c), 2.f);
                bc[c] = b1f * (float)c;
            }
            // groups of Q slots: every load of a group is issued before any
            // of its branches (a slot outside the volume loads a clamped row)
            constexpr int Q = NC <= 3 ? 6 : (NC == 4 ? 5 : (NC == 5 ? 4 : 3));  // at most 20 loaded values
#pragma unroll
            for (int s0 = 0; s0 < NS; s0 += Q) {
                float v[Q][NC], fr[Q], gk[Q];
                int xs[Q], xf[Q];
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    const int s = s0 + q, k = s - 1;
                    if (s >= NS) break;
                    const float tk = fmaf(-rkf, (float)k, f0);
                    const float flt = floorf(tk);
                    xf[q] = X1 + (int)flt;
                    fr[q] = __builtin_amdgcn_fractf(tk);
                    xs[q] = EDGE ? min(max(xf[q], 0), g.X - NC) : xf[q];
                    // g of the slot's first column
                    gk[q] = EDGE ? fmaf(b1f, (float)(xs[q] - X1), fmaf(-ef, (float)k, gl))
                                 : fmaf(b1f, flt, fmaf(-ef, (float)k, gl));
                    if constexpr (!LANE) {
                        // the wave's row (clamped; a slot outside the volume is skipped)
                        const int yc = min(max(ylo + k, 0), g.Y - 1);
                        const float *p = reinterpret_cast<const float *>(
                            reinterpret_cast<const char *>(vlayer + (size_t)yc * sy) + (unsigned)xs[q] * 4u);
#pragma unroll
                        for (int c = 0; c < NC; ++c) v[q][c] = ONES ? 1.f : p[c];
                    } else if constexpr (EDGE) {
                        // the lane's row, clamped into the volume (outside it adds
                        // nothing), as an offset of at most 2 dmax rows past the
                        // slot's least row in the volume
                        const int row = ylo + dl + k;
                        const bool in = (unsigned)row < (unsigned)g.Y;
                        const int rb = __builtin_amdgcn_readfirstlane(min(max(ylo - dmax + k, 0), g.Y - 1));
                        const unsigned off = (unsigned)(min(max(row, rb), g.Y - 1) - rb) * sy4 + (unsigned)xs[q] * 4u;
                        const float *p = reinterpret_cast<const float *>(
                            reinterpret_cast<const char *>(vlayer + (size_t)rb * sy) + off);
#pragma unroll
                        for (int c = 0; c < NC; ++c) v[q][c] = in ? (ONES ? 1.f : p[c]) : 0.f;
                    } else if constexpr (ONES) {
#pragma unroll
                        for (int c = 0; c < NC; ++c) v[q][c] = 1.f;
                    } else {
                        // every lane's rows lie inside the volume: slot s of the
                        // lane is s rows past its first, dl + dmax rows past
                        // the layer's base row ylo - dmax - 1
                        const float *p = reinterpret_cast<const float *>(
                            vbase + (roff + (unsigned)s * sy4 + (unsigned)xs[q] * 4u));
#pragma unroll
                        for (int c = 0; c < NC; ++c) v[q][c] = p[c];
                    }
                }
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    const int s = s0 + q, k = s - 1;
                    if (s >= NS) break;
                    // no lane's row of this slot, ylo + k - dmax .. ylo + k +
                    // dmax, lies inside the volume (one unsigned comparison)
                    if ((unsigned)(ylo + dmax + k) >= (unsigned)(g.Y + 2 * dmax)) continue;
                    float w[NC];
                    if constexpr (EDGE) {
                        const float d0 = fmaf(a1f, (float)(xs[q] - xf[q]), fmaf(-a1f, fr[q], a1f - 1.f));
#pragma unroll
                        for (int c = 0; c < NC; ++c) w[c] = clamp01(1.f - fabsf(fmaf(a1f, (float)c, d0)));
                    } else {
#pragma unroll
                        for (int c = 0; c < NC; ++c) {
                            if (c == 0) w[c] = fmaf(-a1f, fr[q], a1f);
                            else if (c == 1) w[c] = 1.f - fabsf(fmaf(-a1f, fr[q], kp[1]));
                            else if (2 * c >= NC) w[c] = clamp01(fmaf(a1f, fr[q], kn[c]));
                            else w[c] = clamp01(1.f - fabsf(fmaf(-a1f, fr[q], kp[c])));
                        }
                    }
                    // S and S1 from the chain last candidate to first
                    float S = 0.f, S1 = 0.f;
#pragma unroll
                    for (int c = NC - 1; c >= 0; --c) {
                        S = c == NC - 1 ? w[c] * v[q][c] : fmaf(w[c], v[q][c], S);
                        if (c >= 1) S1 = c == NC - 1 ? S : S1 + S;
                    }
                    const float G = fmaf(gk[q], S, b1f * S1);
                    const bool pos = gk[q] >= 0.f;
                    const bool straddle = pos != (gk[q] + bc[NC - 1] >= 0.f);
                    float dn, dm, dp;  // rows k - 1, k, k + 1
                    if (__builtin_amdgcn_ballot_w64(straddle) == 0) {
                        const float tA = pos ? G : -G;
                        dn = pos ? 0.f : tA;
                        dm = S - tA;
                        dp = pos ? G : 0.f;
                    } else {
                        float A = 0.f, Gx = 0.f;
#pragma unroll
                        for (int c = 0; c < NC; ++c) {
                            const float gc = gk[q] + bc[c];
                            const float pc = w[c] * v[q][c];
                            A = fmaf(pc, fabsf(gc), A);
                            Gx = fmaf(pc, gc, Gx);
                        }
                        dn = 0.5f * (A - Gx);
                        dm = S - A;
                        dp = 0.5f * (A + Gx);
                    }
                    if (s >= 2 && s - 2 < RJ) acc[s - 2] += dn;
                    if (s >= 1 && s - 1 < RJ) acc[s - 1] += dm;
                    if (s < RJ) acc[s] += dp;
                }
            }
        };
        auto by_nc = [&](auto edge_tag) {
            switch (nct) {
            case 3: slots(edge_tag, std::integral_constant<int, 3>{}); break;
            case 4: slots(edge_tag, std::integral_constant<int, 4>{}); break;
            case 5: slots(edge_tag, std::integral_constant<int, 5>{}); break;
            default: slots(edge_tag, std::integral_constant<int, 6>{}); break;
            }
        };
        if (mode == 2) by_nc(std::false_type{});
        else by_nc(std::true_type{});
    }
    if (i >= g.X) return;
    float *out = nzc > 1 ? g.part + ((size_t)zc * ntheta + (th - theta0)) * g.Y * g.X
                         : (ONES ? g.rwt : g.rval) + (size_t)th * g.Y * g.X;
#pragma unroll
    for (int r = 0; r < RJ; ++r)
        if (j0 + r < g.Y) out[(size_t)(j0 + r) * g.X + i] = acc[r];
}

// Forward projection as the reference computes it, a bilinear splat of
// every voxel (Reproject, main-mpi.cpp:286-344), for tilts with a1 >= 1/3
// and 0.55 <= b2 <= 1: any pitch with or without an offset.  A wave owns two
// neighbouring tiles of SP_R detector rows by t.si columns of one tilt,
// lanes 0-31 the first and lanes 32-63 the second, and sums them in LDS over
// its z layers.  Per layer each half walks the box of voxels whose taps can
// reach its tile, one volume row of up to 32 ncp voxels per pass (loads
// coalesced per half), and every voxel adds its four weighted taps to the
// tile.  A tap outside the tile lands in padding that is never stored (the
// tap position is clamped into the padded tile, which moves no voxel onto a
// stored cell), so the box needs no exact edge.  The voxel's position is
// ValCoef's u, v (main-mpi.cpp:204-226) in double at the box's first row of
// each lane's column and advanced in float over the rows; its weights are
// the reference's (1 - xd, xd) x (1 - yd, yd).
//
// No atomics: a pass takes its voxels in t.ncp phases, and in phase p lane l
// of a half holds voxel ncp l + p, so the voxels of one read-add-write lie
// ncp columns apart, at least 1 + 1e-4 pixels in u, and no two lanes of a
// half touch the same cell; every lane is busy in every phase.  Every cell
// therefore receives its adds in program order (layers, rows, phases, taps)
// on every run.  (LDS float atomics would need no phases, but gfx950 runs
// them lane by lane, some 30 times slower.)
constexpr int SP_R = 16;      // detector rows per tile (SR): 8, 24 and 32 were slower
constexpr int SP_CP = 6;      // padding columns on each side of a tile row
constexpr int SP_RP = 2;      // padding rows above and below
constexpr int SP_PITCH = 72;  // LDS floats per tile row
constexpr int SP_SIMAX = SP_PITCH - 2 * SP_CP;
static_assert(SP_SIMAX <= 64, "a tile row is stored by one wave");
template <int SR, bool ONES>
__global__ __launch_bounds__(64) void k_general_splat(const GeneralJob g, int theta0, int ntheta, int nzc,
                                                      unsigned ntile)
{
    constexpr int SP_ROWS = SR + 2 * SP_RP;
    __shared__ float acc[2][SP_ROWS * SP_PITCH];
    const unsigned w = xcd_order(blockIdx.x, gridDim.x);
    const unsigned tw = w % ntile, rest = w / ntile;
    const unsigned nrb = (g.Y + SR - 1) / SR;
    const int rb = (int)(rest % nrb), zc = (int)(rest / nrb);
    // the tilt of tile pair tw: the last whose first pair is at most tw
    const int base = g.tg[theta0].tile0;
    int lo = theta0, hi = theta0 + ntheta - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((unsigned)(g.tg[mid].tile0 - base) <= tw) lo = mid;
        else hi = mid - 1;
    }
    const int th = lo;
    const TiltGeneral &t = g.tg[th];
    const int si = t.si, ncp = t.ncp;
    // lanes 0-31 take the pair's first tile, lanes 32-63 its second
    const int lane = (int)threadIdx.x, half = lane >> 5, hl = lane & 31;
    const int ia = (int)(2 * (tw - (unsigned)(t.tile0 - base))) * si;  // the pair's first column
    const int i0 = ia + half * si, i1 = min(g.X, i0 + si);              // this lane's tile
    const int j0 = rb * SR, j1 = min(g.Y, j0 + SR);
    const int zlen = (g.T + nzc - 1) / nzc;
    const int z0 = zc * zlen, z1 = min(g.T, z0 + zlen);
    for (int k = lane; k < 2 * SP_ROWS * SP_PITCH; k += 64) (&acc[0][0])[k] = 0.f;
    __syncthreads();
    float *tile = acc[half];

    const float oxf = (float)t.ox, oyf = (float)t.oy, ozf = (float)t.oz;
    const double oxd = oxf, oyd = oyf;
    const double a1 = t.a[1], a2 = t.a[2], b1 = t.b[1], b2 = t.b[2];
    const float a2f = (float)a2, c2f = (float)(1.0 - b2);  // v advances by 1 - c2 per row
    // b1 X over the detector's columns
    const double bxa = b1 * (double)(0.f - oxf), bxb = b1 * (double)((float)(g.X - 1) - oxf);
    const double bl = fmin(bxa, bxb), bh = fmax(bxa, bxb);
    const size_t sy = g.layout == TREC_LAYOUT_Y ? (size_t)g.T * g.X : (size_t)g.X;
    const size_t sz = g.layout == TREC_LAYOUT_Y ? (size_t)g.X : (size_t)g.Y * g.X;
    for (int z = z0; z < z1; ++z) {
        const double Zc = (double)((float)z - ozf);
        const double cu = t.a[0] + t.a[3] * Zc + oxd;  // u = cu + a1 X + a2 Y
        const double cv = t.b[0] + t.b[3] * Zc + oyd;  // v = cv + b1 X + b2 Y
        // rows whose v can fall in [j0 - 1, j1), columns whose u can fall in
        // [i0 - 1, i1) on those rows (per half).  A voxel the rounding of
        // these bounds leaves out lies within rounding of the tile's edge,
        // where its weight on the tile is of that order.
        const int ya = __builtin_amdgcn_readfirstlane(max((int)floor(((double)(j0 - 1) - cv - bh) / b2 + oyd), 0));
        const int yb = __builtin_amdgcn_readfirstlane(min((int)floor(((double)j1 - cv - bl) / b2 + oyd) + 1, g.Y));
        if (ya >= yb) continue;
        const double ya2 = a2 * (double)((float)ya - oyf), yb2 = a2 * (double)((float)(yb - 1) - oyf);
        const double al = fmin(ya2, yb2), ah = fmax(ya2, yb2);
        const int xa = max((int)floor(((double)(i0 - 1) - cu - ah) / a1 + oxd), 0);
        const int xb = i0 < g.X ? min((int)floor(((double)i1 - cu - al) / a1 + oxd) + 1, g.X) : xa;
        // passes of 32 ncp voxels per half, as many as the wider box needs
        const int need = (max(xb - xa, 0) + 32 * ncp - 1) / (32 * ncp);
        const int npass = max(__builtin_amdgcn_readlane(need, 0), __builtin_amdgcn_readlane(need, 32));
        const int nrow = yb - ya;
        const double Ya = (double)((float)ya - oyf);
        const float *layer = g.vol + (size_t)z * sz + (size_t)ya * sy;
        for (int ps = 0; ps < npass; ++ps) {
            // phase p: voxel xa + 32 ncp ps + ncp hl + p of each half; the
            // voxels of a half lie ncp apart, at least 1 + 1e-4 pixels in u
            for (int p = 0; p < ncp; ++p) {
                const int x = xa + 32 * ncp * ps + ncp * hl + p;
                const bool in = x < xb;
                const double Xc = (double)((float)x - oxf);
                const double U0 = cu + a1 * Xc + a2 * Ya, V0 = cv + b1 * Xc + b2 * Ya;
                const double fU = floor(U0), fV = floor(V0);
                const float fu0 = (float)(U0 - fU), fv0 = (float)(V0 - fV);
                // LDS column and row of the first row's upper-left tap
                const int cu0 = (int)fU - i0 + SP_CP, rv0 = (int)fV - j0 + SP_RP;
                // row bases are uniform (scalar registers); a load adds the lane's column
                const unsigned xoff = (unsigned)min(x, g.X - 1) * 4u;
                for (int k0 = 0; k0 < nrow; k0 += 4) {
                    float vals[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        vals[q] = (ONES || k0 + q >= nrow)
                                      ? 1.f
                                      : *reinterpret_cast<const float *>(
                                            reinterpret_cast<const char *>(layer + (size_t)(k0 + q) * sy) + xoff);
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int k = k0 + q;
                        if (k >= nrow) break;
                        // lanes past their box add nothing
                        const float val = in ? vals[q] : 0.f;
                        const float fu = fmaf(a2f, (float)k, fu0), fv = fmaf(-c2f, (float)k, fv0);
                        const float flu = floorf(fu), flv = floorf(fv);
                        const float fx = fu - flu, fy = fv - flv;
                        const int cc = min(max(cu0 + (int)flu, 0), SP_PITCH - 2);
                        const int rr = min(max(rv0 + k + (int)flv, 0), SP_ROWS - 2);
                        float *cell = tile + rr * SP_PITCH + cc;
                        const float t1 = val * fy, t0 = val - t1;
                        const float w01 = t0 * fx, w00 = t0 - w01, w11 = t1 * fx, w10 = t1 - w11;
                        // the left taps of both rows, then the right ones:
                        // within each group the lanes of a half hit distinct
                        // cells of their own tile, and each group's reads see
                        // the writes before them (a wave's LDS accesses run in
                        // order; the wave barriers keep the compiler from
                        // merging the groups, as lane x's right tap may be
                        // lane x + 1's left one)
                        cell[0] += w00;
                        cell[SP_PITCH] += w10;
                        __builtin_amdgcn_wave_barrier();
                        cell[1] += w01;
                        cell[SP_PITCH + 1] += w11;
                        __builtin_amdgcn_wave_barrier();
                    }
                }
            }
        }
    }
    __syncthreads();
    float *out = nzc > 1 ? g.part + ((size_t)zc * ntheta + (th - theta0)) * g.Y * g.X
                         : (ONES ? g.rwt : g.rval) + (size_t)th * g.Y * g.X;
    // each half stores its tile, 32 columns per row at a time
    for (int c = hl; c < i1 - i0; c += 32)
        for (int r = 0; r < j1 - j0; ++r) out[(size_t)(j0 + r) * g.X + i0 + c] = tile[(r + SP_RP) * SP_PITCH + SP_CP + c];
}

// Back-projection for tilts that share b and the origins.  Each thread takes
// GB_Q voxels 64 apart along x in one (y, z) row of the volume (one row per
// wave).  v is evaluated once per voxel with ValCoef's operation order, u
// once per voxel and tilt.  The bilinear taps are read as two column pairs
// (c0, c0 + 1) of rows (r0, r0 + 1), with c0 = min(max(floor u, 0), X - 2)
// and r0 likewise, and each tap is weighted hat(u - c) hat(v - r): inside
// the projection that is (1 - xd, xd) x (1 - yd, yd), and a tap the
// reference skips because it lies outside gets 0 (its neighbour inside
// keeps its weight), so no tap is ever read out of bounds.
constexpr int GB_Q = 4;
template <int MODE>
__global__ __launch_bounds__(WG) void k_general_back_rows(const GeneralJob g, int theta0, int ntheta,
                                                          const float *__restrict__ img)
{
    const unsigned nxt = (g.X + GB_Q * 64 - 1) / (GB_Q * 64);
    const unsigned row = __builtin_amdgcn_readfirstlane((blockIdx.x / nxt) * (WG / 64) + threadIdx.x / 64);
    if (row >= (unsigned)g.Y * (unsigned)g.T) return;
    const int xb = (int)(blockIdx.x % nxt) * GB_Q * 64 + (int)(threadIdx.x & 63);
    const int y = g.layout == TREC_LAYOUT_Y ? (int)(row / g.T) : (int)(row % g.Y);
    const int z = g.layout == TREC_LAYOUT_Y ? (int)(row % g.T) : (int)(row / g.Y);
    const TiltGeneral &t0 = g.tg[theta0];
    const double Yc = (double)((float)y - (float)t0.oy);
    const double Zc = (double)((float)z - (float)t0.oz);
    const unsigned X = g.X, Y = g.Y;
    double Xc[GB_Q];
    float wy0[GB_Q], wy1[GB_Q], val[GB_Q], wt[GB_Q], out[GB_Q];
    unsigned o0[GB_Q], o1[GB_Q];  // byte offsets of rows r0, r0 + 1 in a projection
    float *vout = g.vol + (size_t)row * g.X;
#pragma unroll
    for (int q = 0; q < GB_Q; ++q) {
        const int x = xb + 64 * q;
        Xc[q] = (double)((float)x - (float)t0.ox);
        const double n1 = __dadd_rn(__dadd_rn(__dadd_rn(t0.b[0], __dmul_rn(t0.b[1], Xc[q])), __dmul_rn(t0.b[2], Yc)),
                                    __dmul_rn(t0.b[3], Zc));
        const double v = __dadd_rn(n1, t0.oy);
        const int r0 = min(max((int)v, 0), (int)Y - 2);
        const float dy = (float)(v - (double)r0);
        wy0[q] = hat(dy);
        wy1[q] = hat(dy - 1.f);
        o0[q] = (unsigned)r0 * X * 4u;
        o1[q] = o0[q] + X * 4u;
        val[q] = 0.f;
        wt[q] = 0.f;
        out[q] = x < g.X ? vout[x] : 0.f;
    }
    for (int th = theta0; th < theta0 + ntheta; ++th) {
        const TiltGeneral &t = g.tg[th];
        const double ub = t.a[0] + t.a[2] * Yc + t.a[3] * Zc + t.ox;  // u = ub + a1 X
        const double a1 = t.a[1];
        const char *im = (const char *)(img + (size_t)th * Y * X);
#pragma unroll
        for (int q = 0; q < GB_Q; ++q) {
            const double u = a1 * Xc[q] + ub;
            const int c0 = min(max((int)u, 0), (int)X - 2);
            const float dx = (float)(u - (double)c0);
            const float w0 = hat(dx), w1 = hat(dx - 1.f);
            const unsigned oc = (unsigned)c0 * 4u;
            const float2 p0 = *(const float2 *)(im + (o0[q] + oc));
            const float2 p1 = *(const float2 *)(im + (o1[q] + oc));
            const float s = wy0[q] * fmaf(w1, p0.y, w0 * p0.x) + wy1[q] * fmaf(w1, p1.y, w0 * p1.x);
            const float c = (w0 + w1) * (wy0[q] + wy1[q]);
            if (MODE == BACK_SIRT || MODE == BACK_LIN) {
                val[q] += s;
                wt[q] += c;
            } else if (c > g.thr_upd) {
                out[q] += (MODE == BACK_SART) ? (s / c) * g.gamma : s / c;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < GB_Q; ++q) {
        const int x = xb + 64 * q;
        if (MODE == BACK_LIN) {
            if (x < g.X) g.out[(size_t)row * g.X + x] = g.lin_a * val[q] + g.lin_b * out[q];
            continue;
        }
        if (MODE == BACK_SIRT && wt[q] > g.thr_upd) out[q] += val[q] / wt[q] * g.gamma;
        if (x < g.X) vout[x] = out[q];
    }
}

// Back-projection for tilts that share b, with the taps taken from LDS.
// With b1 = 0 (NR = 2) v, and with it the two detector rows r0, r0 + 1 and
// their weights, is the same for every voxel of a wave's (y, z) row.  With
// b1 != 0 v moves by |b1| per voxel, less than NR - 2 rows over the wave's
// GL_Q * 64 voxels, so their rows lie among the NR rows from the least r0;
// each voxel keeps its own rows and weights (v does not depend on the tilt).
// Per tilt the wave stages the columns of those rows that its voxels reach
// into its own LDS, with coalesced loads: from c = min(max(trunc(u), 0),
// X - 2) at its least u, at most (64 GL_Q - 1) |a1| + 3 <= GL_W columns.  It
// takes its taps from there: half the bytes of reading them from memory,
// where neighbouring lanes read the same columns again.  The next tilt's
// window is loaded while the current one is used.  The weights are
// k_general_back_rows'; SIRT and ADMM sum each voxel's two detector rows'
// column-weighted taps over the tilts and apply the row weights, the same in
// every tilt, once per voxel.
//
// SH: the four waves of a work-group take four neighbouring layers of one
// volume row (layout Y with T a multiple of 4), whose detector rows and
// columns nearly coincide: v moves by b3 and u by a3 per layer.  The group
// then stages the union of their windows once per tilt, NR + 1 rows (3 |b3|
// < 1) of at most GL_WS columns, double-buffered with one barrier per tilt,
// and each wave takes its taps from its own part of it.  Every wave
// evaluates the union itself, with the same expressions, so they agree.
// With two staged rows and the shared window the kernel runs five waves per
// SIMD (96 VGPRs, a few spilled), which hides more of the barrier per tilt
// than four; with three rows four, and six waves or the per-wave window at
// five spill more than they gain.
constexpr int GL_Q = 8;        // voxels per lane, 64 apart
constexpr int GL_BUF_DWORD3 = 0x00020000;  // raw buffer resource, gfx9 family
constexpr int GL_W = 516;      // staged columns per row
constexpr int GL_PITCH = 520;  // floats between two staged rows
constexpr int GL_WS = 528;     // columns of a shared window: GL_W plus the shift of 3 layers, 3 |a3| <= 3
template <int MODE, int NR, bool SH>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(NR <= 2 && SH ? 5 : (NR <= 3 ? 4 : 3)))) void k_general_back_lds(const GeneralJob g, int theta0, int ntheta,
                                                         const float *__restrict__ img)
{
    constexpr int NRS = SH ? NR + 1 : NR;        // staged rows
    constexpr int PITCH = SH ? GL_WS : GL_PITCH;  // floats between two staged rows
    constexpr int NBUF = SH ? 2 : WG / 64;        // shared: two buffers; else one per wave
    __shared__ __attribute__((aligned(16))) float stage[NBUF][NRS * PITCH];
    const unsigned nxt = (g.X + GL_Q * 64 - 1) / (GL_Q * 64);
    const unsigned wave = threadIdx.x / 64;
    const unsigned rows = (unsigned)g.Y * (unsigned)g.T;
    const unsigned row0 = __builtin_amdgcn_readfirstlane((blockIdx.x / nxt) * (WG / 64));  // the group's first row
    const unsigned row_w = __builtin_amdgcn_readfirstlane(row0 + wave);
    // a wave past the last row shares the group's barriers but computes nothing
    if (!SH && row_w >= rows) return;  // the whole wave: it shares nothing with the others
    const bool active = row_w < rows;
    const unsigned row = min(row_w, rows - 1);
    const int lane = (int)(threadIdx.x & 63);
    const int xb0 = (int)(blockIdx.x % nxt) * GL_Q * 64;
    const int y = g.layout == TREC_LAYOUT_Y ? (int)(row / g.T) : (int)(row % g.Y);
    const int z = g.layout == TREC_LAYOUT_Y ? (int)(row % g.T) : (int)(row / g.Y);
    const TiltGeneral &t0 = g.tg[theta0];
    const double Yc = (double)((float)y - (float)t0.oy);
    const double Zc = (double)((float)z - (float)t0.oz);
    const int X = g.X, Y = g.Y;
    // ValCoef's v, n1 = ((b0 + b1 X) + b2 Y) + b3 Z and v = n1 + oy; its
    // double rounding is monotone in X, so the wave's least row comes from
    // its first or its last voxel
    auto vofr = [&](double Xc, double Ycr, double Zcr) {
        return __dadd_rn(
            __dadd_rn(__dadd_rn(__dadd_rn(t0.b[0], __dmul_rn(t0.b[1], Xc)), __dmul_rn(t0.b[2], Ycr)),
                      __dmul_rn(t0.b[3], Zcr)),
            t0.oy);
    };
    auto vof = [&](double Xc) { return vofr(Xc, Yc, Zc); };
    const double X0 = (double)((float)xb0 - (float)t0.ox);                  // X of the wave's first voxel
    const double X1 = (double)((float)(xb0 + GL_Q * 64 - 1) - (float)t0.ox);  // and of its last
    const int rlo = min(max((int)fmin(vof(X0), vof(X1)), 0), Y - 2);
    // the group's rows (SH): Yc and Zc of its wave k (formed where they are
    // used, so that they hold no registers across the tilt loop), and the
    // least row of all
    auto ycg = [&](int k) { return (double)((float)(int)(min(row0 + k, rows - 1) / g.T) - (float)t0.oy); };
    auto zcg = [&](int k) { return (double)((float)(int)(min(row0 + k, rows - 1) % g.T) - (float)t0.oz); };
    int rlo_g = rlo;
    if constexpr (SH) {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            rlo_g = min(rlo_g, min(max((int)fmin(vofr(X0, ycg(k), zcg(k)), vofr(X1, ycg(k), zcg(k))), 0), Y - 2));
    }
    unsigned orow[NRS];  // staged rows rlo_g .. rlo_g + NRS - 1 (clamped into the projection)
#pragma unroll
    for (int k = 0; k < NRS; ++k) orow[k] = (unsigned)min(rlo_g + k, Y - 1) * (unsigned)X;
    // per voxel: the staged offset of row r0 (0 with NR = 2) and the weights
    // of r0, r0 + 1, which the tilt loop does not need: they are evaluated
    // again at the end.  BPT and SART update the volume per tilt and keep it
    // in registers; SIRT and ADMM read it at the end.
    constexpr bool PER_TILT = MODE == BACK_BPT || MODE == BACK_SART;
    auto row_weights = [&](int q, float &w0, float &w1) {
        const double v = NR == 2 ? vof(X0) : vof((double)((float)(xb0 + lane + 64 * q) - (float)t0.ox));
        const int r0 = min(max((int)v, 0), Y - 2);
        const float dy = (float)(v - (double)r0);
        w0 = hat(dy);
        w1 = hat(dy - 1.f);
        return r0;
    };
    const double Xc0 = (double)((float)(xb0 + lane) - (float)t0.ox);  // X of the lane's first voxel
    float h0[GL_Q], h1[GL_Q], hw[GL_Q], out[PER_TILT ? GL_Q : 1], wy0[PER_TILT ? GL_Q : 1], wy1[PER_TILT ? GL_Q : 1];
    int ro[GL_Q];
    float *vout = g.vol + (size_t)row * g.X;
#pragma unroll
    for (int q = 0; q < GL_Q; ++q) {
        float w0, w1;
        const int r0 = row_weights(q, w0, w1);
        ro[q] = (rlo - rlo_g + (NR == 2 ? 0 : min(max(r0 - rlo, 0), NR - 2))) * PITCH;
        if constexpr (PER_TILT) {
            const int x = xb0 + lane + 64 * q;
            wy0[q] = w0;
            wy1[q] = w1;
            out[q] = x < X ? vout[x] : 0.f;
        }
        h0[q] = 0.f;
        h1[q] = 0.f;
        hw[q] = 0.f;
    }
    // window of tilt th: u = ub + a1 X, its first column, and the staged
    // values of columns cmin + lane + 64 m.
    static_assert(GL_W <= GL_PITCH && GL_W + 12 <= GL_WS, "staged columns fit a row");
    // the window is staged in chunks of 4 columns, at most GC per row, of
    // which lane l loads chunks l, l + 64, ... (SH: the group's window, at
    // most GCS chunks per row, of which thread t loads chunks t, t + WG, ...
    // counted row by row)
    constexpr int GC = GL_W / 4, GK = (GC + 63) / 64;
    constexpr int GCS = GL_WS / 4, GKS = (NRS * GCS + WG - 1) / WG;
    static_assert(GL_W % 4 == 0 && GL_PITCH % 4 == 0 && GL_WS % 4 == 0, "16-byte chunks");
    float4 sv[SH ? 1 : NR][SH ? GKS : GK];
    // Constants of tilts tb .. tb + 63, lane l holding tilt tb + l: u = ub +
    // a1 X, the window's first column cmin, whether every voxel of the wave
    // lies inside the detector (1 <= u <= X - 3), 64 a1 split into its
    // integer and fractional parts, and the chunks the taps can reach
    // (columns up to (64 GL_Q - 1) |a1| + 3 past cmin, with a chunk of
    // margin).  Each tilt reads them with v_readlane.
    // SH: also the group window's first column cu and its chunks per row.
    double ub_v = 0.0, a1_v = 0.0;
    int cmin_v = 0, in_v = 0, ia_v = 0, nch_v = 0, cu_v = 0, nchu_v = 0;
    float fa_v = 0.f;
    auto tilt_block = [&](int tb) {
        const TiltGeneral &t = g.tg[min(tb + lane, theta0 + ntheta - 1)];
        ub_v = t.a[0] + t.a[2] * Yc + t.a[3] * Zc + t.ox;
        a1_v = t.a[1];
        const double u0 = __fma_rn(a1_v, X0, ub_v), u1 = __fma_rn(a1_v, X1, ub_v);
        cmin_v = min(max((int)fmin(u0, u1), 0), X - 2);
        in_v = fmin(u0, u1) >= 1.0 && fmax(u0, u1) <= (double)(X - 3);
        const double A = 64.0 * a1_v, fA_ = floor(A);
        ia_v = (int)fA_;
        fa_v = (float)(A - fA_);
        nch_v = min(GC, (int)(((64.0 * GL_Q - 1.0) * fabs(a1_v) + 4.0) * 0.25) + 2);
        if constexpr (SH) {
            int lo = X, hi = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const double ubk = t.a[0] + t.a[2] * ycg(k) + t.a[3] * zcg(k) + t.ox;
                const double v0 = __fma_rn(a1_v, X0, ubk), v1 = __fma_rn(a1_v, X1, ubk);
                const int ck = min(max((int)fmin(v0, v1), 0), X - 2);
                lo = min(lo, ck);
                hi = max(hi, ck);
            }
            cu_v = lo;
            nchu_v = min(GCS, (hi - lo + 3) / 4 + nch_v);
        }
    };
    auto lane_d = [](double v, int l) {
        const unsigned long long b = (unsigned long long)__double_as_longlong(v);
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
        return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    };
    // buffer loads: the row base as the scalar offset, the lane and chunk in
    // the instruction; columns past the row (never used: taps stop at X - 1)
    // read the next row, or 0 past the projection
    auto load = [&](int th, int cnext, int nch) {
        const auto rs = __builtin_amdgcn_make_buffer_rsrc((void *)(img + (size_t)th * Y * X), (short)0,
                                                          (int)((unsigned)Y * (unsigned)X * 4u), GL_BUF_DWORD3);
        if constexpr (SH) {
#pragma unroll
            for (int m = 0; m < GKS; ++m) {
                const int idx = (int)threadIdx.x + WG * m, r = idx / GCS, c = idx % GCS;
                if (r < NRS && c < nch) {
                    const auto v = __builtin_amdgcn_raw_buffer_load_b128(
                        rs, (int)((orow[min(r, NRS - 1)] + (unsigned)(cnext + 4 * c)) * 4u), 0, 0);
                    sv[0][m] = make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                                           __uint_as_float(v[3]));
                }
            }
            return;
        }
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const int soff = (int)((orow[k] + (unsigned)cnext) * 4u);
#pragma unroll
            for (int m = 0; m < GK; ++m) {
                if (lane + 64 * m < nch) {
                    const auto r = __builtin_amdgcn_raw_buffer_load_b128(rs, lane * 16 + 1024 * m, soff, 0);
                    sv[k][m] = make_float4(__uint_as_float(r[0]), __uint_as_float(r[1]), __uint_as_float(r[2]),
                                           __uint_as_float(r[3]));
                }
            }
        }
    };
    // the window a load stages: the wave's own, or the group's (SH)
    auto wcol = [&](int l) { return __builtin_amdgcn_readlane(SH ? cu_v : cmin_v, l); };
    auto wch = [&](int l) { return __builtin_amdgcn_readlane(SH ? nchu_v : nch_v, l); };
    int tb = theta0;
    tilt_block(tb);
    load(theta0, wcol(0), wch(0));
    int nin = 0;  // SIRT and ADMM: tilts in which every voxel of the wave lies inside
    for (int th = theta0; th < theta0 + ntheta; ++th) {
        const int tl = th - tb;
        const int cmin = __builtin_amdgcn_readlane(cmin_v, tl);
        const bool inside = __builtin_amdgcn_readlane(in_v, tl) != 0;
        const int IA = __builtin_amdgcn_readlane(ia_v, tl);
        const float fA = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fa_v), tl));
        const int nch = wch(tl);
        const double ub = lane_d(ub_v, tl), a1 = lane_d(a1_v, tl);
        float *lds;
        if constexpr (SH) {
            // buffer (th - theta0) & 1: the group finished reading it two
            // tilts ago, before the barrier of the last tilt
            float *buf = stage[(th - theta0) & 1];
#pragma unroll
            for (int m = 0; m < GKS; ++m) {
                const int idx = (int)threadIdx.x + WG * m, r = idx / GCS, c = idx % GCS;
                if (r < NRS && c < nch) *reinterpret_cast<float4 *>(buf + r * PITCH + 4 * c) = sv[0][m];
            }
            __syncthreads();  // the group's window of tilt th is staged
            lds = buf + (cmin - wcol(tl));
        } else {
            lds = stage[wave];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int m = 0; m < GK; ++m) {
                if (lane + 64 * m < nch) {
#pragma unroll
                    for (int k = 0; k < NR; ++k) *reinterpret_cast<float4 *>(lds + k * PITCH + 4 * (lane + 64 * m)) = sv[k][m];
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (th + 1 < theta0 + ntheta) {
            if (th + 1 - tb == 64) {
                tb = th + 1;
                tilt_block(tb);
            }
            load(th + 1, wcol(th + 1 - tb), wch(th + 1 - tb));
        }
        // u - cmin = ubr + a1 X; the staged column li = c0 - cmin of tap c0 =
        // min(max(trunc u, 0), X - 2), which is at most (64 GL_Q - 1) |a1| + 1
        // (the bound GL_W - 2 only guards it)
        const double ubr = ub - (double)cmin;
        const int lhi = min(X - 2 - cmin, GL_W - 2);
        // u - cmin of voxel q is U0 + q A, A = 64 a1: in float from the
        // integer and fractional parts of U0 and A (|f0 + q fA| < 8, so the
        // fraction is good to about 5e-7).  Inside the detector (every voxel
        // of the wave at 1 <= u <= X - 3) the tap is floor(u), at most
        // (64 GL_Q - 1) |a1| + 1 columns past cmin, with the weights 1 - dx and
        // dx, which sum to 1 (SIRT and ADMM count those tilts instead of
        // adding them); at its edges the tap is clamped into [0, X - 2] and
        // the fraction measured from there.
        const double U0 = __fma_rn(a1, Xc0, ubr);
        const double fU0 = floor(U0);
        const int I0 = (int)fU0;
        const float f0 = (float)(U0 - fU0);
        auto voxels = [&](auto edge_tag) {
            constexpr bool EDGE = decltype(edge_tag)::value;
#pragma unroll
            for (int q = 0; q < GL_Q; ++q) {
                // tq >= 0 (f0 and fA are fractions), so truncation is its floor
                const float tq = fmaf((float)q, fA, f0);
                const int base = I0 + q * IA;
                const int lr = q * IA + (int)tq;  // the tap's column past I0
                const int li = EDGE ? min(max(I0 + lr, 0), lhi) : I0 + lr;
                const float dx = EDGE ? tq - (float)(li - base) : __builtin_amdgcn_fractf(tq);
                const float w0 = EDGE ? hat(dx) : 1.f - dx, w1 = EDGE ? hat(dx - 1.f) : dx;
                const float *l0 = EDGE ? lds + ro[q] + li : (lds + ro[q] + I0) + lr;
                if constexpr (!PER_TILT) {
                    h0[q] = fmaf(w1, l0[1], fmaf(w0, l0[0], h0[q]));
                    h1[q] = fmaf(w1, l0[PITCH + 1], fmaf(w0, l0[PITCH], h1[q]));
                    if (EDGE) hw[q] += w0 + w1;
                } else {
                    const float s = wy0[q] * fmaf(w1, l0[1], w0 * l0[0]) +
                                    wy1[q] * fmaf(w1, l0[PITCH + 1], w0 * l0[PITCH]);
                    const float c = (w0 + w1) * (wy0[q] + wy1[q]);
                    if (c > g.thr_upd) out[q] += (MODE == BACK_SART) ? (s / c) * g.gamma : s / c;
                }
            }
        };
        if (!active) {
        } else if (inside) {
            voxels(std::false_type{});
            ++nin;
        } else {
            voxels(std::true_type{});
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (!active) return;
#pragma unroll
    for (int q = 0; q < GL_Q; ++q) hw[q] += (float)nin;
#pragma unroll
    for (int q = 0; q < GL_Q; ++q) {
        const int x = xb0 + lane + 64 * q;
        if constexpr (PER_TILT) {
            if (x < X) vout[x] = out[q];
        } else {
            float w0, w1;
            row_weights(q, w0, w1);
            const float vq = x < X ? vout[x] : 0.f;
            if (MODE == BACK_LIN) {
                if (x < X) g.out[(size_t)row * g.X + x] = g.lin_a * (w0 * h0[q] + w1 * h1[q]) + g.lin_b * vq;
            } else {
                const float val = w0 * h0[q] + w1 * h1[q], wt = (w0 + w1) * hw[q];
                if (x < X) vout[x] = wt > g.thr_upd ? vq + val / wt * g.gamma : vq;
            }
        }
    }
}

// ------------------------------------------------------------------ launchers

namespace {

bool same_rows(const TiltGeneral &a, const TiltGeneral &b)
{
    return std::memcmp(a.b, b.b, sizeof a.b) == 0 && a.ox == b.ox && a.oy == b.oy && a.oz == b.oz;
}

// three: k_general_rows3 (a1 > 0 in every tilt), else k_general_rows
template <int RJ, int NC>
void launch_rows(const GeneralJob &g, int theta0, int ntheta, bool ones, bool three, int nzc, unsigned nblk,
                 hipStream_t st)
{
    if (three && ones)
        hipLaunchKernelGGL((k_general_rows3<RJ, NC, true>), dim3(nblk), dim3(WG), 0, st, g, theta0, ntheta, nzc);
    else if (three)
        hipLaunchKernelGGL((k_general_rows3<RJ, NC, false>), dim3(nblk), dim3(WG), 0, st, g, theta0, ntheta, nzc);
    else if (ones)
        hipLaunchKernelGGL((k_general_rows<RJ, NC, true>), dim3(nblk), dim3(WG), 0, st, g, theta0, ntheta, nzc);
    else
        hipLaunchKernelGGL((k_general_rows<RJ, NC, false>), dim3(nblk), dim3(WG), 0, st, g, theta0, ntheta, nzc);
}

template <int RJ>
void launch_rows_nc(const GeneralJob &g, int theta0, int ntheta, bool ones, bool three, int nc, int nzc,
                    unsigned nblk, hipStream_t st)
{
    switch (nc) {
    case 2: launch_rows<RJ, 2>(g, theta0, ntheta, ones, three, nzc, nblk, st); break;
    case 3: launch_rows<RJ, 3>(g, theta0, ntheta, ones, three, nzc, nblk, st); break;
    case 4: launch_rows<RJ, 4>(g, theta0, ntheta, ones, three, nzc, nblk, st); break;
    default: launch_rows<RJ, 6>(g, theta0, ntheta, ones, three, nzc, nblk, st); break;
    }
}

template <int RJ, bool LANE>
void launch_rows3b(const GeneralJob &g, int theta0, int ntheta, bool ones, int nzc, unsigned nblk, double dlo,
                   int dmax, hipStream_t st)
{
    if (ones)
        hipLaunchKernelGGL((k_general_rows3b<RJ, true, LANE>), dim3(nblk), dim3(WG), 0, st, g, theta0, ntheta, nzc,
                           dlo, dmax);
    else
        hipLaunchKernelGGL((k_general_rows3b<RJ, false, LANE>), dim3(nblk), dim3(WG), 0, st, g, theta0, ntheta, nzc,
                           dlo, dmax);
}

template <int RJ>
void launch_rows3b_lane(const GeneralJob &g, int theta0, int ntheta, bool ones, int nzc, unsigned nblk, double dlo,
                        int dmax, bool lane, hipStream_t st)
{
    if (lane) launch_rows3b<RJ, true>(g, theta0, ntheta, ones, nzc, nblk, dlo, dmax, st);
    else launch_rows3b<RJ, false>(g, theta0, ntheta, ones, nzc, nblk, dlo, dmax, st);
}

// candidate columns a row window needs at a1: 2 / a1 rounded up (a1 just
// under 1/2, as at 60 degrees with an offset, needs 5)
int rows_nc(double aa1) { return std::max(2, (int)std::ceil(2.0 / aa1 - 1e-6)); }

// A bound on |b1 (x - xr)| over the candidate columns x of one wave of
// k_general_rows3b when its lanes take the wave's rows (xr its reference
// column).  The candidates span at most 63 / a1 + 2 drift + NC columns; xr,
// taken at the rows near j0, sits within |a2 / a1| |Ye - Yc0| + 1 / a1 +
// (NC + 3) / 2 of their middle, and |Ye - Yc0| <= (1 + |b1 Xr|) / b2 + 1.
double rows3b_wave_bound(const TiltGeneral &t, int RJ, int NC, int X)
{
    const double a1 = t.a[1], ar = std::fabs(t.a[2] / a1), b1 = std::fabs(t.b[1]), b2 = t.b[2];
    const double drift = ar * RJ + 0.01;
    const double span = 63.0 / a1 + 2.0 * drift + NC;
    const double xr_max = (double)X + std::fabs((double)t.ox) + span + 10.0;  // |Xr| of a wave that is used
    const double off = ar * ((1.0 + b1 * xr_max) / b2 + 1.0) + 1.0 / a1 + 0.5 * (NC + 3);
    return b1 * (0.5 * span + off + 1.0);
}

// The bounds of k_general_rows3b<..., true> (each lane's rows) with RJ rows
// and NC candidates in tilt t:
// dmax, the most a lane shifts its rows, and D, the most |b1 (x - X1)| over
// a lane's candidates x.  A wave's candidates span at most 63 / a1 + 2 drift
// + NC columns; its reference column xr, taken at the rows near j0, sits
// within |a2 / a1| |Ye - Yc0| + 1 / a1 + (NC + 3) / 2 of their middle, and
// |Ye - Yc0| <= (1 + |b1 Xr|) / b2 + 1.  So a lane's first candidate has g
// within |b1| (span / 2 + that) of the wave's g0, and |dl| is at most that
// over b2, plus one.  A lane's candidates lie within |a2 / a1| (RJ + 1) + NC
// of X1, which the row shift moves by at most |a2 / a1| dmax + 1 columns.
void rows3b_bounds(const TiltGeneral &t, int RJ, int NC, int X, double &d, int &dmax)
{
    const double a1 = t.a[1], ar = std::fabs(t.a[2] / a1), b1 = std::fabs(t.b[1]), b2 = t.b[2];
    const double drift = ar * RJ + 0.01;
    const double span = 63.0 / a1 + 2.0 * drift + NC;
    const double xr_max = (double)X + std::fabs((double)t.ox) + span + 10.0;  // |Xr| of a wave that is used
    const double off = ar * ((1.0 + b1 * xr_max) / b2 + 1.0) + 1.0 / a1 + 0.5 * (NC + 3);
    const double dl = b1 * (0.5 * span + off) / b2 + 1.0;
    dmax = (int)std::ceil(dl);
    d = b1 * (ar * (RJ + 1 + dmax) + NC + 2) + 1e-4;
}

template <int NR, bool SH>
void launch_back_lds_sh(const GeneralJob &g, int mode, int theta0, int ntheta, const float *img, dim3 grid,
                        hipStream_t st)
{
    if (mode == BACK_SIRT)
        hipLaunchKernelGGL((k_general_back_lds<BACK_SIRT, NR, SH>), grid, dim3(WG), 0, st, g, theta0, ntheta, img);
    else if (mode == BACK_LIN)
        hipLaunchKernelGGL((k_general_back_lds<BACK_LIN, NR, SH>), grid, dim3(WG), 0, st, g, theta0, ntheta, img);
    else if (mode == BACK_SART)
        hipLaunchKernelGGL((k_general_back_lds<BACK_SART, NR, SH>), grid, dim3(WG), 0, st, g, theta0, ntheta, img);
    else
        hipLaunchKernelGGL((k_general_back_lds<BACK_BPT, NR, SH>), grid, dim3(WG), 0, st, g, theta0, ntheta, img);
}

template <int NR>
void launch_back_lds(const GeneralJob &g, int mode, int theta0, int ntheta, const float *img, dim3 grid, bool sh,
                     hipStream_t st)
{
    if (sh) launch_back_lds_sh<NR, true>(g, mode, theta0, ntheta, img, grid, st);
    else launch_back_lds_sh<NR, false>(g, mode, theta0, ntheta, img, grid, st);
}

}  // namespace

bool general_tiles(TiltGeneral *tg, int P, int X)
{
    bool all = X >= 2;
    int tile = 0;
    for (int i = 0; i < P; ++i) {
        TiltGeneral &t = tg[i];
        const double a1 = t.a[1], b2 = t.b[2];
        t.si = t.ncp = t.ntx = 0;
        t.tile0 = tile;
        if (!(a1 >= 1.0 / 3.0 && b2 >= 0.55 && b2 <= 1.0) || X < 2) {
            all = false;
            continue;
        }
        // a tile's box: at most (SR + 1 + |b1| X) / b2 + 2 rows, and
        // (si + 1 + |a2| rows) / a1 + 2 columns, which should fit 64 lanes
        // (a wider box takes a second pass)
        const double rows = (SP_R + 1 + std::fabs(t.b[1]) * X) / b2 + 4.0;
        const int si = std::min(SP_SIMAX, (int)std::floor(a1 * 62.0 - 1.0 - std::fabs(t.a[2]) * rows));
        if (si < 8) {
            all = false;
            continue;
        }
        int ncp = 1;
        while (a1 * ncp < 1.0 + 1e-4) ++ncp;
        t.si = si;
        t.ncp = ncp;
        t.ntx = (X + 2 * si - 1) / (2 * si);  // pairs of tiles: a wave takes two
        tile += t.ntx;
    }
    return all;
}

// The row windows for a pitch with an offset (b1 != 0, the same b in every
// tilt, a1 >= 1/3, 1/2 <= b2 <= 1): k_general_rows3b with 16, 8 or 4 rows
// per thread (4 for SART's few-tilt launches) while e (RJ + 2) + 2 D <= 1,
// each tilt at its own candidate count (at least 3).  Per RJ the wave's
// rows where their bound holds, else each lane's (a few more instructions
// per layer, but |b1| up to about 0.05); TREC_GENERAL_ROWS_LANE=1 takes each
// lane's rows wherever they apply.
bool rows3b_plan(const TiltGeneral *host_tg, int theta0, int ntheta, int X, int T, int layout, Rows3bPlan &rp)
{
    rp = Rows3bPlan{};
    if (ntheta <= 0) return false;
    int ncb = 2;
    for (int th = theta0; th < theta0 + ntheta; ++th) {
        const TiltGeneral &t = host_tg[th];
        if (!(t.b[1] != 0.0 && t.b[2] <= 1.0 && t.b[2] >= 0.5 && same_rows(t, host_tg[theta0]) && t.a[1] >= 1.0 / 3.0))
            return false;
        ncb = std::max(ncb, rows_nc(t.a[1]));
    }
    if (X < std::max(3, ncb)) return false;
    const double e = 1.0 - host_tg[theta0].b[2];
    const size_t sy = layout == TREC_LAYOUT_Y ? (size_t)T * X : (size_t)X;
    const char *flane = getenv("TREC_GENERAL_ROWS_LANE");
    const bool force_lane = flane && !strcmp(flane, "1");
    for (int r = ntheta <= 4 ? 4 : 16; r >= 4; r /= 2) {
        double dw = 0.0, d = 0.0;
        int dm = 0;
        for (int th = theta0; th < theta0 + ntheta; ++th) {
            const int nct = std::max(3, rows_nc(host_tg[th].a[1]));
            double dt;
            int dmt;
            rows3b_bounds(host_tg[th], r, nct, X, dt, dmt);
            d = std::max(d, dt);
            dm = std::max(dm, dmt);
            dw = std::max(dw, rows3b_wave_bound(host_tg[th], r, nct, X));
        }
        // a lane's load offset, (dl + dmax + s) sy + xs floats, fits 32 bits
        const bool off32 = ((size_t)(2 * dm + r + 2) * sy + (size_t)X) * 4u <= 0xffffffffu;
        const bool off32w = ((size_t)(r + 2) * sy + (size_t)X) * 4u <= 0xffffffffu;
        if (!force_lane && e * (r + 2) + 2.0 * dw <= 1.0 - 1e-6 && off32w) {
            rp.rj = r;
            rp.dlo = -1.0 + e * r + dw;
            return true;
        }
        if (e * (r + 2) + 2.0 * d <= 1.0 - 1e-6 && off32) {
            rp.rj = r;
            rp.dlo = -1.0 + e * r + d;
            rp.dmax = dm;
            rp.lane = true;
            return true;
        }
    }
    return false;
}

hipError_t launch_general_forward(const GeneralJob &g, const TiltGeneral *host_tg, int theta0, int ntheta, bool ones,
                                  hipStream_t st, int *nzc_used)
{
    if (nzc_used) *nzc_used = 1;
    // TREC_GENERAL_FORWARD=rows|rows5|splat|gather puts that kernel first where it
    // applies (tests and timing).  By default: the row windows when v does
    // not depend on x or moves slowly along it, else the splat, else the
    // per-pixel gather.
    const char *force = getenv("TREC_GENERAL_FORWARD");
    const bool no_rows = force && !strcmp(force, "gather");
    const bool no_splat = force && (!strcmp(force, "gather") || !strcmp(force, "rows") || !strcmp(force, "rows5"));
    // a single-tilt value projection whose caller takes partials splits the
    // layers into chunks, for GZ_CHUNKS times the waves; the splat of all
    // tilts takes g.nzc_all chunks, for L2 locality
    const int nzc = nzc_used && g.part && ntheta == 1 && !ones ? std::min(GZ_CHUNKS, g.T) : 1;
    const int nzc_splat = nzc_used && g.part && !ones && ntheta == g.P && g.nzc_all > 1 ? g.nzc_all : nzc;
    // row windows: v independent of x and the same in every tilt launched,
    // 7/8 <= b2 <= 1, at most 6 candidate columns (|a1| >= 1/3)
    bool rows = ntheta > 0 && !no_rows;
    int nc = 2;
    double b2min = 1.0;
    // the three-row slots take a1 > 0 (TREC_GENERAL_FORWARD=rows5: the five-row ones)
    bool three = !(force && !strcmp(force, "rows5"));
    for (int th = theta0; th < theta0 + ntheta && rows; ++th) {
        const TiltGeneral &t = host_tg[th];
        const double aa1 = fabs(t.a[1]);
        rows = t.b[1] == 0.0 && t.b[2] <= 1.0 && same_rows(t, host_tg[theta0]) && aa1 >= 1.0 / 3.0;
        if (!rows) break;
        three = three && t.a[1] > 0.0;
        nc = std::max(nc, (int)std::ceil(2.0 / aa1 - 1e-6));
        b2min = std::min(b2min, t.b[2]);
    }
    nc = nc <= 4 ? nc : 6;
    // 16 detector rows per thread, 8 for pitches between about 20 and 29
    // degrees; SART's launches of one tilt take 4, which gives the GPU four
    // times the waves of 16 at twice the slots per row
    int rj = b2min >= 15.0 / 16.0 ? 16 : (b2min >= 7.0 / 8.0 ? 8 : 0);
    if (ntheta <= 4 && b2min >= 3.0 / 4.0) rj = 4;
    if (three) {
        // three rows per slot need (1 - b2)(RJ + 2) <= 1: 16 rows up to about
        // 19 degrees, 8 up to about 25, 4 up to about 33
        int r3 = ntheta <= 4 ? 4 : 16;
        while (r3 >= 4 && (1.0 - b2min) * (r3 + 2) > 1.0 - 1e-9) r3 /= 2;
        if (r3 >= 4) rj = r3;
        else three = false;
    }
    rows = rows && rj > 0 && g.X >= nc;
    const size_t nib = cdiv((size_t)g.X, WG);
    // row windows for a pitch with an offset (rows3b_plan)
    Rows3bPlan rp;
    const bool rowsb = !rows && ntheta > 0 && !no_rows && !(force && !strcmp(force, "rows5")) &&
                       rows3b_plan(host_tg, theta0, ntheta, g.X, g.T, g.layout, rp);
    const int rjb = rp.rj, dmaxb = rp.dmax;
    const double dlo = rp.dlo;
    const bool laneb = rp.lane;
    if (rowsb && !(force && !strcmp(force, "splat"))) {
        const size_t nblk = nib * cdiv((size_t)g.Y, rjb) * (size_t)ntheta * (size_t)nzc;
        if (nblk <= 0x7fffffffu) {
            if (rjb == 16) launch_rows3b_lane<16>(g, theta0, ntheta, ones, nzc, (unsigned)nblk, dlo, dmaxb, laneb, st);
            else if (rjb == 8) launch_rows3b_lane<8>(g, theta0, ntheta, ones, nzc, (unsigned)nblk, dlo, dmaxb, laneb, st);
            else launch_rows3b_lane<4>(g, theta0, ntheta, ones, nzc, (unsigned)nblk, dlo, dmaxb, laneb, st);
            if (nzc_used) *nzc_used = nzc;
            return hipGetLastError();
        }
    }
    // the splat: every tilt launched has its tiles
    bool splat = ntheta > 0 && !no_splat && (!rows || (force && !strcmp(force, "splat")));
    for (int th = theta0; th < theta0 + ntheta && splat; ++th) splat = host_tg[th].si > 0;
    if (splat) {
        const TiltGeneral &tl = host_tg[theta0 + ntheta - 1];
        const unsigned ntile = (unsigned)(tl.tile0 + tl.ntx - host_tg[theta0].tile0);
        const size_t nblk = (size_t)ntile * cdiv((size_t)g.Y, SP_R) * (size_t)nzc_splat;
        if (nblk <= 0x7fffffffu) {
            if (ones)
                hipLaunchKernelGGL((k_general_splat<SP_R, true>), dim3((unsigned)nblk), dim3(64), 0, st, g, theta0,
                                   ntheta, nzc_splat, ntile);
            else
                hipLaunchKernelGGL((k_general_splat<SP_R, false>), dim3((unsigned)nblk), dim3(64), 0, st, g, theta0,
                                   ntheta, nzc_splat, ntile);
            if (nzc_used) *nzc_used = nzc_splat;
            return hipGetLastError();
        }
    }
    if (rows) {
        const size_t nblk = nib * cdiv((size_t)g.Y, rj) * (size_t)ntheta * (size_t)nzc;
        if (nblk <= 0x7fffffffu) {
            if (rj == 16) launch_rows_nc<16>(g, theta0, ntheta, ones, three, nc, nzc, (unsigned)nblk, st);
            else if (rj == 8) launch_rows_nc<8>(g, theta0, ntheta, ones, three, nc, nzc, (unsigned)nblk, st);
            else launch_rows_nc<4>(g, theta0, ntheta, ones, three, nc, nzc, (unsigned)nblk, st);
            if (nzc_used) *nzc_used = nzc;
            return hipGetLastError();
        }
    }
    // one work-group per 256 pixels of a row: fewer than 2^31 for any stack
    // that fits in memory (X, Y <= 2^20 and P * Y * X floats in HBM)
    const size_t nblk = nib * (size_t)g.Y;
    if (nblk > 0x7fffffffu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_general_gather, dim3((unsigned)nblk, ntheta), dim3(WG), 0, st, g, theta0, ones ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_general_residual(const GeneralJob &g, int theta0, int ntheta, bool quot, hipStream_t st, int nzc)
{
    const size_t n = (size_t)ntheta * g.Y * g.X;
    if (quot) hipLaunchKernelGGL(k_general_residual<true>, dim3(cdiv(n, WG)), dim3(WG), 0, st, g, theta0, n, nzc);
    else hipLaunchKernelGGL(k_general_residual<false>, dim3(cdiv(n, WG)), dim3(WG), 0, st, g, theta0, n, nzc);
    return hipGetLastError();
}

hipError_t launch_general_back(const GeneralJob &g, const TiltGeneral *host_tg, int mode, int theta0, int ntheta,
                               const float *img, hipStream_t st)
{
    // byte offsets within a projection fit 32 bits; two columns and rows
    bool rows = ntheta > 0 && (size_t)g.X * g.Y * 4 <= 0xffffffffu && g.X >= 2 && g.Y >= 2;
    for (int th = theta0; th < theta0 + ntheta && rows; ++th) rows = same_rows(host_tg[th], host_tg[theta0]);
    const size_t nrows_wg = cdiv((size_t)g.Y * g.T, WG / 64);
    const size_t nblk = cdiv((size_t)g.X, GB_Q * 64) * nrows_wg;
    // staged rows: 2 when v does not depend on x, else enough for the rows
    // a wave's voxels reach (|b1| per voxel over 64 GL_Q - 1 voxels)
    int nr = 0;
    if (rows) {
        const double span = std::fabs(host_tg[theta0].b[1]) * (64.0 * GL_Q - 1.0);
        nr = host_tg[theta0].b[1] == 0.0 ? 2 : (span < 0.999 ? 3 : (span < 1.999 ? 4 : 0));
    }
    const size_t nblk_lds = cdiv((size_t)g.X, GL_Q * 64) * nrows_wg;
    // the group's four waves share their window when they take four layers
    // of one volume row whose rows differ by less than one (3 |b3| < 1)
    // (TREC_GENERAL_BACK=wave: every wave stages its own)
    const char *force = getenv("TREC_GENERAL_BACK");
    const bool sh = nr && g.layout == TREC_LAYOUT_Y && g.T % 4 == 0 && 3.0 * std::fabs(host_tg[theta0].b[3]) < 1.0 &&
                    !(force && !strcmp(force, "wave"));
    if (nr && nblk_lds <= 0x7fffffffu) {
        const dim3 grid((unsigned)nblk_lds);
        switch (nr) {
        case 2: launch_back_lds<2>(g, mode, theta0, ntheta, img, grid, sh, st); break;
        case 3: launch_back_lds<3>(g, mode, theta0, ntheta, img, grid, sh, st); break;
        default: launch_back_lds<4>(g, mode, theta0, ntheta, img, grid, sh, st); break;
        }
        return hipGetLastError();
    }
    if (rows && nblk <= 0x7fffffffu) {
        const dim3 grid((unsigned)nblk);
        if (mode == BACK_SIRT)
            hipLaunchKernelGGL(k_general_back_rows<BACK_SIRT>, grid, dim3(WG), 0, st, g, theta0, ntheta, img);
        else if (mode == BACK_LIN)
            hipLaunchKernelGGL(k_general_back_rows<BACK_LIN>, grid, dim3(WG), 0, st, g, theta0, ntheta, img);
        else if (mode == BACK_SART)
            hipLaunchKernelGGL(k_general_back_rows<BACK_SART>, grid, dim3(WG), 0, st, g, theta0, ntheta, img);
        else
            hipLaunchKernelGGL(k_general_back_rows<BACK_BPT>, grid, dim3(WG), 0, st, g, theta0, ntheta, img);
        return hipGetLastError();
    }
    const size_t n = (size_t)g.X * g.T * g.Y;
    hipLaunchKernelGGL(k_general_back, dim3(cdiv(n, WG)), dim3(WG), 0, st, g, mode, theta0, ntheta, img);
    return hipGetLastError();
}

}  // namespace trec
