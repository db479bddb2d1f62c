// geopar.hip -- GeoParGen on the GPU: the thickness, centre and tilt of the
// specimen slab in a trial reconstruction (src/GeoParGen/src/main/main.cpp:
// 71-215; ViewSlicer, mrcview.h:43-147; mrcview.cpp:13-127).
//
// Per view of the volume (x: the ny x nz plane at column j; y: section i),
// the reference normalises the view to [0, 1] (cv::normalize NORM_MINMAX,
// CV_32F), sums chips of 50 columns every 50 into row profiles (chip * ones),
// and per profile tries every pair (l, r), r >= l + 30, for the best
// |NCC| with a rectangle of ones that is zero on rows [l, r], keeping a
// float running maximum; z = (l + r) / 2 and thickness = r - l of the
// winner.  The slope of z over the chip positions, weighted by (ncc /
// sum ncc)^2, gives the view's angle; the same weights average thickness
// and z; main averages the three over views in float.
//
// Exactness.  The reference's pair search costs O(n) per pair (n = rows),
// so O(n^3) per chip.  Here every pair gets an O(1) approximation from
// prefix sums with a bounded error; only the pairs whose approximation lies
// within that bound of the best are evaluated with the reference's own
// sequential double loop, and the reference's float-maximum scan rule is
// applied to them in scan order.  Any pair that can win under that rule
// lies in this window (its NCC rounds to the float of the maximum), and
// pairs outside the window are below every window member's rounding, so the
// window alone reproduces the full scan.  Chips whose window is larger than
// GP_WCAP (or whose profile is degenerate) take the full exact scan.
#include <hip/hip_runtime.h>
#include <float.h>
#include <math.h>

#include "trec_internal.h"

namespace trec {

// ------------------------------------------------------------ min / max

// x views: view j is column j of every row of every section.
__global__ __launch_bounds__(WG) void k_gp_minmax_x(GpLaunch a)
{
    const int j = blockIdx.x * WG + threadIdx.x;
    if (j >= a.nx) return;  // no barrier below
    const size_t rows = (size_t)a.nz * a.ny;
    const size_t per = (rows + a.splits - 1) / a.splits;
    const size_t r0 = (size_t)blockIdx.y * per, r1 = min(rows, r0 + per);
    float lo = INFINITY, hi = -INFINITY;
    int bad = 0;
#pragma unroll 8
    for (size_t r = r0; r < r1; ++r) {
        const float v = a.vol[r * a.nx + j];
        bad |= !__builtin_isfinite(v);
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
    }
    if (bad) atomicOr(&a.flags[0], 1);
    a.pmin[(size_t)blockIdx.y * a.nviews + j] = lo;
    a.pmax[(size_t)blockIdx.y * a.nviews + j] = hi;
}

// y views: view i is section i; gridDim.x work-groups share a section.
__global__ __launch_bounds__(WG) void k_gp_minmax_y(GpLaunch a)
{
    const int i = blockIdx.y;
    const size_t n = (size_t)a.ny * a.nx;
    const float *s = a.vol + (size_t)i * n;
    float lo = INFINITY, hi = -INFINITY;
    int bad = 0;
    for (size_t k = (size_t)blockIdx.x * WG + threadIdx.x; k < n; k += (size_t)gridDim.x * WG) {
        const float v = s[k];
        bad |= !__builtin_isfinite(v);
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, o, 64));
        hi = fmaxf(hi, __shfl_xor(hi, o, 64));
    }
    __shared__ float red[2][WG / 64];
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = lo;
        red[1][threadIdx.x >> 6] = hi;
    }
    if (bad) atomicOr(&a.flags[0], 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < WG / 64; ++w) {
            lo = fminf(lo, red[0][w]);
            hi = fmaxf(hi, red[1][w]);
        }
        a.pmin[(size_t)blockIdx.x * a.nviews + i] = lo;
        a.pmax[(size_t)blockIdx.x * a.nviews + i] = hi;
    }
}

// cv::normalize(view, view, 0, 1, NORM_MINMAX, CV_32F): scale =
// (float)(1 / (max - min)), or 0 when max - min <= DBL_EPSILON; shift =
// (float)0 - (float)(min * scale); convertTo then forms x * scale + shift.
__global__ __launch_bounds__(WG) void k_gp_scale(GpLaunch a)
{
#pragma clang fp contract(off)
    // 64 views per work-group (lanes), the partials split over the 4 waves,
    // so each load of a wave reads 64 neighbouring views of one partial
    __shared__ float red[2][WG / 64][64];
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int v = blockIdx.x * 64 + lane;
    float lo = INFINITY, hi = -INFINITY;
    if (v < a.nviews) {
#pragma unroll 8
        for (int s = g; s < a.splits; s += WG / 64) {
            lo = fminf(lo, a.pmin[(size_t)s * a.nviews + v]);
            hi = fmaxf(hi, a.pmax[(size_t)s * a.nviews + v]);
        }
    }
    red[0][g][lane] = lo;
    red[1][g][lane] = hi;
    __syncthreads();
    if (g != 0 || v >= a.nviews) return;
    for (int w = 1; w < WG / 64; ++w) {
        lo = fminf(lo, red[0][w][lane]);
        hi = fmaxf(hi, red[1][w][lane]);
    }
    const double smin = lo, smax = hi;
    double scale = 1.0 * (smax - smin > DBL_EPSILON ? 1. / (smax - smin) : 0);
    scale = (float)scale;
    const double shift = (float)0.0 - (float)(smin * scale);
    a.ab[v] = make_float2((float)scale, (float)shift);
}

// ------------------------------------------------------------ profiles

// x views: a 64-column x 64-row tile of chip q; columns are lanes, so each
// load reads 256 contiguous bytes of a section row.  The tile is turned in
// LDS so each profile row leaves as 256 contiguous bytes too.
__global__ __launch_bounds__(WG) void k_gp_profiles_x(GpLaunch a)
{
#pragma clang fp contract(off)
    __shared__ float tile[64][65];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int j0 = blockIdx.x * 64, r0 = blockIdx.y * 64, q = blockIdx.z;
    const int j = j0 + lane;
    const bool live = j < a.nx;
    const float2 ab = live ? a.ab[j] : make_float2(0.f, 0.f);
    const size_t sec = (size_t)a.ny * a.nx;
    for (int rr = w; rr < 64; rr += 4) {
        const int r = r0 + rr;
        float res = 0.f;
        if (live && r < a.ny) {
            const float *p = a.vol + ((size_t)q * GP_STEP * a.ny + r) * a.nx + j;
            double s = 0.0;
#pragma unroll 10
            for (int u = 0; u < GP_WIDTH; ++u) s = s + (double)__builtin_fmaf(p[u * sec], ab.x, ab.y);
            res = (float)s;
        }
        tile[lane][rr] = res;
    }
    __syncthreads();
    for (int jj = w; jj < 64; jj += 4) {
        const int jg = j0 + jj, r = r0 + lane;
        if (jg < a.nx && r < a.ny) a.prof[((size_t)jg * a.nchips + q) * a.ny + r] = tile[jj][lane];
    }
}

// y views: one work-group per (row, section); lanes take chips.
__global__ __launch_bounds__(64) void k_gp_profiles_y(GpLaunch a)
{
#pragma clang fp contract(off)
    const int r = blockIdx.x, i = blockIdx.y;
    const float2 ab = a.ab[i];
    const float *row = a.vol + ((size_t)i * a.ny + r) * a.nx;
    for (int q = threadIdx.x; q < a.nchips; q += 64) {
        double s = 0.0;
        for (int u = 0; u < GP_WIDTH; ++u) s = s + (double)__builtin_fmaf(row[q * GP_STEP + u], ab.x, ab.y);
        a.prof[((size_t)i * a.nchips + q) * a.ny + r] = (float)s;
    }
}

// ------------------------------------------------------------ search

// The reference's NCC of the profile (centred: p1 = a - mean) with the
// rectangle that is zero on [l, r], term by term in row order
// (calculateNCC, mrcview.h:116-129).
__device__ double gp_exact(const double *p1, int n, int l, int r, double sum1, double invn)
{
#pragma clang fp contract(off)
    const double bm = (double)(n - (r - l + 1)) * invn;
    const double one = 1.0 - bm, zero = 0.0 - bm;
    const double one2 = one * one, zero2 = zero * zero;
    double s = 0.0, s2 = 0.0;
    int i = 0;
    for (; i < l; ++i) {
        s = s + p1[i] * one;
        s2 = s2 + one2;
    }
    for (; i <= r; ++i) {
        s = s + p1[i] * zero;
        s2 = s2 + zero2;
    }
    for (; i < n; ++i) {
        s = s + p1[i] * one;
        s2 = s2 + one2;
    }
    return fabs(s / __builtin_sqrt(sum1 * s2));
}

// Pair t of the scan order (l ascending, then r): row l holds m - l pairs
// and starts at l m - l (l - 1) / 2.
__device__ void gp_pair(long long t, int m, int &l, int &r)
{
    const double B = 2.0 * m + 1.0;
    int L = (int)((B - __builtin_sqrt(B * B - 8.0 * (double)t)) * 0.5);
    L = max(0, min(L, m - 1));
    auto start = [m](long long x) { return x * m - x * (x - 1) / 2; };
    while (L + 1 < m && start(L + 1) <= t) ++L;
    while (L > 0 && start(L) > t) --L;
    l = L;
    r = L + GP_MINT + (int)(t - start(L));
}

// The window threshold for the largest approximate value mx, rounded down
// to a float: ((mx - 2E)(1 - 2^-20)(1 - 2 er) - 2E) with er = 2^-18 bounding
// the float tail of the approximation (conversions, the hardware reciprocal
// square root, the product).  Comparing floats against a value rounded down
// can only keep more pairs.
__device__ float gp_theta_down(float mx, double E)
{
    const double er = 0x1p-18;
    const double t = ((double)mx - 2.0 * E) * (1.0 - 0x1p-20) * (1.0 - 2.0 * er) - 2.0 * E;
    float f = (float)t;
    if ((double)f > t) f = nextafterf(f, -INFINITY);
    return f;
}

// Entries past P[n]: lanes beyond the pair triangle read them.
constexpr int GP_PAD = 64;

__device__ int wg_max_int(int v, int *red)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    int m = red[0];
    for (int w = 1; w < WG / 64; ++w) m = max(m, red[w]);
    return m;
}

__device__ float wg_max_float(float v, float *red)
{
    for (int o = 32; o > 0; o >>= 1) {
        const float u = __shfl_xor(v, o, 64);
        v = u > v ? u : v;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float m = red[0];
    for (int w = 1; w < WG / 64; ++w) m = red[w] > m ? red[w] : m;
    return m;
}

// One work-group per (view, chip).  Dynamic LDS: p1[n], P[n + 1 + GP_PAD]
// (prefix sums of p1, padded with the total), CK[n + 1] and rs[n + 1] (per count k of zero rows: (1 -
// mean of the rectangle) * sum p1 and 1 / sqrt(sum1 * sum2)), the window
// list and its exact values, and a scan buffer.  The approximation of pair
// (l, r), k = r - l + 1, is |CK[k] - (P[r + 1] - P[l])| * rs[k].
__global__ __launch_bounds__(WG) void k_gp_search(GpLaunch a)
{
#pragma clang fp contract(off)
    extern __shared__ double gsm[];
    __shared__ double sh_amean, sh_sum1, sh_ptot, sh_best;
    __shared__ int sh_cnt, ired[WG / 64];
    __shared__ float fred[WG / 64];
    __shared__ float bmax[WG / 64][GP_MAXN / 64];  // per wave and block of 64 l
    const int n = a.ny, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double *p1 = gsm, *P = gsm + n, *CK = P + (n + 1 + GP_PAD), *resw = CK + (n + 1), *scan = resw + GP_WCAP;
    float *rs = (float *)(scan + WG);
    int *list = (int *)(rs + (n + 1));
    const float *prof = a.prof + (size_t)blockIdx.x * n;
    float4 *out = a.chips + blockIdx.x;

    for (int i = tid; i < n; i += WG) p1[i] = (double)prof[i];
    __syncthreads();
    const double invn = 1. / n;
    // The sums below are sequential in row order (the reference's loops).
    // Thread 0 reads GP_CH values at a time into registers so that the LDS
    // loads of a chunk overlap each other instead of each waiting behind
    // the previous step's store.
    constexpr int GP_CH = 16;
    if (tid == 0) {
        // cv::mean: the double sum times 1 / n
        double s = 0.0;
        for (int i0 = 0; i0 < n; i0 += GP_CH) {
            double x[GP_CH];
#pragma unroll
            for (int k = 0; k < GP_CH; ++k) x[k] = i0 + k < n ? p1[i0 + k] : 0.0;
#pragma unroll
            for (int k = 0; k < GP_CH; ++k)
                if (i0 + k < n) s = s + x[k];
        }
        sh_amean = s * invn;
    }
    __syncthreads();
    const double amean = sh_amean;
    for (int i = tid; i < n; i += WG) p1[i] = p1[i] - amean;
    __syncthreads();
    if (tid == 0) {
        double s1 = 0.0, c = 0.0;
        P[0] = 0.0;
        for (int i0 = 0; i0 < n; i0 += GP_CH) {
            double x[GP_CH], pc[GP_CH];
#pragma unroll
            for (int k = 0; k < GP_CH; ++k) x[k] = i0 + k < n ? p1[i0 + k] : 0.0;
#pragma unroll
            for (int k = 0; k < GP_CH; ++k) {
                if (i0 + k < n) {
                    s1 = s1 + x[k] * x[k];
                    c = c + x[k];
                }
                pc[k] = c;
            }
#pragma unroll
            for (int k = 0; k < GP_CH; ++k)
                if (i0 + k < n) P[i0 + k + 1] = pc[k];
        }
        for (int i = n + 1; i < n + 1 + GP_PAD; ++i) P[i] = c;
        sh_sum1 = s1;
        sh_ptot = c;
        sh_cnt = 0;
    }
    __syncthreads();
    const double sum1 = sh_sum1, ptot = sh_ptot;
    const int m = n - GP_MINT;
    // no pair, or a constant profile (every NCC is 0 / 0): nothing is found
    if (m <= 0 || !(sum1 > 0.0)) {
        if (tid == 0) *out = make_float4(-1.f, -1.f, 0.f, -1.f);
        return;
    }
    const long long npairs = (long long)m * (m + 1) / 2;
    bool full = a.exhaustive || !(sum1 >= 1e-30 && sum1 <= 1e30);
    int cnt = 0;
    if (!full) {
        for (int k = tid; k <= n; k += WG) {
            const double bm = (double)(n - k) * invn, one = 1.0 - bm;
            CK[k] = one * ptot;
            const double s2 = (double)(n - k) * (one * one) + (double)k * (bm * bm);
            rs[k] = __builtin_amdgcn_rsqf((float)(sum1 * s2));
        }
        __syncthreads();
        // |approximate - exact| <= er * exact + 2 E (E: rounding of the
        // double sums, er: the float tail); the window keeps every pair whose
        // exact NCC is within 2^-20 of the maximum
        const double E = 0x1p-53 * (16.0 * n * __builtin_sqrt(2.0 * n) + 2.0 * n + 64.0);
        // Lanes take 64 neighbouring l, the waves interleave the lengths
        // k = r - l + 1, so P[l] stays in a register and the per-k terms are
        // wave-uniform.  Pass 1 finds the largest value, pass 2 collects the
        // pairs at or above the window threshold.  Lanes past the triangle
        // read the padding of P and drop out through the select.  k = n
        // (the all-zero rectangle) is 0 / 0 in the reference: skipped.
        const int lmax = n - GP_MINT - 1;  // l + 30 < n
        float mx = -1.f;
        for (int b = 0; 64 * b <= lmax; ++b) {
            const int l = 64 * b + lane;
            const double Pl = P[l];
            const int kend = min(n - 64 * b, n - 1);
            const int lim = l <= lmax ? n - l : -1;  // the lane's pairs: k <= lim
            float mb = -1.f;
#pragma unroll 4
            for (int k = GP_MINT + 1 + wave; k <= kend; k += WG / 64) {
                const double num = CK[k] - (P[l + k] - Pl);
                const float v = fabsf((float)num) * rs[k];
                mb = fmaxf(mb, k <= lim ? v : -1.f);
            }
            // the wave's largest value in this block lets pass 2 skip it
            for (int o = 32; o > 0; o >>= 1) mb = fmaxf(mb, __shfl_xor(mb, o, 64));
            if (lane == 0) bmax[wave][b] = mb;
            mx = fmaxf(mx, mb);
        }
        const float Mf = wg_max_float(mx, fred);
        if (Mf < 0.f) full = true;  // no finite NCC: let the exact scan say so
        else {
            const float thg = gp_theta_down(Mf, E);
            for (int b = 0; 64 * b <= lmax; ++b) {
                if (bmax[wave][b] < thg) continue;
                const int l = 64 * b + lane;
                const double Pl = P[l];
                const int kend = min(n - 64 * b, n - 1);
                const int lim = l <= lmax ? n - l : -1;
#pragma unroll 4
                for (int k = GP_MINT + 1 + wave; k <= kend; k += WG / 64) {
                    const double num = CK[k] - (P[l + k] - Pl);
                    const float v = fabsf((float)num) * rs[k];
                    if (k <= lim && v >= thg) {
                        const int at = atomicAdd(&sh_cnt, 1);
                        if (at < GP_WCAP) list[at] = l * n + l + k - 1;
                    }
                }
            }
            __syncthreads();
            cnt = sh_cnt;
            if (cnt > GP_WCAP) full = true;
        }
    }

    int best = -1;
    if (!full) {
        // exact values of the window, then the scan rule among them: a pair
        // is a record when its NCC exceeds the float of the largest NCC
        // before it in scan order
        double re = 0.0;
        int code = -1;
        if (tid < cnt) {
            code = list[tid];
            re = gp_exact(p1, n, code / n, code % n, sum1, invn);
            resw[tid] = re;
        }
        __syncthreads();
        int key = -1;
        if (tid < cnt) {
            double before = 0.0;
            for (int s = 0; s < cnt; ++s)
                if (list[s] < code) before = fmax(before, resw[s]);
            if (re > (double)(float)before) key = code;
        }
        best = wg_max_int(key, ired);
        if (best >= 0 && key == best) sh_best = re;
    } else {
        if (tid == 0) atomicAdd(&a.flags[1], 1);
        // the literal scan: pairs in scan order, WG at a time, with the
        // maximum of everything before carried across chunks
        double carry = 0.0;
        for (long long base = 0; base < npairs; base += WG) {
            const long long t = base + tid;
            double re = -1.0;
            int code = -1;
            if (t < npairs) {
                int l, r;
                gp_pair(t, m, l, r);
                code = l * n + r;
                re = gp_exact(p1, n, l, r, sum1, invn);
                if (!(re == re)) re = -1.0;  // NaN: never a record, never a maximum
            }
            __syncthreads();
            scan[tid] = re;
            for (int o = 1; o < WG; o <<= 1) {
                __syncthreads();
                const double x = tid >= o ? scan[tid - o] : -1.0;
                __syncthreads();
                scan[tid] = fmax(scan[tid], x);
            }
            __syncthreads();
            const double before = fmax(carry, tid > 0 ? scan[tid - 1] : -1.0);
            const int key = re > (double)(float)before ? code : -1;
            carry = fmax(carry, scan[WG - 1]);
            const int kb = wg_max_int(key, ired);
            if (kb >= 0) {
                best = kb;
                if (key == kb) sh_best = re;
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        if (best < 0) {
            *out = make_float4(-1.f, -1.f, 0.f, -1.f);
        } else {
            const float lcnt = (float)(best / n), rcnt = (float)(best % n);
            const float z = (float)(int)((lcnt + rcnt) / 2), th = (float)(int)(rcnt - lcnt);
            *out = make_float4(z, th, (float)sh_best, lcnt);
        }
    }
}

// ------------------------------------------------------------ per view

// calculateAngle and calculateAvgThickAndZ (mrcview.cpp:13-66) over the
// view's chips; chip q sits at x = 50 q + 25.
__global__ __launch_bounds__(WG) void k_gp_views(GpLaunch a)
{
#pragma clang fp contract(off)
    const int v = blockIdx.x * WG + threadIdx.x;
    if (v >= a.nviews) return;
    const float4 *c = a.chips + (size_t)v * a.nchips;
    float total = 0.f;
    for (int q = 0; q < a.nchips; ++q) total = total + (c[q].z < 0.f ? 0.f : c[q].z);
    double wxy = 0, wx = 0, wy = 0, wx2 = 0, ws = 0, wt = 0, wz = 0;
    for (int q = 0; q < a.nchips; ++q) {
        const int pa = GP_STEP * q + GP_WIDTH / 2, pb = (int)c[q].x, th = (int)c[q].y;
        double w = (c[q].z < 0.f ? 0.f : c[q].z) / total;
        w = w * w;
        ws = ws + w;
        wx = wx + w * pa;
        wy = wy + w * pb;
        wxy = wxy + w * pa * pb;
        wx2 = wx2 + w * pa * pa;
        wt = wt + w * th;
        wz = wz + w * pb;
    }
    const double ans = (ws * wxy - wx * wy) / (ws * wx2 - wx * wx);
    const double ang = atan(ans) * 180 / M_PI;
    float *o = a.views + 3 * (size_t)v;
    o[0] = (float)ang;
    if (ws == 0) {
        o[1] = 0.f;
        o[2] = 0.f;
    } else {
        o[1] = (float)(wt / ws);
        o[2] = (float)(wz / ws);
    }
}

// main.cpp:206-208: float sums over views in order, divided by the count.
// The work-group stages 1024 views at a time in LDS; thread 0 sums them in
// order, reading a chunk into registers at a time.
__global__ __launch_bounds__(WG) void k_gp_average(GpLaunch a)
{
#pragma clang fp contract(off)
    __shared__ float buf[3 * 1024];
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    for (int base = 0; base < a.nviews; base += 1024) {
        const int cnt = min(1024, a.nviews - base);
        __syncthreads();
        for (int k = threadIdx.x; k < 3 * cnt; k += WG) buf[k] = a.views[3 * (size_t)base + k];
        __syncthreads();
        if (threadIdx.x == 0) {
            int i = 0;
            for (; i + 8 <= cnt; i += 8) {
                float x[24];
#pragma unroll
                for (int k = 0; k < 24; ++k) x[k] = buf[3 * i + k];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    s0 = s0 + x[3 * k];
                    s1 = s1 + x[3 * k + 1];
                    s2 = s2 + x[3 * k + 2];
                }
            }
            for (; i < cnt; ++i) {
                s0 = s0 + buf[3 * i];
                s1 = s1 + buf[3 * i + 1];
                s2 = s2 + buf[3 * i + 2];
            }
        }
    }
    if (threadIdx.x == 0) {
        const float nv = (float)a.nviews;
        a.avg[0] = s0 / nv;
        a.avg[1] = s1 / nv;
        a.avg[2] = s2 / nv;
    }
}

// ------------------------------------------------------------ launchers

hipError_t launch_gp_scale(const GpLaunch &a, hipStream_t st)
{
    if (a.axis == 0)
        hipLaunchKernelGGL(k_gp_minmax_x, dim3((a.nx + WG - 1) / WG, a.splits), dim3(WG), 0, st, a);
    else
        hipLaunchKernelGGL(k_gp_minmax_y, dim3(a.splits, a.nz), dim3(WG), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gp_scale, dim3((a.nviews + 63) / 64), dim3(WG), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_gp_profiles(const GpLaunch &a, hipStream_t st)
{
    if (a.nchips == 0) return hipSuccess;
    if (a.axis == 0)
        hipLaunchKernelGGL(k_gp_profiles_x, dim3((a.nx + 63) / 64, (a.ny + 63) / 64, a.nchips), dim3(WG), 0, st, a);
    else
        hipLaunchKernelGGL(k_gp_profiles_y, dim3(a.ny, a.nz), dim3(64), 0, st, a);
    return hipGetLastError();
}

size_t gp_search_lds(int n)
{
    return sizeof(double) * ((size_t)n + 2 * ((size_t)n + 1) + GP_PAD + GP_WCAP + WG) +
           sizeof(float) * ((size_t)n + 1) + sizeof(int) * GP_WCAP;
}

hipError_t launch_gp_search(const GpLaunch &a, hipStream_t st)
{
    if (a.nchips == 0) return hipSuccess;
    const size_t lds = gp_search_lds(a.ny);
    if (lds > 64 * 1024) {
        static hipError_t attr = hipFuncSetAttribute((const void *)k_gp_search,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (attr != hipSuccess) return attr;
    }
    hipLaunchKernelGGL(k_gp_search, dim3((unsigned)((size_t)a.nviews * a.nchips)), dim3(WG), lds, st, a);
    return hipGetLastError();
}

hipError_t launch_gp_views(const GpLaunch &a, hipStream_t st)
{
    hipLaunchKernelGGL(k_gp_views, dim3((a.nviews + WG - 1) / WG), dim3(WG), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gp_average, dim3(1), dim3(WG), 0, st, a);
    return hipGetLastError();
}

}  // namespace trec
