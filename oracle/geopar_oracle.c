/*
 * geopar_oracle.c -- CPU restatement of GeoParGen, the geometry estimate
 * AuTom2 runs on a trial reconstruction.
 *
 * TEST INFRASTRUCTURE ONLY.  This file checks the product's results and is never part of it:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load the library built from it.  The HIP library in autom2_amd/ never
 * links, loads or falls back to it.
 *
 * What it restates (paths relative to the AuTom2 source tree):
 *   - main: src/GeoParGen/src/main/main.cpp:71-215 -- views of the volume,
 *     one ViewSlicer(view, 50, 50) per view, per-view angle, thickness and
 *     z, and the three averages (float std::accumulate / view count);
 *   - MrcData::getXView / getYView: mrcview.cpp:99-127;
 *   - ViewSlicer: mrcview.h:43-147 -- NORM_MINMAX to [0, 1] as CV_32F
 *     (line 112), chips of 50 columns every 50 (SubImage::init, 50-65),
 *     column sums (SubImage::project: chip * ones, 66-71), calculateNCC
 *     (116-129), getAllTAndZ (131-141);
 *   - calculateAngle (mrcview.cpp:13-45), calculateAvgThickAndZ (47-66),
 *     getThicknessAndZ (68-97).
 *   OpenCV is absent here, so the library calls are restated from OpenCV
 *   4.x's published sources, choosing the x86-64 AVX2 build's arithmetic;
 *   the exact float bits of these three steps are parity unpinned against
 *   an OpenCV binary:
 *       * cv::normalize(NORM_MINMAX, CV_32F): scale = (float)(1 / (max -
 *         min)) (0 when max - min <= DBL_EPSILON), shift = (float)0 -
 *         (float)(min * scale), then convertTo's fused x * scale + shift
 *         in float (modules/core/src/norm.cpp, convert_scale.simd.hpp);
 *       * chip * ones: GEMMSingleMul<float, double>, a double sum in column
 *         order rounded to float (modules/core/src/matmul.simd.hpp);
 *       * cv::mean: a double sum times (1. / n); the sum is taken in index
 *         order here (OpenCV's SIMD lanes may order it differently).
 *   The search and the regressions are plain C++ in the reference and are
 *   restated operation for operation, including the float running maximum
 *   of getThicknessAndZ, whose rounding decides near-ties.
 *
 * Built with -ffp-contract=off: every operation is rounded on its own, as
 * the reference's x86-64 build (no FMA target) evaluates it; the one fused
 * operation (convertTo) is written as fmaf.
 */
#include <float.h>
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define GP_STEP 50    /* ViewSlicer(view, 50, 50): chip step */
#define GP_WIDTH 50   /* chip width (SubImage offset) */
#define GP_MINT 30    /* getAllTAndZ(30, 60): minimum thickness */

/* Chips of a view `w` columns wide: i = 0, 50, ... while i < w - 50
 * (SubImage::init). */
int orc_gp_nchips(int w)
{
    return w > GP_WIDTH ? (w - GP_WIDTH - 1) / GP_STEP + 1 : 0;
}

/* Pairs (l, r) getThicknessAndZ tries on a profile of n rows. */
long long orc_gp_npairs(int n)
{
    const long long m = n - GP_MINT;
    return m > 0 ? m * (m + 1) / 2 : 0;
}

/* View v as a rows x cols float matrix (row-major): axis 0 ("x", the
 * default) gives view j in [0, nx): element (r, c) = section c, row r,
 * column j, so ny rows and nz columns; axis 1 ("y") gives section v itself,
 * ny rows and nx columns. */
void orc_gp_view(const float *vol, int nx, int ny, int nz, int axis, int v, float *out)
{
    if (axis == 0) {
        for (int r = 0; r < ny; ++r)
            for (int c = 0; c < nz; ++c) out[(size_t)r * nz + c] = vol[((size_t)c * ny + r) * nx + v];
    } else {
        memcpy(out, vol + (size_t)v * ny * nx, sizeof(float) * (size_t)ny * nx);
    }
}

/* cv::normalize(src, dst, 0, 1, NORM_MINMAX, CV_32F). */
void orc_gp_normalize(const float *src, long long count, float *dst)
{
    double smin = src[0], smax = src[0];
    for (long long i = 1; i < count; ++i) {
        if (src[i] < smin) smin = src[i];
        if (src[i] > smax) smax = src[i];
    }
    double scale = 1.0 * (smax - smin > DBL_EPSILON ? 1. / (smax - smin) : 0);
    scale = (float)scale;
    const double shift = (float)0.0 - (float)(smin * scale);
    const float a = (float)scale, b = (float)shift;
    for (long long i = 0; i < count; ++i) dst[i] = fmaf(src[i], a, b);
}

/* SubImage::init + project: prof[q][r] = sum over the chip's 50 columns of
 * row r, in double, rounded to float. */
void orc_gp_profiles(const float *norm, int rows, int cols, float *prof)
{
    const int nc = orc_gp_nchips(cols);
    for (int q = 0; q < nc; ++q)
        for (int r = 0; r < rows; ++r) {
            double s = 0.0;
            for (int u = 0; u < GP_WIDTH; ++u) s += (double)norm[(size_t)r * cols + q * GP_STEP + u] * 1.0;
            prof[(size_t)q * rows + r] = (float)(s * 1.0);
        }
}

/* cv::mean of a float column. */
static double mean_f(const float *a, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += (double)a[i];
    return s * (n ? 1. / n : 0);
}

/* calculateNCC(a, rect) with rect = ones except zeros on rows [l, r]. */
double orc_gp_ncc(const float *a, int n, int l, int r)
{
    const double amean = mean_f(a, n);
    double ones = 0.0;
    for (int i = 0; i < n; ++i) ones += (i < l || i > r) ? 1.0 : 0.0;
    const double bmean = ones * (n ? 1. / n : 0);
    double sum = 0, sum1 = 0, sum2 = 0;
    for (int i = 0; i < n; ++i) {
        const double p1 = (double)a[i] - amean;
        const double p2 = (double)((i < l || i > r) ? 1.0f : 0.0f) - bmean;
        sum += p1 * p2;
        sum1 += p1 * p1;
        sum2 += p2 * p2;
    }
    const double ans = sum / sqrt(sum1 * sum2);
    return fabs(ans);
}

/* getThicknessAndZ(profile, 30, 60, z, thickness): the literal scan with
 * its float running maximum.  Returns the float it returns (0 when nothing
 * was found, with z = thickness = -1); *lo, *hi receive l and r (-1). */
float orc_gp_search(const float *a, int n, int *z, int *thick, int *lo, int *hi)
{
    /* calculateNCC recomputes amean, sum1 and the p1 terms per pair with
     * the same operations in the same order, so they are hoisted here;
     * each pair's sum and sum2 are still formed term by term in order */
    const double amean = mean_f(a, n);
    double *p1 = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double sum1 = 0;
    for (int i = 0; i < n; ++i) {
        p1[i] = (double)a[i] - amean;
        sum1 += p1[i] * p1[i];
    }
    float maxx = 0, lcnt = 0, rcnt = 0;
    for (int l = 0; l + GP_MINT < n; l++) {
        for (int r = l + GP_MINT; r < n; r++) {
            const double bmean = (double)(n - (r - l + 1)) * (1. / n);
            const double one = 1.0 - bmean, zero = 0.0 - bmean;
            double sum = 0, sum2 = 0;
            for (int i = 0; i < n; ++i) {
                const double p2 = (i < l || i > r) ? one : zero;
                sum += p1[i] * p2;
                sum2 += p2 * p2;
            }
            const double result = fabs(sum / sqrt(sum1 * sum2));
            if (fabs(result) > fabsf(maxx)) {
                maxx = (float)result;
                lcnt = (float)l;
                rcnt = (float)r;
            }
        }
    }
    free(p1);
    if (maxx == 0.0) {
        *z = -1;
        *thick = -1;
        *lo = *hi = -1;
        return 0;
    }
    *z = (int)((lcnt + rcnt) / 2);
    *thick = (int)(rcnt - lcnt);
    *lo = (int)lcnt;
    *hi = (int)rcnt;
    return maxx;
}

/* calculateAngle and calculateAvgThickAndZ over one view's chips (z, thick,
 * ncc as getAllTAndZ stores them): out = {angle in degrees, thickness, z},
 * each rounded to float as main() stores them. */
void orc_gp_view_result(const int *z, const int *thick, const float *nc, int nchips, float *out)
{
    float totalncc = 0;
    for (int q = 0; q < nchips; ++q) totalncc += nc[q] < 0 ? 0 : nc[q];
    double wxysum = 0, wxsum = 0, wysum = 0, wx2sum = 0, wsum = 0;
    for (int q = 0; q < nchips; ++q) {
        const int pa = GP_STEP * q + GP_WIDTH / 2, pb = z[q];
        double w = (nc[q] < 0 ? 0 : nc[q]) / totalncc;
        w = w * w;
        wsum += w;
        wxsum += w * pa;
        wysum += w * pb;
        wxysum += w * pa * pb;
        wx2sum += w * pa * pa;
    }
    const double ans_a = (wsum * wxysum - wxsum * wysum) / (wsum * wx2sum - wxsum * wxsum);
    const double ang = atan(ans_a) * 180 / M_PI;
    double total = 0, wt = 0, wz = 0;
    for (int q = 0; q < nchips; ++q) {
        double w = (nc[q] < 0 ? 0 : nc[q]) / totalncc;
        w = w * w;
        wt += w * thick[q];
        wz += w * z[q];
        total += w;
    }
    out[0] = (float)ang;
    if (total == 0) {
        out[1] = 0.f;
        out[2] = 0.f;
    } else {
        out[1] = (float)(wt / total);
        out[2] = (float)(wz / total);
    }
}

/* The whole of GeoParGen on a volume of nz sections of ny x nx floats.
 * views: [nviews][3] (angle, thickness, z); chips (may be NULL):
 * [nviews][nchips][4] = (z, thickness, ncc, l) as floats -- l lets a test
 * compare the chosen pair itself; avg: the three printed averages. */
void orc_geopar(const float *vol, int nx, int ny, int nz, int axis, float *views, float *chips, float *avg)
{
    const int nviews = axis == 0 ? nx : nz, cols = axis == 0 ? nz : nx, rows = ny;
    const int nc = orc_gp_nchips(cols);
    float *v = (float *)malloc(sizeof(float) * (size_t)rows * cols);
    float *prof = (float *)malloc(sizeof(float) * (size_t)rows * (nc > 0 ? nc : 1));
    int *zz = (int *)malloc(sizeof(int) * (nc > 0 ? nc : 1));
    int *tt = (int *)malloc(sizeof(int) * (nc > 0 ? nc : 1));
    float *cc = (float *)malloc(sizeof(float) * (nc > 0 ? nc : 1));
    float sa = 0.f, st = 0.f, sz = 0.f;
    for (int k = 0; k < nviews; ++k) {
        orc_gp_view(vol, nx, ny, nz, axis, k, v);
        orc_gp_normalize(v, (long long)rows * cols, v);
        orc_gp_profiles(v, rows, cols, prof);
        for (int q = 0; q < nc; ++q) {
            int lo, hi;
            cc[q] = orc_gp_search(prof + (size_t)q * rows, rows, &zz[q], &tt[q], &lo, &hi);
            if (chips) {
                float *c = chips + ((size_t)k * nc + q) * 4;
                c[0] = (float)zz[q];
                c[1] = (float)tt[q];
                c[2] = cc[q];
                c[3] = (float)lo;
            }
        }
        orc_gp_view_result(zz, tt, cc, nc, views + 3 * (size_t)k);
        sa = sa + views[3 * (size_t)k];
        st = st + views[3 * (size_t)k + 1];
        sz = sz + views[3 * (size_t)k + 2];
    }
    avg[0] = sa / (float)nviews;
    avg[1] = st / (float)nviews;
    avg[2] = sz / (float)nviews;
    free(v);
    free(prof);
    free(zz);
    free(tt);
    free(cc);
}
