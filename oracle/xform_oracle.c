/*
 * xform_oracle.c -- CPU restatement of mrcstack's .xf application.
 *
 * TEST INFRASTRUCTURE ONLY.  This file checks the product's results and is never part of it:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load the library built from it.  The HIP library in autom2_amd/ never
 * links, loads or falls back to it.
 *
 * What it restates (paths relative to the AuTom2 source tree):
 *   - the per-section loop: src/mrctools/src/mrcstack/main.cpp:40-80
 *     (M = [[a11, a12, cx + dx - a11 cx - a12 cy],
 *           [a21, a22, cy + dy - a21 cx - a22 cy]] in float, cx = W * .5,
 *      cy = H * .5; border value = the section's mean; INTER_LANCZOS4 for
 *      -m 2, INTER_CUBIC otherwise)
 *   - the mean: GetImgAvgSdvCutoffMinMax (src/mrctools/src/mrcimg/
 *     img_util.h:353-368) through cv::meanStdDev
 *   - cv::warpAffine itself lives in OpenCV (4.x; the reference's Dockerfile
 *     builds the 4.x branch, its manual names 4.8.0), which is absent here.
 *     This is a restatement of OpenCV's published algorithm
 *     (modules/imgproc/src/imgwarp.cpp: warpAffine, WarpAffineInvoker,
 *     remapBicubic, remapLanczos4, interpolateCubic, interpolateLanczos4,
 *     initInterTab2D) for single-channel float images:
 *       * the forward matrix is widened to double and inverted;
 *       * source coordinates are fixed point: 10 fractional bits per row
 *         term and per column term (each rounded to nearest even), plus a
 *         rounding bias of half a 1/32 step, reduced to 5 bits: integer
 *         part saturated to short, fraction indexes a 32 x 32 weight table;
 *       * weights are float products of two 1-D float kernels (Keys cubic
 *         with A = -0.75, or the normalised Lanczos-4 window);
 *       * fully inside the source: plain row-by-row float sums;
 *         touching the border: value + sum of (tap - value) * weight over
 *         taps inside; entirely outside: the border value.
 *     Agreement with OpenCV's own binary is therefore parity unpinned; the
 *     restatement is pinned in tests/test_xform_oracle.py by closed forms
 *     (identity, integer shifts, quarter turns, constant images, the first
 *     moment of each kernel on a ramp) and by an independent numpy
 *     evaluation of both kernels.
 *
 * Built with -ffp-contract=off: every float and double operation is rounded
 * on its own, as the reference's x86-64 build (no FMA target) evaluates it.
 */
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define XF_INTER_BITS 5
#define XF_TAB (1 << XF_INTER_BITS) /* 32 sub-pixel positions */
#define XF_AB_BITS 10
#define XF_AB_SCALE (1 << XF_AB_BITS)
#define XF_ROUND_DELTA (XF_AB_SCALE / XF_TAB / 2)

/* Keys cubic kernel at fraction x (OpenCV interpolateCubic). */
static void cubic_coeffs(float x, float *c)
{
    const float A = -0.75f;
    c[0] = ((A * (x + 1) - 5 * A) * (x + 1) + 8 * A) * (x + 1) - 4 * A;
    c[1] = ((A + 2) * x - (A + 3)) * x * x + 1;
    c[2] = ((A + 2) * (1 - x) - (A + 3)) * (1 - x) * (1 - x) + 1;
    c[3] = 1.f - c[0] - c[1] - c[2];
}

/* Normalised Lanczos-4 window at fraction x (OpenCV interpolateLanczos4):
 * sin(y_i) / y_i^2 through the angle-sum identity with one sin/cos pair,
 * the common factor dropped and the 8 taps normalised to sum 1.  A tap at
 * distance 0 gets 1e30 before normalisation, which makes it 1 and the rest
 * negligible. */
static void lanczos4_coeffs(float x, float *c)
{
    static const double s45 = 0.70710678118654752440084436210485;
    static const double cs[8][2] = {{1, 0}, {-s45, -s45}, {0, 1}, {s45, -s45},
                                    {-1, 0}, {s45, s45}, {0, -1}, {-s45, s45}};
    const double pi = 3.1415926535897932384626433832795;
    float sum = 0;
    const double y0 = -(double)(x + 3) * pi * 0.25, s0 = sin(y0), c0 = cos(y0);
    for (int i = 0; i < 8; i++) {
        const float y0f = (x + 3 - i);
        if (fabsf(y0f) >= 1e-6f) {
            const double y = -(double)y0f * pi * 0.25;
            c[i] = (float)((cs[i][0] * s0 + cs[i][1] * c0) / (y * y));
        } else {
            c[i] = 1e30f;
        }
        sum += c[i];
    }
    sum = 1.f / sum;
    for (int i = 0; i < 8; i++) c[i] *= sum;
}

/* 1-D tables at the 32 fractions i / 32: tab[i * K + k], K = 4 (cubic) or 8
 * (Lanczos-4) (initInterTab1D). */
void orc_xf_table(int lanczos, float *tab)
{
    const float scale = 1.f / XF_TAB;
    for (int i = 0; i < XF_TAB; i++) {
        if (lanczos) lanczos4_coeffs(i * scale, tab + 8 * i);
        else cubic_coeffs(i * scale, tab + 4 * i);
    }
}

/* mrcstack's forward matrix (float arithmetic, main.cpp:56-60) followed by
 * warpAffine's inversion in double.  xf = a11 a12 a21 a22 dx dy. */
void orc_xf_inverse(const float *xf, int W, int H, double *M)
{
    const float a11 = xf[0], a12 = xf[1], a21 = xf[2], a22 = xf[3], sx = xf[4], sy = xf[5];
    const float cx = W * .5, cy = H * .5;
    const float f[6] = {a11, a12, cx + sx - a11 * cx - a12 * cy, a21, a22, cy + sy - a21 * cx - a22 * cy};
    for (int i = 0; i < 6; i++) M[i] = f[i];
    double D = M[0] * M[4] - M[1] * M[3];
    D = D != 0 ? 1. / D : 0;
    const double A11 = M[4] * D, A22 = M[0] * D;
    M[0] = A11;
    M[1] *= -D;
    M[3] *= -D;
    M[4] = A22;
    const double b1 = -M[0] * M[2] - M[1] * M[5];
    const double b2 = -M[3] * M[2] - M[4] * M[5];
    M[2] = b1;
    M[5] = b2;
}

/* Mean of n floats in double: sum, then times 1/n (meanStdDev's order of
 * operations; its summation order inside OpenCV is not pinned). */
double orc_image_mean(const float *img, long n)
{
    double s = 0;
    for (long i = 0; i < n; i++) s += img[i];
    return s * (n ? 1. / n : 0.);
}

/* Round to nearest, ties to even, as x86-64 cvtsd2si does for cvRound;
 * results outside int give INT_MIN, the instruction's indefinite value. */
static int cv_round(double v)
{
    const double r = rint(v);
    if (!(r >= -2147483648.0 && r <= 2147483647.0)) return INT_MIN;
    return (int)r;
}

static int sat_short(int v) { return v < -32768 ? -32768 : (v > 32767 ? 32767 : v); }

/* Source position of output pixel (x, y): integer tap origin (before the
 * kernel's -1 / -3 offset) and the table index fy * 32 + fx. */
static void src_coord(const double *M, int x, int y, int *ix, int *iy, int *tab)
{
    const int adelta = cv_round(M[0] * x * XF_AB_SCALE);
    const int bdelta = cv_round(M[3] * x * XF_AB_SCALE);
    const int X0 = cv_round((M[1] * y + M[2]) * XF_AB_SCALE) + XF_ROUND_DELTA;
    const int Y0 = cv_round((M[4] * y + M[5]) * XF_AB_SCALE) + XF_ROUND_DELTA;
    /* two's-complement wrap, as the reference's int arithmetic does */
    const int X = (int)((unsigned)X0 + (unsigned)adelta) >> (XF_AB_BITS - XF_INTER_BITS);
    const int Y = (int)((unsigned)Y0 + (unsigned)bdelta) >> (XF_AB_BITS - XF_INTER_BITS);
    *ix = sat_short(X >> XF_INTER_BITS);
    *iy = sat_short(Y >> XF_INTER_BITS);
    *tab = (Y & (XF_TAB - 1)) * XF_TAB + (X & (XF_TAB - 1));
}

/* Exposed for tests: the fixed-point source coordinate of one pixel. */
void orc_xf_coord(const double *M, int x, int y, int *out3)
{
    src_coord(M, x, y, &out3[0], &out3[1], &out3[2]);
}

/* One section: dst = warpAffine(src, M^-1 given as M, K x K kernel, border
 * constant cval). */
void orc_warp_affine(const float *src, int W, int H, const double *M, int lanczos, float cval, float *dst)
{
    const int K = lanczos ? 8 : 4, o = lanczos ? 3 : 1;
    float t1[XF_TAB * 8];
    orc_xf_table(lanczos, t1);
    const int width1 = W - (K - 1) > 0 ? W - (K - 1) : 0;
    const int height1 = H - (K - 1) > 0 ? H - (K - 1) : 0;
    for (int y = 0; y < H; y++) {
        for (int x = 0; x < W; x++) {
            int ix, iy, t;
            src_coord(M, x, y, &ix, &iy, &t);
            const float *ty = t1 + K * (t / XF_TAB), *tx = t1 + K * (t % XF_TAB);
            float w[64];
            for (int k1 = 0; k1 < K; k1++)
                for (int k2 = 0; k2 < K; k2++) w[k1 * K + k2] = ty[k1] * tx[k2];
            const int sx = ix - o, sy = iy - o;
            float sum;
            if ((unsigned)sx < (unsigned)width1 && (unsigned)sy < (unsigned)height1) {
                const float *S = src + (long)sy * W + sx;
                if (!lanczos) {
                    sum = S[0] * w[0] + S[1] * w[1] + S[2] * w[2] + S[3] * w[3];
                    for (int r = 1; r < 4; r++) {
                        const float *R = S + (long)r * W;
                        const float *q = w + 4 * r;
                        sum += R[0] * q[0] + R[1] * q[1] + R[2] * q[2] + R[3] * q[3];
                    }
                } else {
                    sum = 0;
                    for (int r = 0; r < 8; r++) {
                        const float *R = S + (long)r * W;
                        const float *q = w + 8 * r;
                        sum += R[0] * q[0] + R[1] * q[1] + R[2] * q[2] + R[3] * q[3] + R[4] * q[4] +
                               R[5] * q[5] + R[6] * q[6] + R[7] * q[7];
                    }
                }
            } else if (sx >= W || sx + K <= 0 || sy >= H || sy + K <= 0) {
                sum = cval;
            } else {
                sum = cval;
                for (int k1 = 0; k1 < K; k1++) {
                    const int yy = sy + k1;
                    if (yy < 0 || yy >= H) continue;
                    for (int k2 = 0; k2 < K; k2++) {
                        const int xx = sx + k2;
                        if (xx < 0 || xx >= W) continue;
                        sum += (src[(long)yy * W + xx] - cval) * w[k1 * K + k2];
                    }
                }
            }
            dst[(long)y * W + x] = sum;
        }
    }
}

/* The whole stack: section i of `in` ([P][H][W]) warped by line i of the
 * .xf table (xf[6 i .. 6 i + 5]); mode 2 selects Lanczos-4, any other value
 * the cubic kernel (main.cpp:69-74).  cval_out (may be NULL) receives the
 * border value of each section. */
void orc_apply_xf(const float *in, int W, int H, int P, const float *xf, int mode, float *out, float *cval_out)
{
    const long n = (long)W * H;
    for (int i = 0; i < P; i++) {
        double M[6];
        orc_xf_inverse(xf + 6 * i, W, H, M);
        const float cval = (float)orc_image_mean(in + n * i, n);
        if (cval_out) cval_out[i] = cval;
        orc_warp_affine(in + n * i, W, H, M, mode == 2, cval, out + n * i);
    }
}
