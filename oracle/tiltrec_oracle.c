/*
 * tiltrec_oracle.c -- CPU restatement of TiltRec's reconstruction arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  This file checks the product's results and is never part of it:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load the library built from it.  The HIP library in autom2_amd/ never
 * links, loads or falls back to it.
 *
 * What it restates (paths relative to the AuTom2 source tree):
 *   - coefficients:  src/TiltRec/src/main/main-mpi.cpp:85-183 (double storage,
 *                    float trig) and main-cuda.cpp:111-168 (float storage)
 *   - ValCoef:       main-mpi.cpp:204-226
 *   - Reproject:     main-mpi.cpp:286-344 (the y-major variant SIRT/SART use)
 *   - BilinearValue: main-mpi.cpp:378-410
 *   - BackProject:   main-mpi.cpp:411-450 (BPT, and WBP after filtering)
 *   - SART:          main-mpi.cpp:492-524, 578-612
 *   - SIRT:          main-mpi.cpp:526-576, 614-664
 *   - R-weighting:   src/TiltRec/src/filter/filter_prj.cpp:636-736 (mode 2 of
 *                    ApplyFilterInplace, :738-802), evaluated in double with an
 *                    explicit DFT instead of FFTW/cuFFT.
 *
 * Two arithmetic forms are provided for every method:
 *   *_f32 -- the reference's own types: double coordinates (ValCoef), float
 *            weights and float accumulators, in the reference's loop order.
 *            Built with -ffp-contract=off so no FMA changes the rounding.
 *   *_f64 -- the same loops with every weight and accumulator in double: the
 *            FP64 oracle that states the error of a float result.
 *
 * Numeric rules.  Every update and residual test takes a threshold that the
 * weight sum must exceed, compared in double as the reference compares a
 * float with a double literal:
 *   TiltRec-mpi:  0 everywhere (its `if (c)` / `if (rw[n])` tests; the sums
 *                 are never negative), main-mpi.cpp:442,571,600,645;
 *   TiltRec-cuda: BPT/WBP/FBP c > 1e-5 (basic.cu:409), SIRT wtvol > 10e-6
 *                 (basic.cu:184) with an exact residual test (:152), SART
 *                 c > 10e-7 (:434) and reproj_wt > 1e-7 (:170).
 * TiltRec-cuda also puts the y origin at Y/2 (main-cuda.cpp:259-261), which
 * callers pass in org[1].
 *
 * Data layouts follow the TiltRec-mpi y path after RotateX
 * (mrcstack.cpp:566-579): projections are [P][rows][X] and the volume is
 * [rows][T][X] (one XZ section per row, the order WriteBlock('z') writes).
 *
 * Parity of each function is pinned in tests/test_oracle.py against closed
 * forms, a brute-force FP64 system matrix, adjointness and numpy's FFT.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

#define ORC_PI_180 0.01745329252f /* PI_180, main-mpi.cpp:9 */

/* ------------------------------------------------------------------ */
/* coefficients                                                         */
/* ------------------------------------------------------------------ */

/* main-mpi.cpp:85-106 then :110-183, linear terms only (the quadratic
 * coefficients a4..a9/b4..b9 start at zero and ValCoef never reads them).
 * coef[8*i + 0..3] = a0..a3, coef[8*i + 4..7] = b0..b3.
 * cos()/sin() of a float argument resolve to the float overloads in the
 * reference (mrcstack.h includes <math.h>, which brings std::cos(float) into
 * scope), so the tilt trig is cosf/sinf widened to double. */
void orc_coeffs_mpi(const float *angles, int P, float offset, float pitch,
                    float zshift, double *coef)
{
    double alpha = -(pitch * ORC_PI_180);
    double beta = offset * ORC_PI_180;
    double t = -zshift;
    double ca = cos(alpha), sa = sin(alpha), cb = cos(beta), sb = sin(beta);
    for (int i = 0; i < P; i++) {
        float r = angles[i] * ORC_PI_180;
        double a[4] = {0.0, (double)cosf(r), 0.0, (double)(-sinf(r))};
        double b[4] = {0.0, 0.0, 1.0, 0.0};
        double na[4], nb[4];
        na[0] = a[0];
        na[1] = (a[2] * sa * sb + a[3] * ca * sb + a[1] * cb);
        na[2] = (a[2] * ca - a[3] * sa);
        na[3] = (a[3] * ca * cb + a[2] * cb * sa - a[1] * sb);
        nb[0] = b[0];
        nb[1] = (b[2] * sa * sb + b[3] * ca * sb + b[1] * cb);
        nb[2] = (b[2] * ca - b[3] * sa);
        nb[3] = (b[3] * ca * cb + b[2] * cb * sa - b[1] * sb);
        /* z shift, main-mpi.cpp:167-182 (a9 = b9 = 0, a5 = a6 = 0) */
        double *c = coef + 8 * (size_t)i;
        c[0] = na[0] + na[3] * t;
        c[1] = na[1];
        c[2] = na[2];
        c[3] = na[3];
        c[4] = nb[0] + nb[3] * t;
        c[5] = nb[1];
        c[6] = nb[2];
        c[7] = nb[3];
    }
}

/* main-cuda.cpp:111-168: the same map held in float (SimCoeff). */
void orc_coeffs_cuda(const float *angles, int P, float offset, float pitch,
                     float zshift, float *coef)
{
    float alpha = -(pitch * ORC_PI_180), beta = offset * ORC_PI_180, t = -zshift;
    float ca = cosf(alpha), sa = sinf(alpha), cb = cosf(beta), sb = sinf(beta);
    for (int i = 0; i < P; i++) {
        float r = angles[i] * ORC_PI_180;
        float a[4] = {0.f, cosf(r), 0.f, -sinf(r)};
        float b[4] = {0.f, 0.f, 1.f, 0.f};
        float *c = coef + 8 * (size_t)i;
        float a3 = (a[3] * ca * cb + a[2] * cb * sa - a[1] * sb);
        float b3 = (b[3] * ca * cb + b[2] * cb * sa - b[1] * sb);
        c[0] = a[0] + a3 * t;
        c[1] = (a[2] * sa * sb + a[3] * ca * sb + a[1] * cb);
        c[2] = (a[2] * ca - a[3] * sa);
        c[3] = a3;
        c[4] = b[0] + b3 * t;
        c[5] = (b[2] * sa * sb + b[3] * ca * sb + b[1] * cb);
        c[6] = (b[2] * ca - b[3] * sa);
        c[7] = b3;
    }
}

/* ------------------------------------------------------------------ */
/* ValCoef / taps                                                        */
/* ------------------------------------------------------------------ */

typedef struct {
    int xm, ym;
    float xd, yd;   /* Weight::x_min_del / y_min_del are float (mrcstack.h:291-298) */
    double xdd, ydd; /* the same fractions before the float rounding (FP64 oracle) */
} orc_w;

/* main-mpi.cpp:204-226.  coord - origin is int - float, evaluated in float. */
static inline void valcoef(const float org[3], int cx, int cy, int cz,
                           const double *c, orc_w *w)
{
    double X = (float)((float)cx - org[0]);
    double Y = (float)((float)cy - org[1]);
    double Z = (float)((float)cz - org[2]);
    double n0 = c[0] + c[1] * X + c[2] * Y + c[3] * Z;
    double n1 = c[4] + c[5] * X + c[6] * Y + c[7] * Z;
    double x = n0 + (double)org[0];
    double y = n1 + (double)org[1];
    w->xm = (int)floor(x);
    w->ym = (int)floor(y);
    w->xdd = x - (double)w->xm;
    w->ydd = y - (double)w->ym;
    w->xd = (float)w->xdd;
    w->yd = (float)w->ydd;
}

/* ------------------------------------------------------------------ */
/* FP32 form                                                             */
/* ------------------------------------------------------------------ */

/* Reproject, main-mpi.cpp:286-344.  vol [len][T][W]; rv/rw [len][W], added to. */
static void reproject_f32(const float org[3], const float *vol, const double *c,
                          float *rv, float *rw, int T, int len, int W)
{
    for (int y = 0; y < len; y++) {
        const float *vy = vol + (size_t)y * W * T;
        for (int z = 0; z < T; z++) {
            const float *vz = vy + (size_t)z * W;
            for (int x = 0; x < W; x++) {
                float v = vz[x];
                orc_w w;
                valcoef(org, x, y, z, c, &w);
                int n;
                if (w.xm >= 0 && w.xm < W && w.ym >= 0 && w.ym < len) {
                    n = w.xm + w.ym * W;
                    rv[n] += (1 - w.xd) * (1 - w.yd) * v;
                    rw[n] += (1 - w.xd) * (1 - w.yd);
                }
                if ((w.xm + 1) >= 0 && (w.xm + 1) < W && w.ym >= 0 && w.ym < len) {
                    n = w.xm + 1 + w.ym * W;
                    rv[n] += w.xd * (1 - w.yd) * v;
                    rw[n] += w.xd * (1 - w.yd);
                }
                if (w.xm >= 0 && w.xm < W && (w.ym + 1) >= 0 && (w.ym + 1) < len) {
                    n = w.xm + (w.ym + 1) * W;
                    rv[n] += (1 - w.xd) * w.yd * v;
                    rw[n] += (1 - w.xd) * w.yd;
                }
                if ((w.xm + 1) >= 0 && (w.xm + 1) < W && (w.ym + 1) >= 0 && (w.ym + 1) < len) {
                    n = (w.xm + 1) + (w.ym + 1) * W;
                    rv[n] += w.xd * w.yd * v;
                    rw[n] += w.xd * w.yd;
                }
            }
        }
    }
}

/* BilinearValue, main-mpi.cpp:378-410 */
static inline void bilinear_f32(int W, int len, const float *img, const orc_w *w,
                                float *val, float *vwt)
{
    int n;
    if (w->xm >= 0 && w->xm < W && w->ym >= 0 && w->ym < len) {
        n = w->xm + w->ym * W;
        *val += (1 - w->xd) * (1 - w->yd) * img[n];
        *vwt += (1 - w->xd) * (1 - w->yd);
    }
    if ((w->xm + 1) >= 0 && (w->xm + 1) < W && w->ym >= 0 && w->ym < len) {
        n = w->xm + 1 + w->ym * W;
        *val += w->xd * (1 - w->yd) * img[n];
        *vwt += w->xd * (1 - w->yd);
    }
    if (w->xm >= 0 && w->xm < W && (w->ym + 1) >= 0 && (w->ym + 1) < len) {
        n = w->xm + (w->ym + 1) * W;
        *val += (1 - w->xd) * w->yd * img[n];
        *vwt += (1 - w->xd) * w->yd;
    }
    if ((w->xm + 1) >= 0 && (w->xm + 1) < W && (w->ym + 1) >= 0 && (w->ym + 1) < len) {
        n = w->xm + 1 + (w->ym + 1) * W;
        *val += w->xd * w->yd * img[n];
        *vwt += w->xd * w->yd;
    }
}

/* One forward projection of one tilt: rv/rw [len][W] are overwritten. */
void orc_reproject_f32(const double *coef, int W, int len, int T,
                       const float org[3], const float *vol, float *rv, float *rw)
{
    memset(rv, 0, sizeof(float) * (size_t)W * len);
    memset(rw, 0, sizeof(float) * (size_t)W * len);
    reproject_f32(org, vol, coef, rv, rw, T, len, W);
}

/* BackProject, main-mpi.cpp:411-450 (BPT).  proj [P][len][W], vol [len][T][W]
 * is added to (the reference starts from zero or from --initial). */
void orc_bpt_f32(const double *coef, int P, int W, int len, int T,
                 const float org[3], const float *proj, float *vol, double thr_upd)
{
    for (int idx = 0; idx < P; idx++) {
        const float *one = proj + (size_t)W * len * idx;
        for (int y = 0; y < len; y++) {
            float *vy = vol + (size_t)y * W * T;
            for (int z = 0; z < T; z++) {
                float *vz = vy + (size_t)z * W;
                for (int x = 0; x < W; x++) {
                    orc_w w;
                    float s = 0, c = 0;
                    valcoef(org, x, y, z, coef + 8 * (size_t)idx, &w);
                    bilinear_f32(W, len, one, &w, &s, &c);
                    if ((double)c > thr_upd) vz[x] += (float)(s / c);
                }
            }
        }
    }
}

/* SIRT, main-mpi.cpp:614-664 with UpdateWeightsByProjDiff (:526-562). */
void orc_sirt_f32(const double *coef, int P, int W, int len, int T,
                  const float org[3], const float *proj, int iters, float gamma,
                  float *vol, double thr_upd, double thr_res)
{
    size_t px = (size_t)len * W, size = (size_t)len * W * T;
    float *rv = (float *)malloc(sizeof(float) * px);
    float *rw = (float *)malloc(sizeof(float) * px);
    float *valvol = (float *)malloc(sizeof(float) * size);
    float *wtvol = (float *)malloc(sizeof(float) * size);
    for (int it = 0; it < iters; it++) {
        memset(valvol, 0, sizeof(float) * size);
        memset(wtvol, 0, sizeof(float) * size);
        for (int idx = 0; idx < P; idx++) {
            const double *c = coef + 8 * (size_t)idx;
            memset(rv, 0, sizeof(float) * px);
            memset(rw, 0, sizeof(float) * px);
            reproject_f32(org, vol, c, rv, rw, T, len, W);
            const float *pd = proj + (size_t)W * len * idx;
            for (size_t n = 0; n < px; n++) {
                /* diff = 0 unless the ray weight passes the test (basic.cu:142-175) */
                const float diff = (double)rw[n] > thr_res ? rv[n] / rw[n] : 0.f;
                rv[n] = pd[n] - diff;
            }
            for (int y = 0; y < len; y++) {
                float *vy = valvol + (size_t)y * W * T;
                float *wy = wtvol + (size_t)y * W * T;
                for (int z = 0; z < T; z++) {
                    for (int x = 0; x < W; x++) {
                        orc_w w;
                        float s = 0, cc = 0;
                        valcoef(org, x, y, z, c, &w);
                        bilinear_f32(W, len, rv, &w, &s, &cc);
                        vy[(size_t)z * W + x] += s;
                        wy[(size_t)z * W + x] += cc;
                    }
                }
            }
        }
        for (size_t i = size; i--;) {
            if ((double)wtvol[i] > thr_upd) vol[i] += valvol[i] / wtvol[i] * gamma;
        }
    }
    free(rv);
    free(rw);
    free(valvol);
    free(wtvol);
}

/* SART, main-mpi.cpp:578-612 with UpdateVolumeByProjDiff (:492-524). */
void orc_sart_f32(const double *coef, int P, int W, int len, int T,
                  const float org[3], const float *proj, int iters, float gamma,
                  float *vol, double thr_upd, double thr_res)
{
    size_t px = (size_t)len * W;
    float *rv = (float *)malloc(sizeof(float) * px);
    float *rw = (float *)malloc(sizeof(float) * px);
    for (int it = 0; it < iters; it++) {
        for (int idx = 0; idx < P; idx++) {
            const double *c = coef + 8 * (size_t)idx;
            memset(rv, 0, sizeof(float) * px);
            memset(rw, 0, sizeof(float) * px);
            reproject_f32(org, vol, c, rv, rw, T, len, W);
            const float *pd = proj + (size_t)W * len * idx;
            for (size_t n = 0; n < px; n++) {
                /* diff = 0 unless the ray weight passes the test (basic.cu:142-175) */
                const float diff = (double)rw[n] > thr_res ? rv[n] / rw[n] : 0.f;
                rv[n] = pd[n] - diff;
            }
            for (int y = 0; y < len; y++) {
                float *vy = vol + (size_t)y * W * T;
                for (int z = 0; z < T; z++) {
                    for (int x = 0; x < W; x++) {
                        orc_w w;
                        float s = 0, cc = 0;
                        valcoef(org, x, y, z, c, &w);
                        bilinear_f32(W, len, rv, &w, &s, &cc);
                        if ((double)cc > thr_upd) vy[(size_t)z * W + x] += (float)(s / cc) * gamma;
                    }
                }
            }
        }
    }
    free(rv);
    free(rw);
}

/* ------------------------------------------------------------------ */
/* FP64 form: identical loops, double weights and accumulators          */
/* ------------------------------------------------------------------ */

static void reproject_f64(const float org[3], const double *vol, const double *c,
                          double *rv, double *rw, int T, int len, int W)
{
    for (int y = 0; y < len; y++) {
        const double *vy = vol + (size_t)y * W * T;
        for (int z = 0; z < T; z++) {
            const double *vz = vy + (size_t)z * W;
            for (int x = 0; x < W; x++) {
                double v = vz[x];
                orc_w w;
                valcoef(org, x, y, z, c, &w);
                double xd = w.xdd, yd = w.ydd;
                int n;
                if (w.xm >= 0 && w.xm < W && w.ym >= 0 && w.ym < len) {
                    n = w.xm + w.ym * W;
                    rv[n] += (1 - xd) * (1 - yd) * v;
                    rw[n] += (1 - xd) * (1 - yd);
                }
                if ((w.xm + 1) >= 0 && (w.xm + 1) < W && w.ym >= 0 && w.ym < len) {
                    n = w.xm + 1 + w.ym * W;
                    rv[n] += xd * (1 - yd) * v;
                    rw[n] += xd * (1 - yd);
                }
                if (w.xm >= 0 && w.xm < W && (w.ym + 1) >= 0 && (w.ym + 1) < len) {
                    n = w.xm + (w.ym + 1) * W;
                    rv[n] += (1 - xd) * yd * v;
                    rw[n] += (1 - xd) * yd;
                }
                if ((w.xm + 1) >= 0 && (w.xm + 1) < W && (w.ym + 1) >= 0 && (w.ym + 1) < len) {
                    n = (w.xm + 1) + (w.ym + 1) * W;
                    rv[n] += xd * yd * v;
                    rw[n] += xd * yd;
                }
            }
        }
    }
}

static inline void bilinear_f64(int W, int len, const double *img, const orc_w *w,
                                double *val, double *vwt)
{
    double xd = w->xdd, yd = w->ydd;
    int n;
    if (w->xm >= 0 && w->xm < W && w->ym >= 0 && w->ym < len) {
        n = w->xm + w->ym * W;
        *val += (1 - xd) * (1 - yd) * img[n];
        *vwt += (1 - xd) * (1 - yd);
    }
    if ((w->xm + 1) >= 0 && (w->xm + 1) < W && w->ym >= 0 && w->ym < len) {
        n = w->xm + 1 + w->ym * W;
        *val += xd * (1 - yd) * img[n];
        *vwt += xd * (1 - yd);
    }
    if (w->xm >= 0 && w->xm < W && (w->ym + 1) >= 0 && (w->ym + 1) < len) {
        n = w->xm + (w->ym + 1) * W;
        *val += (1 - xd) * yd * img[n];
        *vwt += (1 - xd) * yd;
    }
    if ((w->xm + 1) >= 0 && (w->xm + 1) < W && (w->ym + 1) >= 0 && (w->ym + 1) < len) {
        n = w->xm + 1 + (w->ym + 1) * W;
        *val += xd * yd * img[n];
        *vwt += xd * yd;
    }
}

void orc_reproject_f64(const double *coef, int W, int len, int T,
                       const float org[3], const double *vol, double *rv, double *rw)
{
    memset(rv, 0, sizeof(double) * (size_t)W * len);
    memset(rw, 0, sizeof(double) * (size_t)W * len);
    reproject_f64(org, vol, coef, rv, rw, T, len, W);
}

/* The per-voxel sums of one tilt's back-projection (BilinearValue over a
 * whole slab): s[len][T][W] and c[len][T][W] are overwritten. */
void orc_backproject_sums_f64(const double *coef, int W, int len, int T,
                              const float org[3], const double *img,
                              double *s, double *c)
{
    for (int y = 0; y < len; y++)
        for (int z = 0; z < T; z++)
            for (int x = 0; x < W; x++) {
                orc_w w;
                double sv = 0, cv = 0;
                valcoef(org, x, y, z, coef, &w);
                bilinear_f64(W, len, img, &w, &sv, &cv);
                size_t i = ((size_t)y * T + z) * W + x;
                s[i] = sv;
                c[i] = cv;
            }
}

void orc_bpt_f64(const double *coef, int P, int W, int len, int T,
                 const float org[3], const double *proj, double *vol, double thr_upd)
{
    for (int idx = 0; idx < P; idx++) {
        const double *one = proj + (size_t)W * len * idx;
        for (int y = 0; y < len; y++)
            for (int z = 0; z < T; z++)
                for (int x = 0; x < W; x++) {
                    orc_w w;
                    double s = 0, c = 0;
                    valcoef(org, x, y, z, coef + 8 * (size_t)idx, &w);
                    bilinear_f64(W, len, one, &w, &s, &c);
                    if (c > thr_upd) vol[((size_t)y * T + z) * W + x] += s / c;
                }
    }
}

void orc_sirt_f64(const double *coef, int P, int W, int len, int T,
                  const float org[3], const double *proj, int iters, double gamma,
                  double *vol, double thr_upd, double thr_res)
{
    size_t px = (size_t)len * W, size = (size_t)len * W * T;
    double *rv = (double *)malloc(sizeof(double) * px);
    double *rw = (double *)malloc(sizeof(double) * px);
    double *valvol = (double *)malloc(sizeof(double) * size);
    double *wtvol = (double *)malloc(sizeof(double) * size);
    for (int it = 0; it < iters; it++) {
        memset(valvol, 0, sizeof(double) * size);
        memset(wtvol, 0, sizeof(double) * size);
        for (int idx = 0; idx < P; idx++) {
            const double *c = coef + 8 * (size_t)idx;
            memset(rv, 0, sizeof(double) * px);
            memset(rw, 0, sizeof(double) * px);
            reproject_f64(org, vol, c, rv, rw, T, len, W);
            const double *pd = proj + (size_t)W * len * idx;
            for (size_t n = 0; n < px; n++) {
                const double diff = rw[n] > thr_res ? rv[n] / rw[n] : 0.0;
                rv[n] = pd[n] - diff;
            }
            for (int y = 0; y < len; y++)
                for (int z = 0; z < T; z++)
                    for (int x = 0; x < W; x++) {
                        orc_w w;
                        double s = 0, cc = 0;
                        valcoef(org, x, y, z, c, &w);
                        bilinear_f64(W, len, rv, &w, &s, &cc);
                        size_t i = ((size_t)y * T + z) * W + x;
                        valvol[i] += s;
                        wtvol[i] += cc;
                    }
        }
        for (size_t i = size; i--;) {
            if (wtvol[i] > thr_upd) vol[i] += valvol[i] / wtvol[i] * gamma;
        }
    }
    free(rv);
    free(rw);
    free(valvol);
    free(wtvol);
}

void orc_sart_f64(const double *coef, int P, int W, int len, int T,
                  const float org[3], const double *proj, int iters, double gamma,
                  double *vol, double thr_upd, double thr_res)
{
    size_t px = (size_t)len * W;
    double *rv = (double *)malloc(sizeof(double) * px);
    double *rw = (double *)malloc(sizeof(double) * px);
    for (int it = 0; it < iters; it++) {
        for (int idx = 0; idx < P; idx++) {
            const double *c = coef + 8 * (size_t)idx;
            memset(rv, 0, sizeof(double) * px);
            memset(rw, 0, sizeof(double) * px);
            reproject_f64(org, vol, c, rv, rw, T, len, W);
            const double *pd = proj + (size_t)W * len * idx;
            for (size_t n = 0; n < px; n++) {
                const double diff = rw[n] > thr_res ? rv[n] / rw[n] : 0.0;
                rv[n] = pd[n] - diff;
            }
            for (int y = 0; y < len; y++)
                for (int z = 0; z < T; z++)
                    for (int x = 0; x < W; x++) {
                        orc_w w;
                        double s = 0, cc = 0;
                        valcoef(org, x, y, z, c, &w);
                        bilinear_f64(W, len, rv, &w, &s, &cc);
                        if (cc > thr_upd) vol[((size_t)y * T + z) * W + x] += s / cc * gamma;
                    }
        }
    }
    free(rv);
    free(rw);
}

/* ------------------------------------------------------------------ */
/* R-weighting filter (ApplyFilterInplace mode 2), FP64                 */
/* ------------------------------------------------------------------ */

/* filter_prj.cpp:652-736 for `rows` rows of length W, in place, in double:
 * extend each row to N = W + W/10 with the linear ramp of buf2fft_padding_1D,
 * forward real DFT, weight (0.2 at DC, k up to K = floor(0.05 N), a Gaussian
 * fall-off of width 0.5 N beyond), unnormalised inverse real DFT, keep the
 * first W samples and divide by W.  The DFT is evaluated directly (O(N^2)),
 * which is exact up to double rounding; numpy's FFT pins it in the tests. */
void orc_rweight_f64(double *data, int W, int rows)
{
    int N = W + W / 10;
    int nxp = N - W + 1;
    int K = (int)floorf((float)N * 0.05f); /* int(floor(Nx_final * radial)), radial = 0.05f */
    double sig = (double)((float)N * 0.5f); /* sigma_Nx = float(Nx_final) * 0.5f */
    int nb = N / 2 + 1; /* bins the r2c transform produces */
    double *e = (double *)malloc(sizeof(double) * N);
    double *re = (double *)malloc(sizeof(double) * nb);
    double *im = (double *)malloc(sizeof(double) * nb);
    double *ct = (double *)malloc(sizeof(double) * N);
    double *st = (double *)malloc(sizeof(double) * N);
    for (int m = 0; m < N; m++) {
        ct[m] = cos(2.0 * M_PI * m / N);
        st[m] = sin(2.0 * M_PI * m / N);
    }
    for (int r = 0; r < rows; r++) {
        double *row = data + (size_t)r * W;
        for (int j = 0; j < W; j++) e[j] = row[j];
        for (int j = W; j < N; j++)
            e[j] = ((double)(j - W + 1) / nxp) * row[0] + ((double)(N - j) / nxp) * row[W - 1];
        for (int k = 0; k < nb; k++) {
            double sr = 0, si = 0;
            for (int j = 0; j < N; j++) {
                int m = (int)(((long long)k * j) % N);
                sr += e[j] * ct[m];
                si -= e[j] * st[m];
            }
            double wgt;
            if (k == 0)
                wgt = 0.2;
            else if (k <= K)
                wgt = k;
            else
                wgt = (double)K * exp(-((double)(k - K) * (k - K)) / (sig * sig));
            re[k] = sr * wgt;
            im[k] = si * wgt;
        }
        for (int j = 0; j < W; j++) {
            /* unnormalised c2r: Hermitian extension of the half spectrum;
             * the imaginary parts of DC and (even N) Nyquist are dropped */
            double acc = re[0];
            for (int k = 1; k < nb; k++) {
                int m = (int)(((long long)k * j) % N);
                double term = re[k] * ct[m] - im[k] * st[m];
                if ((N % 2 == 0) && k == N / 2)
                    acc += re[k] * ct[m];
                else
                    acc += 2.0 * term;
            }
            row[j] = acc / W;
        }
    }
    free(e);
    free(re);
    free(im);
    free(ct);
    free(st);
}

/* ------------------------------------------------------------------ FBP
 * ApplyFilterInplace modes 0 (Shepp-Logan) and 1 (Ram-Lak), the filter of
 * TiltRec's FBP (filter_prj.cpp:137-178 taps, 266-446 the filtering, 738-771
 * the dispatch; main-mpi.cpp:835-840).  Rows are filtered independently (the
 * reference transforms with CV_DXT_ROWS, so its y padding never mixes
 * rows).  Per row of W pixels:
 *   pad  = row with its first and last pixel repeated (W + 2 cells, the
 *          symmetrised 1-pixel border);
 *   ext  = pad, then its mirror image pad[2 (W+2) - 1 - j], up to N =
 *          cvGetOptimalDFTSize(W + 2 + L - 1) cells or 2 (W+2), zeros beyond;
 *   y    = circular convolution of ext with the L = 251 taps (DFT product,
 *          inverse DFT scaled by 1/N);
 *   out[i] = y[L/2 + 1 + i].
 * The reference evaluates it with OpenCV's float DFT; this is the same
 * convolution summed directly in double, the FP64 form. */

/* cvGetOptimalDFTSize: the smallest 2^a 3^b 5^c >= n. */
int orc_optimal_dft_size(int n)
{
    for (int m = n > 1 ? n : 1;; m++) {
        int r = m;
        while (r % 2 == 0) r /= 2;
        while (r % 3 == 0) r /= 3;
        while (r % 5 == 0) r /= 5;
        if (r == 1) return m;
    }
}

/* The L taps, centred: SheppLogan / RamLak of filter_prj.cpp:137-178. */
void orc_spatial_taps(int mode, int L, float *h)
{
    if (mode == 0) {
        float n = -((float)L - 1.0) / 2.0;
        for (int i = 0; i < L; i++) {
            h[i] = (float)(-2.0 / (M_PI * (4.0 * n * n - 1.0)));
            n = n + 1.0;
        }
    } else {
        int n = -(L - 1) / 2;
        for (int i = 0; i < L; i++) {
            if (n == 0)
                h[i] = M_PI / 4;
            else if (abs(n % 2) == 1)
                h[i] = -1.0 / (M_PI * n * n);
            else
                h[i] = 0;
            n++;
        }
    }
}

void orc_fbp_filter_f64(double *data, int W, int rows, int mode)
{
    const int L = 251, npad = W + 2;
    const int N = orc_optimal_dft_size(npad + L - 1);
    float h[251];
    orc_spatial_taps(mode, L, h);
    double *ext = (double *)malloc(sizeof(double) * N);
    double *out = (double *)malloc(sizeof(double) * W);
    for (int r = 0; r < rows; r++) {
        double *row = data + (size_t)r * W;
        for (int j = 0; j < N; j++) {
            int p;
            if (j < npad)
                p = j;
            else if (j < 2 * npad)
                p = 2 * npad - 1 - j;
            else
                p = -1;
            if (p < 0) {
                ext[j] = 0;
            } else {
                int x = p - 1;
                x = x < 0 ? 0 : (x > W - 1 ? W - 1 : x);
                ext[j] = row[x];
            }
        }
        for (int i = 0; i < W; i++) {
            double acc = 0;
            for (int k = 0; k < L; k++) {
                int j = (i + L / 2 + 1 - k) % N;
                if (j < 0) j += N;
                acc += (double)h[k] * ext[j];
            }
            out[i] = acc;
        }
        for (int i = 0; i < W; i++) row[i] = out[i];
    }
    free(ext);
    free(out);
}

/* ------------------------------------------------------------------ f2b
 * --f2b byte output as TiltRecZ-cuda writes it (recon_z.cu:253-262,
 * CufloatToByteKernel basic.cu:442-474): mean and standard deviation of the
 * volume (here summed in double and rounded to float; the reference sums
 * in float through thrust), then per value
 *   z = (v - mean) / sd, clamped to [-10, 10], byte = (uchar)(((z + 10) / 20) * 255).
 * stats receives {mean, sd}. */
void orc_bytes_from_stats(const float *v, long n, float mean, float sd, unsigned char *out)
{
    const float scale = 10.0f;
    for (long i = 0; i < n; i++) {
        float z = (v[i] - mean) / sd;
        z = fmaxf(-scale, fminf(scale, z));
        z = ((z + scale) / (2.0f * scale)) * 255.0f;
        out[i] = (unsigned char)z;
    }
}

void orc_float_to_byte(const float *v, long n, unsigned char *out, float *stats)
{
    double s = 0;
    for (long i = 0; i < n; i++) s += v[i];
    const float mean = (float)(s / (double)n);
    double q = 0;
    for (long i = 0; i < n; i++) {
        const float d = v[i] - mean;
        q += (double)d * (double)d;
    }
    const float sd = (float)sqrt(q / (double)n);
    orc_bytes_from_stats(v, n, mean, sd, out);
    stats[0] = mean;
    stats[1] = sd;
}

/* ------------------------------------------------------------------ ADMM
 * TiltRecZ-mpi's ADMM (main-mpiz.cpp:603-1013), in the y-major layout of
 * this file (the reference's z-major loops visit the same terms):
 *   htb  = sum_theta A_theta^T b_theta          (Reproject_admm_htb: tap
 *          weight times pixel, no normalisation)
 *   op(x) = sum_theta A_theta^T [(A_theta x) / w_theta] + mu x
 *          (Reproject_admm_atax + ATAmuLTL; w_theta = A_theta 1, and a tap
 *          contributes only where |w| > 10e-6)
 *   outer loop: htb += (u - d) mu; x0 = op(vol); applycg; u = soft(vol + d);
 *               d += vol - u
 *   applycg: r = htb - x0, p = r; per iteration h = op(p),
 *            alpha = (r.r) / (h.p), vol += alpha p, r = htb - op(vol),
 *            beta = (r.h) / (h.p), p = r + beta p; alpha, beta = 0 when
 *            |h.p| <= 10e-6; dot products of float products summed in
 *            double, the updates evaluated in double.
 * htb keeps its updates across outer iterations, as the reference's does. */

static void admm_op_f32(const float org[3], const double *coef, int P, int W, int len, int T, const float *x,
                        float mu, float *out, float *ax, float *w)
{
    size_t px = (size_t)len * W, size = px * T;
    memset(out, 0, sizeof(float) * size);
    for (int idx = 0; idx < P; idx++) {
        const double *c = coef + 8 * (size_t)idx;
        memset(ax, 0, sizeof(float) * px);
        memset(w, 0, sizeof(float) * px);
        reproject_f32(org, x, c, ax, w, T, len, W);
        for (int y = 0; y < len; y++)
            for (int z = 0; z < T; z++)
                for (int xx = 0; xx < W; xx++) {
                    orc_w t;
                    valcoef(org, xx, y, z, c, &t);
                    float *o = out + ((size_t)y * T + z) * W + xx;
                    const int xs[4] = {t.xm, t.xm + 1, t.xm, t.xm + 1};
                    const int ys[4] = {t.ym, t.ym, t.ym + 1, t.ym + 1};
                    const float wx[4] = {1 - t.xd, t.xd, 1 - t.xd, t.xd};
                    const float wy[4] = {1 - t.yd, 1 - t.yd, t.yd, t.yd};
                    for (int k = 0; k < 4; k++) {
                        if (xs[k] < 0 || xs[k] >= W || ys[k] < 0 || ys[k] >= len) continue;
                        const size_t n = (size_t)xs[k] + (size_t)ys[k] * W;
                        if (fabsf(w[n]) > 10e-6) *o += wx[k] * wy[k] * ax[n] / w[n];
                    }
                }
    }
    for (size_t i = 0; i < size; i++) out[i] += mu * x[i];
}

void orc_admm_f32(const double *coef, int P, int W, int len, int T, const float org[3], const float *proj,
                  int outer, int cgiter, float mu, float soft, float *vol)
{
    size_t px = (size_t)len * W, size = px * T;
    float *htb = (float *)calloc(size, sizeof(float));
    float *u = (float *)calloc(size, sizeof(float));
    float *d = (float *)calloc(size, sizeof(float));
    float *x0 = (float *)malloc(sizeof(float) * size);
    float *r = (float *)malloc(sizeof(float) * size);
    float *p = (float *)malloc(sizeof(float) * size);
    float *h = (float *)malloc(sizeof(float) * size);
    float *ax = (float *)malloc(sizeof(float) * px);
    float *w = (float *)malloc(sizeof(float) * px);
    for (int idx = 0; idx < P; idx++) {
        const double *c = coef + 8 * (size_t)idx;
        const float *one = proj + px * idx;
        for (int y = 0; y < len; y++)
            for (int z = 0; z < T; z++)
                for (int xx = 0; xx < W; xx++) {
                    orc_w t;
                    valcoef(org, xx, y, z, c, &t);
                    float *o = htb + ((size_t)y * T + z) * W + xx;
                    const int xs[4] = {t.xm, t.xm + 1, t.xm, t.xm + 1};
                    const int ys[4] = {t.ym, t.ym, t.ym + 1, t.ym + 1};
                    const float wx[4] = {1 - t.xd, t.xd, 1 - t.xd, t.xd};
                    const float wy[4] = {1 - t.yd, 1 - t.yd, t.yd, t.yd};
                    for (int k = 0; k < 4; k++) {
                        if (xs[k] < 0 || xs[k] >= W || ys[k] < 0 || ys[k] >= len) continue;
                        *o += wx[k] * wy[k] * one[(size_t)xs[k] + (size_t)ys[k] * W];
                    }
                }
    }
    for (int it = 0; it < outer; it++) {
        for (size_t i = 0; i < size; i++) htb[i] += (u[i] - d[i]) * mu;
        admm_op_f32(org, coef, P, W, len, T, vol, mu, x0, ax, w);
        for (size_t i = 0; i < size; i++) {
            r[i] = htb[i] - x0[i];
            p[i] = r[i];
        }
        for (int cg = 0; cg < cgiter; cg++) {
            admm_op_f32(org, coef, P, W, len, T, p, mu, h, ax, w);
            double d0 = 0, d2 = 0, d1 = 0, alpha = 0, beta = 0;
            for (size_t i = 0; i < size; i++) {
                d0 += r[i] * r[i];
                d2 += h[i] * p[i];
            }
            if (d2 > 10e-6 || -d2 > 10e-6) alpha = d0 / d2;
            for (size_t i = 0; i < size; i++) vol[i] = vol[i] + alpha * p[i];
            admm_op_f32(org, coef, P, W, len, T, vol, mu, x0, ax, w);
            for (size_t i = 0; i < size; i++) {
                r[i] = htb[i] - x0[i];
                d1 += r[i] * h[i];
            }
            if (d2 > 10e-6 || -d2 > 10e-6) beta = d1 / d2;
            for (size_t i = 0; i < size; i++) p[i] = r[i] + beta * p[i];
        }
        for (size_t i = 0; i < size; i++) {
            if (vol[i] + d[i] < -soft)
                u[i] = vol[i] + d[i] + soft;
            else if (vol[i] + d[i] > soft)
                u[i] = vol[i] + d[i] - soft;
            else
                u[i] = 0;
        }
        for (size_t i = 0; i < size; i++) d[i] += vol[i] - u[i];
    }
    free(htb);
    free(u);
    free(d);
    free(x0);
    free(r);
    free(p);
    free(h);
    free(ax);
    free(w);
}

/* The same in double throughout (weights before the float rounding). */
static void admm_op_f64(const float org[3], const double *coef, int P, int W, int len, int T, const double *x,
                        double mu, double *out, double *ax, double *w)
{
    size_t px = (size_t)len * W, size = px * T;
    memset(out, 0, sizeof(double) * size);
    for (int idx = 0; idx < P; idx++) {
        const double *c = coef + 8 * (size_t)idx;
        memset(ax, 0, sizeof(double) * px);
        memset(w, 0, sizeof(double) * px);
        reproject_f64(org, x, c, ax, w, T, len, W);
        for (int y = 0; y < len; y++)
            for (int z = 0; z < T; z++)
                for (int xx = 0; xx < W; xx++) {
                    orc_w t;
                    valcoef(org, xx, y, z, c, &t);
                    double *o = out + ((size_t)y * T + z) * W + xx;
                    const int xs[4] = {t.xm, t.xm + 1, t.xm, t.xm + 1};
                    const int ys[4] = {t.ym, t.ym, t.ym + 1, t.ym + 1};
                    const double wx[4] = {1 - t.xdd, t.xdd, 1 - t.xdd, t.xdd};
                    const double wy[4] = {1 - t.ydd, 1 - t.ydd, t.ydd, t.ydd};
                    for (int k = 0; k < 4; k++) {
                        if (xs[k] < 0 || xs[k] >= W || ys[k] < 0 || ys[k] >= len) continue;
                        const size_t n = (size_t)xs[k] + (size_t)ys[k] * W;
                        if (fabs(w[n]) > 10e-6) *o += wx[k] * wy[k] * ax[n] / w[n];
                    }
                }
    }
    for (size_t i = 0; i < size; i++) out[i] += mu * x[i];
}

void orc_admm_f64(const double *coef, int P, int W, int len, int T, const float org[3], const double *proj,
                  int outer, int cgiter, double mu, double soft, double *vol)
{
    size_t px = (size_t)len * W, size = px * T;
    double *htb = (double *)calloc(size, sizeof(double));
    double *u = (double *)calloc(size, sizeof(double));
    double *d = (double *)calloc(size, sizeof(double));
    double *x0 = (double *)malloc(sizeof(double) * size);
    double *r = (double *)malloc(sizeof(double) * size);
    double *p = (double *)malloc(sizeof(double) * size);
    double *h = (double *)malloc(sizeof(double) * size);
    double *ax = (double *)malloc(sizeof(double) * px);
    double *w = (double *)malloc(sizeof(double) * px);
    for (int idx = 0; idx < P; idx++) {
        const double *c = coef + 8 * (size_t)idx;
        const double *one = proj + px * idx;
        for (int y = 0; y < len; y++)
            for (int z = 0; z < T; z++)
                for (int xx = 0; xx < W; xx++) {
                    orc_w t;
                    valcoef(org, xx, y, z, c, &t);
                    double *o = htb + ((size_t)y * T + z) * W + xx;
                    const int xs[4] = {t.xm, t.xm + 1, t.xm, t.xm + 1};
                    const int ys[4] = {t.ym, t.ym, t.ym + 1, t.ym + 1};
                    const double wx[4] = {1 - t.xdd, t.xdd, 1 - t.xdd, t.xdd};
                    const double wy[4] = {1 - t.ydd, 1 - t.ydd, t.ydd, t.ydd};
                    for (int k = 0; k < 4; k++) {
                        if (xs[k] < 0 || xs[k] >= W || ys[k] < 0 || ys[k] >= len) continue;
                        *o += wx[k] * wy[k] * one[(size_t)xs[k] + (size_t)ys[k] * W];
                    }
                }
    }
    for (int it = 0; it < outer; it++) {
        for (size_t i = 0; i < size; i++) htb[i] += (u[i] - d[i]) * mu;
        admm_op_f64(org, coef, P, W, len, T, vol, mu, x0, ax, w);
        for (size_t i = 0; i < size; i++) {
            r[i] = htb[i] - x0[i];
            p[i] = r[i];
        }
        for (int cg = 0; cg < cgiter; cg++) {
            admm_op_f64(org, coef, P, W, len, T, p, mu, h, ax, w);
            double d0 = 0, d2 = 0, d1 = 0, alpha = 0, beta = 0;
            for (size_t i = 0; i < size; i++) {
                d0 += r[i] * r[i];
                d2 += h[i] * p[i];
            }
            if (d2 > 10e-6 || -d2 > 10e-6) alpha = d0 / d2;
            for (size_t i = 0; i < size; i++) vol[i] = vol[i] + alpha * p[i];
            admm_op_f64(org, coef, P, W, len, T, vol, mu, x0, ax, w);
            for (size_t i = 0; i < size; i++) {
                r[i] = htb[i] - x0[i];
                d1 += r[i] * h[i];
            }
            if (d2 > 10e-6 || -d2 > 10e-6) beta = d1 / d2;
            for (size_t i = 0; i < size; i++) p[i] = r[i] + beta * p[i];
        }
        for (size_t i = 0; i < size; i++) {
            if (vol[i] + d[i] < -soft)
                u[i] = vol[i] + d[i] + soft;
            else if (vol[i] + d[i] > soft)
                u[i] = vol[i] + d[i] - soft;
            else
                u[i] = 0;
        }
        for (size_t i = 0; i < size; i++) d[i] += vol[i] - u[i];
    }
    free(htb);
    free(u);
    free(d);
    free(x0);
    free(r);
    free(p);
    free(h);
    free(ax);
    free(w);
}
