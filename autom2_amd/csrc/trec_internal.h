// trec_internal.h -- types shared by the HIP kernels and the host side of
// libtiltrec_hip.so.  Not part of the public ABI (include/tiltrec.h is).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <utility>
#include <vector>

#include "tiltrec.h"

namespace trec {

// Rows of the volume that one thread carries.  At pitch 0 every row y is an
// independent 2-D problem fed by row y of every projection (the detector y
// coordinate equals y exactly), so a thread processes RB rows with one
// evaluation of the geometry.  Volumes and projections are stored row-blocked:
//   volume  [nyb][T][X][RB]      (nyb = ceil(Y / RB))
//   stack   [nyb][P][X][RB]
// so the RB rows of one (z, x) voxel column or one (tilt, pixel) are 64
// contiguous bytes in float.
constexpr int RB = 16;

// Threads per work-group of the row-blocked kernels.
constexpr int WG = 256;
// Direct forward projection (ray weights, and the TREC_FP_DIRECT reference
// path): FP_PIX pixels of one tilt per work-group, gathered straight from L2.
constexpr int FP_PIX = 256;
// Tilt-group forward projection: FG neighbouring tilts (one wave each) of
// the same FW_PIX pixels share an LDS band of FGS cells per plane, loaded in
// chunks of FW_G cells.  Unions wider than that (tilts past about 60 degrees
// far from the tilt axis) take the direct path.
constexpr int FW_PIX = 64;
constexpr int FW_G = 16;
constexpr int FG = 4;
constexpr int FGS = 400;  // 25 KB of LDS per group: six groups (24 waves) fit a CU
static_assert(FGS % FW_G == 0, "the group band is loaded in whole chunks");
// a band of 22.5 KB: seven groups (28 waves) fit a CU, for launches whose
// unions all fit it
constexpr int FGS_SMALL = 352;
static_assert(FGS_SMALL % FW_G == 0 && FGS_SMALL < FGS, "the small band is loaded in whole chunks");
// Back-projection tile: 64 voxels along x (one wave) by BP_Z along z; each
// thread carries BP_ZT voxels of one x column (z, z+4, ...).  Residuals of
// BP_G tilts are staged in LDS per pass, BP_W cells per tilt: with a1 <= 1
// and |a3| <= 1 the taps of a tile span at most 64 + BP_Z cells.
constexpr int BP_X = 64;
constexpr int BP_ZT = 4;
constexpr int BP_Z = (WG / BP_X) * BP_ZT;
constexpr int BP_G = 8;  // 16-bit storage; float passes stage BP_G32
constexpr int BP_G32 = 6;  // fewer tilts, less LDS: four work-groups per CU
constexpr int BP_W = BP_X + BP_Z + 4;

// Largest number of jobs one multi-job launch carries.
constexpr int MAX_JOBS = 64;

// Per-tilt constants.  a0, a1, a3 are TiltRec-mpi's double coefficients
// (main-mpi.cpp:85-183); the fast path needs a2 = 0 and b = (0, 0, 1, 0),
// which holds exactly whenever the pitch is 0.  Voxel (x, z) maps to the
// detector column
//   u = ((a0 + a1 * (x - ox)) + a3 * (z - oz)) + ox      (ValCoef's order)
// and, for the forward projection, u(x, z) = c0 + a1 * x + a3 * z.
struct TiltConst {
    double a0, a1, a3;
    double ox, oz;
    double c0;      // a0 + ox - a1 * ox - a3 * oz
    double inv_a1;  // 1 / a1
    double slope;   // -a3 / a1: shift of a pixel's voxel window per z step
    float a1f;      // a1 rounded to float
    int ncand;      // voxels a pixel can see in one z layer: ceil(2 / a1)
};

// General geometry (any pitch): u = a0 + a1 X + a2 Y + a3 Z + ox,
// v = b0 + b1 X + b2 Y + b3 Z + oy with X = x - ox etc.
struct TiltGeneral {
    double a[4];
    double b[4];
    double ox, oy, oz;
    // tiles of the splat forward projection (k_general_splat), set by
    // general_tiles: si detector columns per tile, ncp lane phases whose
    // voxels are at least one pixel apart, ntx pairs of tiles across the
    // detector (a wave takes a pair), tile0 the pairs of the tilts before
    // this one (0 for every tilt the splat cannot take)
    int si, ncp, ntx, tile0;
};
// Fills the splat tiles of P tilts (host); false when some tilt needs
// another forward kernel (|a1| < 1/3, b2 < 0.55 or a narrow tile).
bool general_tiles(TiltGeneral *tg, int P, int X);

enum PathKind { PATH_ROWS = 0, PATH_GENERAL = 1 };

// One reconstruction job as the kernels see it (device memory).
struct JobDesc {
    int X, Y, T, P;  // Y: rows of the caller's stack and volume
    int y0;        // first caller row of this slab (slab-wise jobs; else 0)
    int nyb;       // row blocks of the slab (the whole job when y0 = 0 and one slab)
    int nzc;       // z chunks of the forward projection (1 = fused residual)
    int zc_len;    // layers per z chunk
    int storage;   // trec_storage of proj / res
    int bp_g, bp_zt;  // SIRT back-projection shape: tilts per pass, voxels per thread along z (0: default)
    float gamma;
    float thr_upd;  // SIRT: C > thr_upd; BPT/SART: c > thr_upd
    float thr_res;  // residual: rwt > thr_res (0: rwt != 0)
    float scale;    // FP16 form: stack divided by scale on the way in
    const TiltConst *tc;     // [P]
    float *vol;              // [nyb][T][X][RB]
    void *proj;              // [nyb][P][X][RB] of the storage type
    void *res;               // [nyb][P][X][RB] of the storage type
    float *part;             // [nzc][nyb][P][X][RB] when nzc > 1
    float *inv_rwt;          // [P][X]
    float *inv_c;            // [T][X]
    float *rwt;              // [P][X] raw ray weights, or NULL
    // half storage (iterative methods): a half copy of vol for the forward
    // projection, as (c, c+1) pairs: entry c + 1 of quad q of slice (yb, z)
    // holds rows 4q .. 4q+3 of cells c and c+1, c in [-1, X); else NULL
    void *hvol;              // [nyb][T][4][X + 1] x 16 bytes
    // BACK_LIN (ADMM): out = lin_a * sum_theta s_theta(res) + lin_b * vol
    float *out;              // [nyb][T][X][RB]
    float lin_a, lin_b;
};

// A multi-job launch: jobs[k] owns work-groups [off[k], off[k+1]).
struct LaunchList {
    int n;
    unsigned off[MAX_JOBS + 1];
    const JobDesc *job[MAX_JOBS];
    short theta0[MAX_JOBS];  // first tilt of the job in this launch
    short ntheta[MAX_JOBS];  // tilts of the job in this launch, 0 = all
};

// ---------------------------------------------------------------- launchers
// (kernels.hip).  All enqueue on `st` and return the launch error.
hipError_t launch_block_stack(const JobDesc &h, const JobDesc *d, const float *stack, hipStream_t st);
hipError_t launch_block_volume(const JobDesc &h, const JobDesc *d, const float *src, int layout, hipStream_t st);
hipError_t launch_zero(float *p, size_t n, hipStream_t st);
hipError_t launch_unblock_volume(const JobDesc &h, const JobDesc *d, float *dst, int layout, hipStream_t st);
hipError_t launch_weights(const JobDesc &h, const JobDesc *d, hipStream_t st);
// builds hvol from vol
hipError_t launch_shadow(const JobDesc &h, const JobDesc *d, hipStream_t st);
// small: the fused forward with the FGS_SMALL band (every job of the launch
// has fg_small set)
hipError_t launch_forward(const LaunchList &L, unsigned nblk, int storage, bool fused, hipStream_t st,
                          bool small = false);
hipError_t launch_combine(const LaunchList &L, unsigned nblk, int storage, hipStream_t st);
// g, zt: the back-projection shape (tilts staged per pass, voxels per thread
// along z; 0: the default).  Only SIRT has other shapes (back_shapes).
hipError_t launch_back(const LaunchList &L, unsigned nblk, int storage, int mode, hipStream_t st, int g = 0,
                       int zt = 0);
// (tilts per pass, voxels per thread along z) of the SIRT back-projection
// for a storage type; the first is the default.
const std::vector<std::pair<int, int>> &back_shapes(int storage);
hipError_t launch_rweight(float *data, const float *hk, int N, int X, long long nrows,
                          long long row_stride, int x_stride, int rows_per_group,
                          long long group_stride, hipStream_t st);
// FBP spatial filter in place (FBP_TAPS taps, ext length N), same addressing.
constexpr int FBP_TAPS = 251;
hipError_t launch_fbp_filter(float *data, const float *taps, int N, int X, long long nrows, long long row_stride,
                             int x_stride, int rows_per_group, long long group_stride, hipStream_t st);
// --f2b: mean and standard deviation of n floats into stats[0..1] (part
// holds 2048 doubles), then the byte mapping into out.
hipError_t launch_to_byte(const float *a, size_t n, double *part, float *stats, unsigned char *out, hipStream_t st);
// trec_convert's element conversion (types are trec_dtype values).
hipError_t launch_convert(const void *in, int in_type, void *out, int out_type, size_t n, hipStream_t st);
// Sum of squares (in double) of n values of the storage type, as one partial
// per work-group in part[0 .. sumsq_blocks(n)) (at most SUMSQ_BLOCKS).
constexpr int SUMSQ_BLOCKS = 1024;
unsigned sumsq_blocks(size_t n);
hipError_t launch_sumsq(const void *a, int storage, size_t n, double *part, hipStream_t st);

// Work-group counts of the multi-job kernels for one job.
unsigned forward_blocks(const JobDesc &h, int ntheta);
unsigned forward_group_blocks(const JobDesc &h, int ntheta);
// work-groups of a value forward projection for the kernel launch_forward picks
unsigned fwd_value_blocks(const JobDesc &h, int ntheta);
unsigned combine_blocks(const JobDesc &h, int ntheta);
unsigned back_blocks(const JobDesc &h);

// Back-projection modes.
enum BackMode {
    BACK_SIRT = 0,  // val += s over all tilts, then vol += gamma * val / C
    BACK_BPT = 1,   // vol += s / c per tilt (c above the threshold)
    BACK_SART = 2,  // one tilt: vol += gamma * s / c
    BACK_LIN = 3    // out = lin_a * sum s + lin_b * vol, float only (ADMM's operator)
};

// General-geometry kernels (general.hip: natural layouts, any pitch).
struct GeneralJob {
    int X, Y, T, P;
    float gamma, thr_upd, thr_res;
    int layout;             // trec_layout of vol
    const TiltGeneral *tg;  // [P]
    const float *stack;     // [P][Y][X]
    float *vol;             // layout Y [Y][T][X] or Z [T][Y][X]
    float *rval;            // [P][Y][X]
    float *rwt;             // [P][Y][X]
    // BACK_LIN (ADMM): out = lin_a * sum_theta s_theta(img) + lin_b * vol
    float *out;
    float lin_a, lin_b;
    float *part;  // layer-chunk partials: [GZ_CHUNKS][Y][X] of a single-tilt forward (SART), or
                  // [nzc_all][P][Y][X] of an all-tilt splat forward (SIRT); or null
    int nzc_all;  // layer chunks of an all-tilt splat forward into part (1: none)
};
constexpr int GZ_CHUNKS = 4;
// host_tg: the same tilt table on the host, for the choice of kernel
// nzc_used (optional): the caller takes layer-chunk partials in g.part; the
// number of chunks written there (1: the projection is in rval itself)
// The row-window forward for a pitch with an offset (k_general_rows3b): whether
// tilts theta0 .. theta0 + ntheta - 1 of a volume X wide and T thick take it,
// and with which rows per thread (rj), slot offset (dlo), most row shift
// (dmax) and row mode (lane: each lane's rows).
struct Rows3bPlan {
    int rj = 0;
    double dlo = 0.0;
    int dmax = 0;
    bool lane = false;
};
bool rows3b_plan(const TiltGeneral *host_tg, int theta0, int ntheta, int X, int T, int layout, Rows3bPlan &rp);
hipError_t launch_general_forward(const GeneralJob &g, const TiltGeneral *host_tg, int theta0, int ntheta, bool ones,
                                  hipStream_t st, int *nzc_used = nullptr);
// quot: -rval / rwt above the threshold, 0 elsewhere (ADMM's operator on a
// zero stack); else stack - rval / rwt
hipError_t launch_general_residual(const GeneralJob &g, int theta0, int ntheta, bool quot, hipStream_t st,
                                   int nzc = 1);
hipError_t launch_general_back(const GeneralJob &g, const TiltGeneral *host_tg, int mode, int theta0, int ntheta,
                               const float *img, hipStream_t st);

// .xf application (xform.hip).  OpenCV warpAffine's fixed point: 10-bit row
// and column terms, 1/32-pixel steps, a bias of half a step.
constexpr int XF_TAB = 32;
constexpr int XF_AB_SCALE = 1024;
constexpr int XF_ROUND_DELTA = XF_AB_SCALE / XF_TAB / 2;
// Output tile of a warp work-group (one wave: lane = column), the table
// entry pitch (floats; 32-byte entries) and the largest LDS window (floats)
// a launch reserves.  Windows of inner tiles have a fixed row pitch of 72
// or 80 floats (the launch's `pitch`): a 64 x 16 tile at unit scale spans
// 68 columns unturned, 62 at 45 degrees.
constexpr int XF_TW = 64;
constexpr int XF_TH = 16;
constexpr int XF_TP = 8;
constexpr int XF_LDS_MAX = 15360;

// Inverse affine map of one section, warpAffine's m[0..5].
struct XfConst {
    double m[6];
};

struct XfLaunch {
    int W, H, P;
    int nb;       // partial sums per section
    int lanczos;  // 0: 4 x 4 cubic, 1: 8 x 8 Lanczos-4
    int win_cap;  // LDS window of a warp work-group, floats
    int pitch;    // LDS row pitch of inner windows: 72 or 80
    const XfConst *xc;  // [P]
    const float *tab;   // [32][XF_TP] 1-D kernel table
    const float *src;   // [P][H][W]
    float *dst;         // [P][H][W]
    int2 *rows;         // [P][H] row terms (x, y)
    int2 *cols;         // [P][W] column terms (x, y)
    double *part;       // [P][nb]
    float *cval;        // [P] border values
};
hipError_t launch_xf_terms(const XfLaunch &a, hipStream_t st);
hipError_t launch_xf_mean(const XfLaunch &a, hipStream_t st);
hipError_t launch_xf_warp(const XfLaunch &a, hipStream_t st);

// ADMM (admm.hip): the volume-sized vectors of one reconstruction, all in
// the row-blocked layout, and its device scalars.
constexpr int ADMM_DOT_BLOCKS = 2048;
enum AdmmStep { ADMM_HTB = 0, ADMM_R0, ADMM_STEP, ADMM_RES, ADMM_DIR, ADMM_SOFT };
struct AdmmVecs {
    float *vol, *htb, *u, *d, *r, *p, *h, *ax;
    size_t n;      // floats per vector
    double *part;  // [2][ADMM_DOT_BLOCKS]
    double *S;     // r.r, h.p, alpha, r.h, beta
    float mu, soft;
};
hipError_t launch_admm_vec(const AdmmVecs &a, int step, hipStream_t st);
// which 0: r.r and h.p, then alpha; which 1: r.h, then beta
hipError_t launch_admm_dots(const AdmmVecs &a, int which, hipStream_t st);

// GeoParGen (geopar.hip): views of a volume, chips of GP_WIDTH columns every
// GP_STEP, a step-function search of at least GP_MINT + 1 rows per chip.
constexpr int GP_STEP = 50;
constexpr int GP_WIDTH = 50;
constexpr int GP_MINT = 30;
constexpr int GP_MAXN = 4096;  // profile rows the search keeps in LDS
constexpr int GP_WCAP = 256;   // window members checked exactly per chip
struct GpLaunch {
    const float *vol;  // [nz][ny][nx]
    int nx, ny, nz, axis, nviews, cols, nchips, exhaustive;
    float *pmin, *pmax;  // [splits][nviews] partial minima and maxima
    int splits;
    float2 *ab;          // per view: x * a + b normalises to [0, 1]
    float *prof;         // [nviews][nchips][ny]
    float4 *chips;       // [nviews][nchips]: z, thickness, ncc, l
    float *views;        // [nviews][3]: angle, thickness, z
    float *avg;          // [3]
    int *flags;          // [0]: non-finite voxels seen; [1]: chips searched exhaustively
};
hipError_t launch_gp_scale(const GpLaunch &a, hipStream_t st);
hipError_t launch_gp_profiles(const GpLaunch &a, hipStream_t st);
hipError_t launch_gp_search(const GpLaunch &a, hipStream_t st);
hipError_t launch_gp_views(const GpLaunch &a, hipStream_t st);

// Small helpers.
// d->scale = sqrt(sum of launch_sumsq's partials / n), summed in a fixed order
hipError_t launch_set_scale(JobDesc *d, const double *part, size_t n, hipStream_t st);
hipError_t launch_unblock_proj(const float *src, int P, int X, int Y, int nyb, float *dst, hipStream_t st);
hipError_t launch_broadcast_rows(const float *src, int P, int X, int Y, float *dst, hipStream_t st);
hipError_t launch_copy(float *dst, const float *src, size_t n, hipStream_t st);

}  // namespace trec
