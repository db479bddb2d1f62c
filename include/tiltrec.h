/*
 * tiltrec.h -- C ABI of the MI355X TiltRec reconstruction library
 * (libtiltrec_hip.so).
 *
 * The reference has no library boundary: its GPU drivers are C++ functions
 * taking references to MPI-backed classes, and they write the output file
 * themselves.  Each entry point below replaces one of them with plain
 * pointers and sizes; paths are relative to the AuTom2 source tree.
 *
 *   reference                                          replaced by
 *   -------------------------------------------------  -------------------------
 *   TranslateAngleToCoefficients + Decorate*           trec_coefficients
 *     (src/TiltRec/src/main/main-mpi.cpp:85-183,
 *      main-cuda.cpp:111-168)
 *   ExeRecY -> CuBackProject / CuFBP / CuSIRT / CuSART trec_reconstruct
 *     (main-cuda.cpp:170-195,
 *      TiltRec_cuda_y/recon_y.cuh:25-40)
 *   CuBackProjectZ / CuSIRTZ / CuSARTZ / CuFBPZ         trec_reconstruct with
 *     (TiltRec_cuda_z/recon_z.cuh:17-39)                 layout TREC_LAYOUT_Z
 *   ApplyFilterInplace(mode 2)                          trec_filter_rweight
 *     (src/TiltRec/src/filter/filter_prj.h:24)
 *   ApplyFilterInplace(modes 0, 1)                      trec_filter_fbp
 *     (filter_prj.cpp:266-446, 738-771)
 *   MRC element reading loops                           trec_convert
 *     (src/TiltRec/src/mrcmx/mrcstack.cpp:354-459,
 *      mrctools/src/mrcimg/mrc2img.cpp:46-126)
 *   --f2b: thrust moments + CufloatToByteKernel         trec_float_to_byte
 *     (TiltRec_cuda_z/recon_z.cu:253-262,
 *      TiltRec_cuda_y/basic.cu:442-474)
 *   CuReprojectKernel + CuReWeightKernel                trec_forward_project
 *     (TiltRec_cuda_y/basic.cu:237-264,357-382)
 *   CuBackProjKernel (basic.cu:383-415)                 trec_backproject
 *   one iteration of CuSIRT's loop (recon_y.cu:347-382) trec_sirt_iterate
 *   (none: AuTom2 launches one process per series)      trec_reconstruct_batch,
 *                                                        trec_server_*
 *   mrcstack's per-section cv::warpAffine loop          trec_apply_xf
 *     (src/mrctools/src/mrcstack/main.cpp:40-80)
 *   GeoParGen's per-view ViewSlicer loop and averages   trec_geopar
 *     (src/GeoParGen/src/main/main.cpp:71-215,
 *      main/mrcview.h:43-147, mrcview.cpp:13-127)
 *
 * Conventions
 *   - Every array argument is a device pointer (hipMalloc'd memory on the
 *     context's device) unless the name says host.  The caller owns all of
 *     them; the library owns only its workspace, which a context caches.
 *   - stack is [nproj][ny][nx] float (an MRC tilt stack, section = tilt).
 *   - volume layout TREC_LAYOUT_Y is [ny][thickness][nx] (TiltRec-*: MRC
 *     (nx, ny, nz) = (X, thickness, Y)); TREC_LAYOUT_Z is
 *     [thickness][ny][nx] (TiltRecZ-*: MRC (X, Y, thickness)).
 *   - stream is a hipStream_t passed as void* (NULL = the null stream).
 *     Work is enqueued on it; calls return before it finishes unless the
 *     description says otherwise.
 *   - A context is used by one host thread at a time.  Different contexts
 *     are independent; a server handle is thread-safe.
 */
#ifndef TILTREC_H
#define TILTREC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    TREC_OK = 0,
    TREC_ERR_INVALID = 1,     /* an argument is out of range or inconsistent */
    TREC_ERR_UNSUPPORTED = 2, /* valid for the reference, not provided here */
    TREC_ERR_NOMEM = 3,       /* device or host allocation failed */
    TREC_ERR_DEVICE = 4,      /* the HIP runtime reported an error */
    TREC_ERR_IO = 5,          /* file could not be read or written */
    TREC_ERR_BUSY = 6,        /* server: over the memory budget or queue full */
    TREC_ERR_NOT_READY = 7    /* server: the request has not finished */
} trec_status;

typedef enum {
    TREC_BPT = 0,  /* plain back-projection */
    TREC_FBP = 1,  /* Shepp-Logan filtered back-projection */
    TREC_WBP = 2,  /* R-weighted back-projection */
    TREC_SIRT = 3,
    TREC_SART = 4,
    TREC_ADMM = 5  /* TiltRecZ-mpi's ADMM (main-mpiz.cpp:958-1013), any geometry */
} trec_method_kind;

typedef enum {
    TREC_STORE_F32 = 0,  /* projections and residuals in float */
    TREC_STORE_F16 = 1,  /* stored as IEEE half, accumulated in float */
    TREC_STORE_BF16 = 2  /* stored as bfloat16, accumulated in float */
} trec_storage;

typedef enum {
    TREC_LAYOUT_Y = 0, /* [ny][thickness][nx] */
    TREC_LAYOUT_Z = 1  /* [thickness][ny][nx] */
} trec_layout;

/* Which binary's numeric rules to follow where the reference binaries
 * differ: the y origin (T/2 for TiltRec-mpi, Y/2 otherwise; it matters only
 * with a pitch) and the weight thresholds (exact zero tests in TiltRec-mpi,
 * main-mpi.cpp:442,571,600; 1e-5 / 1e-6 / 1e-7 in TiltRec-cuda,
 * basic.cu:170,184,409,434). */
typedef enum {
    TREC_RULES_MPI = 0,
    TREC_RULES_CUDA = 1
} trec_rules;

/* Geometry of one tilt series: the stack size, the volume thickness and the
 * --geometry offset,pitch,zshift,thickness option (opts.cpp:232-260). */
typedef struct {
    int nx;          /* X: detector width = volume width */
    int ny;          /* Y: detector height = number of volume rows */
    int nproj;       /* P: number of tilts */
    int thickness;   /* T: volume thickness in voxels */
    float offset;    /* tilt-angle offset, degrees */
    float pitch;     /* tilt of the specimen about x, degrees */
    float zshift;    /* shift along z, voxels */
    int rules;       /* trec_rules */
} trec_geom;

/* --method NAME,iterations,gamma or ADMM,iterations,cgiter,gamma,soft
 * (opts.cpp:287-335) plus the forms. */
typedef struct {
    int method;      /* trec_method_kind */
    int iterations;  /* SIRT/SART iterations; ADMM outer iterations */
    float gamma;     /* relaxation; ADMM's penalty mu */
    int storage;     /* trec_storage (SIRT/SART); BPT/FBP/WBP/ADMM use float */
    int layout;      /* trec_layout of `volume` and `initial` */
    int cg_iterations; /* ADMM: conjugate-gradient steps per outer iteration */
    float soft;        /* ADMM: soft-threshold */
} trec_method;

typedef struct trec_context trec_context;

/* Opens a context on HIP device `device`. */
trec_status trec_context_create(int device, trec_context **out);
void trec_context_destroy(trec_context *ctx);

/* Releases the cached workspace of a context (it is re-created on demand). */
trec_status trec_context_trim(trec_context *ctx);

/* Static description of a status code. */
const char *trec_status_string(trec_status s);

/* Message of the last failure on this context (empty if none). */
const char *trec_last_error(const trec_context *ctx);

/* Library version string, e.g. "tiltrec-hip 1.0 gfx950". */
const char *trec_version(void);

/* Host: per-tilt map coefficients as TiltRec-mpi builds them
 * (main-mpi.cpp:85-183): coef_out[8*i + 0..3] = a0..a3 (detector x),
 * [8*i + 4..7] = b0..b3 (detector y).  angles_deg is host memory. */
trec_status trec_coefficients(const trec_geom *geom, const float *angles_deg,
                              double *coef_out);

/* Device workspace (bytes) trec_reconstruct needs for this geometry and
 * method, beyond the caller's stack and volume. */
trec_status trec_workspace_size(const trec_geom *geom, const trec_method *method,
                                size_t *bytes);

/* Workspace budget of one trec_reconstruct call on this context, bytes
 * (0, the default: 90 % of the device memory free at the call plus the
 * workspace the context already holds).  A pitch-0 series whose workspace
 * (trec_workspace_size) exceeds it is reconstructed in y slabs, one after
 * another in a workspace that fits, as TiltRec-cuda streams y slabs
 * (recon_y.cu:33-62, 312-420); the volume is the same, bit for bit. */
trec_status trec_set_workspace_limit(trec_context *ctx, size_t bytes);

/* How trec_reconstruct will run a series on this context (after any
 * autotuning), for tools and benchmarks. */
typedef struct {
    int path;          /* 0: row-blocked (pitch 0, |tilt| <= ~78 deg), 1: general */
    int zchunks;       /* z chunks of the SIRT forward projection; 1: the
                          forward writes the residual itself */
    int band_cells;    /* LDS band of the tilt-group forward: 352 or 400 cells
                          (0: no forward projection) */
    int back_tilts;    /* tilts the row-path back-projection stages per pass */
    int back_voxels_z; /* voxels per thread along z of its tile (the tile is
                          64 x 4 back_voxels_z voxels of 16 rows) */
    int slabs;         /* y slabs the volume is reconstructed in */
    int slab_rows;     /* volume rows per slab */
    size_t workspace;  /* device workspace bytes of the call */
} trec_plan_info;
trec_status trec_plan(trec_context *ctx, const trec_geom *geom,
                      const float *angles_deg, const trec_method *method,
                      trec_plan_info *info);

/* Reconstructs one tilt series.  stack is read only.  initial (may be NULL)
 * seeds iterative methods and is the starting volume of BPT/WBP, as
 * --initial does in TiltRec-mpi (main-mpi.cpp:778-791); it may alias
 * volume.  angles_deg is host memory with geom->nproj values. */
trec_status trec_reconstruct(trec_context *ctx, const trec_geom *geom,
                             const float *angles_deg, const float *stack,
                             const float *initial, const trec_method *method,
                             float *volume, void *stream);

/* R-weighting of nrows rows of nx floats in place (ApplyFilterInplace mode 2,
 * filter_prj.cpp:652-802). */
trec_status trec_filter_rweight(trec_context *ctx, float *rows, int nx,
                                long long nrows, void *stream);

/* trec_filter_rweight for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The same filter; its kernel
 * table is built on the device in a fixed array of the library (128 KiB per
 * device, held until the process ends), so the call allocates nothing.
 * Rows are limited to nx + nx / 10 <= 8192 taps (nx <= 7448, the widest rows
 * the filter holds in LDS); wider rows return TREC_ERR_INVALID.  All calls on one device share that array and it
 * keeps the last call's table: calls that overlap on one device (different
 * streams or threads) race, so issue them on one stream. */
trec_status trec_filter_rweight_static(trec_context *ctx, float *rows, int nx,
                                       long long nrows, void *stream);

/* FBP's spatial filter of nrows rows of nx floats in place (ApplyFilterInplace
 * modes 0 and 1, filter_prj.cpp:137-178, 266-446): mode 0 Shepp-Logan (what
 * FBP uses), 1 Ram-Lak; 251 taps over the reference's symmetrised and
 * mirror-padded row.  Rows up to about 9,900 pixels. */
trec_status trec_filter_fbp(trec_context *ctx, float *rows, int nx,
                            long long nrows, int mode, void *stream);

/* Element types of MRC data as trec_convert takes them. */
typedef enum {
    TREC_DT_I8 = 0,         /* MRC mode 0 as TiltRec reads it: signed bytes */
    TREC_DT_U8 = 1,         /* unsigned bytes (MRC mode 101 unpacked) */
    TREC_DT_U8_DIV255 = 2,  /* MRC mode 0 as mrctools / GeoParGen read it:
                               unsigned bytes / 255.0f */
    TREC_DT_I16 = 3,        /* MRC mode 1 */
    TREC_DT_U16 = 4,        /* MRC mode 6 */
    TREC_DT_F16 = 5,        /* MRC mode 12 */
    TREC_DT_F32 = 6         /* MRC mode 2 */
} trec_dtype;

/* Converts n device elements of in_type into out_type (TREC_DT_F32 or
 * TREC_DT_F16) on the stream: integers and halves to float exactly, bytes
 * over 255 as the mrctools reader divides them (mrc2img.cpp:60-75), float to
 * half rounded to nearest even.  Replaces the reference readers' per-element
 * host loops (mrcstack.cpp:354-459, mrc2img.cpp:46-126), so a CLI uploads
 * the file's own bytes.  in and out must not overlap. */
trec_status trec_convert(trec_context *ctx, const void *in, int in_type,
                         void *out, int out_type, size_t n, void *stream);

/* --f2b byte output (TiltRecZ-cuda, recon_z.cu:253-262; CufloatToByteKernel,
 * basic.cu:442-474): the mean and standard deviation of the n floats of `in`,
 * then out[i] = (unsigned char)(((clamp((in[i] - mean) / sd, -10, 10) + 10)
 * / 20) * 255).  The moments are summed in double and rounded to float.
 * stats_host (may be NULL) receives {mean, sd}; the call then waits for the
 * stream. */
trec_status trec_float_to_byte(trec_context *ctx, const float *in, size_t n,
                               unsigned char *out, float *stats_host,
                               void *stream);

/* Forward projection of a layout-Y volume: rval[p][y][x] = splat of the
 * volume, rwt[p][y][x] = splat of ones (Reproject, main-mpi.cpp:286-344).
 * rwt may be NULL. */
trec_status trec_forward_project(trec_context *ctx, const trec_geom *geom,
                                 const float *angles_deg, const float *volume,
                                 float *rval, float *rwt, void *stream);

/* Back-projection of a stack into a layout-Y volume, added to it:
 * volume += sum over tilts of s/c (BackProject, main-mpi.cpp:411-450). */
trec_status trec_backproject(trec_context *ctx, const trec_geom *geom,
                             const float *angles_deg, const float *stack,
                             float *volume, void *stream);

/* One SIRT iteration on a layout-Y volume in place (main-mpi.cpp:627-658). */
trec_status trec_sirt_iterate(trec_context *ctx, const trec_geom *geom,
                              const float *angles_deg, const float *stack,
                              float gamma, float *volume, void *stream);

/* .xf application: what mrcstack does before TiltRec runs
 * (src/mrctools/src/mrcstack/main.cpp:40-80).  Section i of `in` ([nsec][ny][nx]
 * float) is warped by line i of an IMOD-style .xf table, xf_host[6 i ..
 * 6 i + 5] = a11 a12 a21 a22 dx dy (host memory), about the centre
 * (nx / 2, ny / 2):
 *   out_i(x, y) = in_i(M^-1 (x, y)),  M = [[a11, a12, cx + dx - a11 cx - a12 cy],
 *                                          [a21, a22, cy + dy - a21 cx - a22 cy]]
 * with cv::warpAffine's arithmetic: 1/32-pixel fixed-point source
 * coordinates, INTER_LANCZOS4 (8 x 8) when mode == 2 and INTER_CUBIC
 * (4 x 4) otherwise, and the mean of in_i as the constant outside the
 * section.  in and out must not overlap.  Returns TREC_ERR_INVALID for
 * sections above 2^29 pixels, a non-finite .xf value or one that maps
 * pixels 2^20 or more pixels away. */
trec_status trec_apply_xf(trec_context *ctx, int nx, int ny, int nsec,
                          const float *xf_host, int mode, const float *in,
                          float *out, void *stream);

/* GeoParGen (src/GeoParGen/src/main/main.cpp:71-215; ViewSlicer,
 * mrcview.h:43-147; mrcview.cpp:13-127): the tilt, thickness and centre of
 * the specimen slab in a volume of nz sections of ny x nx floats (an MRC
 * volume in file order; TiltRec-* output is (X, thickness, Y)).
 *   axis TREC_GP_VIEW_X (GeoParGen's default "x"): nx views, view j the
 *   ny x nz plane at column j; TREC_GP_VIEW_Y ("y"): the nz sections.
 * Each view is normalised to [0, 1], cut into chips of 50 columns every 50,
 * and each chip's row profile is matched against every rectangle that is
 * zero on rows [l, r], r >= l + 30, by |NCC|, with the reference's float
 * running maximum deciding near-ties exactly as its scan does.  The view's
 * angle is the atan, in degrees, of the slope of z = (l + r) / 2 over the
 * chip positions, weighted by (ncc / sum ncc)^2; the same weights average
 * thickness r - l and z.
 * Outputs (all optional):
 *   views  device [nviews][3] floats: angle, thickness, z per view;
 *   chips  device [nviews][nchips][4] floats: z, thickness, ncc, l per
 *          chip (-1, -1, 0, -1 when nothing was found);
 *   result_host  host [3]: the printed averages (float sums over views
 *          divided by the view count);
 *   exhaustive_host  host: chips whose search needed the full exact scan.
 * With result_host or exhaustive_host the call waits for the stream and
 * returns TREC_ERR_INVALID if the volume held non-finite values.
 * flags: TREC_GP_EXHAUSTIVE evaluates every pair of every chip exactly (the
 * reference's own cost; the default gives the same results).
 * Limits: ny <= 4096; y views need nz <= 65535. */
enum { TREC_GP_VIEW_X = 0, TREC_GP_VIEW_Y = 1 };
enum { TREC_GP_EXHAUSTIVE = 1 };
trec_status trec_geopar(trec_context *ctx, const float *volume, int nx,
                        int ny, int nz, int axis, int flags, float *views,
                        float *chips, float *result_host, int *exhaustive_host,
                        void *stream);

/* One job of a batch: a tilt series and where its volume goes. */
typedef struct {
    trec_geom geom;
    trec_method method;
    const float *angles_deg; /* host */
    const float *stack;      /* device [nproj][ny][nx] */
    const float *initial;    /* device or NULL */
    float *volume;           /* device */
} trec_job;

/* Reconstructs njobs series of any mix of sizes and methods together: every
 * phase of every job is issued in one multi-job kernel per phase, and the
 * whole schedule is captured in one HIP graph and launched once.  Each job's
 * result equals what trec_reconstruct gives for it alone. */
trec_status trec_reconstruct_batch(trec_context *ctx, int njobs,
                                   const trec_job *jobs, void *stream);

/* Autotuning, replacing TiltRec-cuda's fixed slab heuristic InitYMAX
 * (recon_y.cu:75-86): for a SIRT geometry, times whole iterations with the
 * forward projection's z chunks 1, 2, 4, ... and then with the
 * back-projection's tile shapes (tilts staged per pass, voxels per thread
 * along z), and keeps the fastest in the context and in the process-wide
 * table later contexts read (trec_plan shows the choice).  A candidate
 * replaces an earlier, preferred one only when it is more than 2 % faster,
 * so the choice is stable from run to run.  Synchronises the device. */
trec_status trec_autotune(trec_context *ctx, const trec_geom *geom,
                          const trec_method *method);

/* Per-kernel timing: while enabled, every launch group of the
 * reconstruction calls is bracketed by HIP events on the stream it runs on
 * (HIP graphs are not used meanwhile).  kernel is one of "prep" (layout
 * transposes, filter, weights), "forward", "combine", "back", "finish",
 * "xf_mean" (trec_apply_xf's coordinate terms and section means),
 * "xf_warp" (its warp), "gp_profile" (trec_geopar's normalisation and chip
 * profiles), "gp_search" (its pair search) and "gp_views" (its
 * regressions and averages);
 * trec_profile_read waits for the recorded events and returns the summed
 * milliseconds and the number of launch groups since enabling. */
trec_status trec_profile_enable(trec_context *ctx, int on);
trec_status trec_profile_read(trec_context *ctx, const char *kernel,
                              double *total_ms, long long *launches);

/* ---------------------------------------------------------------- serving
 * A server keeps a request loop on its own host thread and one device
 * arena of hbm_budget bytes.  Requests carry host stacks; the loop admits
 * queued requests first come, first served while at most max_active are
 * active and the request's footprint (trec_request_footprint) fits the
 * arena, copies the admitted stack in, and then advances every active
 * request by one stage per step (one SIRT iteration, one SART tilt, one
 * BPT/WBP pass) in shared multi-job launches.  Requests arriving meanwhile
 * join at the next step.  A request's volume is the same, bit for bit,
 * whether it ran alone or with others. */
typedef struct trec_server trec_server;

typedef struct {
    trec_geom geom;
    trec_method method;
    const float *angles_deg; /* host, copied at submit */
    const void *stack;       /* host [nproj][ny][nx] float (IEEE half with
                                stack_half); must stay valid until the
                                request is done */
    void *volume_out;        /* host, float (half with volume_half), written
                                when the request completes */
    int stack_half;          /* 1: the stack is IEEE half (MRC mode 12); it
                                crosses PCIe at half the bytes and is widened
                                to float on the device */
    int volume_half;         /* 1: the volume comes back as IEEE half, rounded
                                to nearest even on the device (MRC mode 12
                                output), at half the bytes */
} trec_request;

typedef struct {
    trec_status status;
    double submit_s;   /* host clock when submitted */
    double admit_s;    /* when admitted to the device */
    double done_s;     /* when its volume was back on the host */
    size_t footprint;  /* device bytes it held */
    double staged_s;   /* when its stack was on the device */
    double computed_s; /* when its volume was final on the device (the
                          first chunk of the copy out had arrived) */
} trec_result;

/* hbm_budget = 0 uses 90 % of the device's free memory; max_active is at
 * most 64. */
trec_status trec_server_create(int device, size_t hbm_budget, int max_active,
                               trec_server **out);
/* Queues a request; TREC_ERR_BUSY if its footprint exceeds the whole
 * budget (it could never be admitted). */
trec_status trec_server_submit(trec_server *srv, const trec_request *req,
                               uint64_t *id_out);
/* *done = 1 once the request finished (successfully or not). */
trec_status trec_server_poll(trec_server *srv, uint64_t id, int *done);
/* Blocks until the request finished and returns its record and status;
 * the id is released afterwards. */
trec_status trec_server_wait(trec_server *srv, uint64_t id, trec_result *res);
/* Device bytes a request with float input and output would hold (the
 * admission footprint, an upper bound over the row and the general path);
 * half input or output adds 2 bytes per stack or volume element. */
trec_status trec_request_footprint(const trec_geom *geom,
                                   const trec_method *method, size_t *bytes);
void trec_server_destroy(trec_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* TILTREC_H */
