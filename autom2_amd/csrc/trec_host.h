// trec_host.h -- host-side job plans shared by api.cpp and server.cpp.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "trec_internal.h"

struct trec_context;

namespace trec {

// One reconstruction: its arguments, the per-tilt constants built on the
// host, and (once carved) the device buffers it works in.
struct JobPlan {
    trec_geom g{};
    trec_method m{};
    const float *stack = nullptr;
    const float *initial = nullptr;
    float *volume = nullptr;
    int path = PATH_ROWS;
    int stages = 0;         // SIRT: iterations; SART: iterations * tilts; BPT/WBP: 1
    bool want_rwt = false;  // keep the raw ray weights (forward-projection API)
    std::vector<double> coef;
    // row-blocked path.  A job may run as nslab y slabs of slab_rb row
    // blocks, one after another in the same workspace (single-job calls whose
    // whole workspace exceeds the limit): h and d then describe the slab being
    // issued (select_slab), d_all holds every slab's descriptor.
    JobDesc h{};
    JobDesc *d = nullptr;
    JobDesc *d_all = nullptr;
    int nyb_total = 0;
    int slab_rb = 0;
    int nslab = 1;
    JobDesc *d_ones = nullptr;
    std::vector<TiltConst> tc;
    bool fg_small = false;  // every union of the forward fits the FGS_SMALL band
    double *sumsq = nullptr;
    float *hk = nullptr;
    int N = 0;
    std::vector<float> hk_host;
    // general path
    GeneralJob gj{};
    std::vector<TiltGeneral> tg;
    TiltGeneral *d_tg = nullptr;
    float *stack_copy = nullptr;
    // ADMM: vectors and scalars, and the descriptors of its passes
    // (0: htb = A^T b; 1: op(vol) -> ax; 2: op(p) -> h)
    AdmmVecs av{};
    JobDesc *d_adm[3] = {nullptr, nullptr, nullptr};
    JobDesc h_adm[3]{};
    size_t bytes = 0;
};

trec_status fail(trec_context *ctx, trec_status s, const char *fmt, ...);
void clear_error(trec_context *ctx);
int context_device(const trec_context *ctx);
// rweight_static.hip: takes the device address of the no-allocation entry
// points' file-scope array on `device`, which must be current (once per
// device; trec_context_create calls it).
hipError_t static_scratch_init(int device);

// Coefficients of main-mpi.cpp:85-183 (coef: 8 per tilt).
void coefficients(const trec_geom &g, const float *angles, double *coef);
// Circulant R-weighting kernel of length N = X + X/10.
std::vector<float> rweight_kernel(int X, int &N);
// FBP taps (mode 0 Shepp-Logan, 1 Ram-Lak) and the padded length N.
std::vector<float> spatial_taps(int mode, int X, int &N);

trec_status validate(trec_context *ctx, const trec_geom *g, const trec_method *m, const float *angles);
// force_nzc > 0 fixes the SIRT forward's z chunks (else the tuned or default
// count); force_general takes the general path whatever the geometry.
trec_status plan_job(trec_context *ctx, JobPlan &p, const trec_geom &g, const trec_method &m, const float *angles,
                     const float *stack, const float *initial, float *volume, int force_nzc = 0,
                     bool force_general = false);
// Assigns the plan's device buffers from `base` (nullptr: sizing only).
size_t carve(JobPlan &p, char *base);
// Copies the plans' descriptors to the device through pinned staging.
trec_status upload(trec_context *ctx, const std::vector<JobPlan *> &plans, hipStream_t st);

// Splits a row-path job into y slabs whose workspace fits `limit` bytes
// (false: not even one row block fits).  Only single-job calls use slabs.
bool split_slabs(JobPlan &p, size_t limit);
// Points h and d at slab s.
void select_slab(JobPlan &p, int s);

hipError_t issue_prep(JobPlan &p, hipStream_t st);
// Advances every job whose stage[j] < stages by one stage, merging the jobs'
// kernels into multi-job launches; increments stage[j].
hipError_t issue_step(const std::vector<JobPlan *> &plans, std::vector<int> &stage, hipStream_t st);
hipError_t issue_finish(JobPlan &p, hipStream_t st);

// Bytes a request holds on the device in the serving loop: workspace plus
// its stack and volume, planned with its own tilt angles; *nzc receives the
// z chunks that plan uses (0 on the general path), which the request keeps.
size_t request_bytes(const trec_geom &g, const trec_method &m, const float *angles, int *nzc);
// The same without angles: the larger of the row and the general path.
size_t request_bytes_bound(const trec_geom &g, const trec_method &m);

}  // namespace trec
