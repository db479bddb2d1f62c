// api.cpp -- host side of libtiltrec_hip.so: the C ABI of include/tiltrec.h.
//
// A call validates its arguments, builds each job's per-tilt constants on
// the host (the coefficient arithmetic of main-mpi.cpp:85-183), carves the
// job's buffers out of the context's cached device arena, uploads the small
// descriptors through pinned staging, and issues the kernels of kernels.hip.
// Work is organised in stages (one SIRT iteration, one SART tilt, one
// back-projection pass); jobs of a batch run stage by stage, every stage as
// one multi-job launch per kernel, and the whole schedule is captured in a
// HIP graph that is cached on the context and replayed while nothing
// changes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "trec_internal.h"
#include "trec_host.h"

using namespace trec;

namespace {

constexpr float PI_180 = 0.01745329252f;  // main-mpi.cpp:9
constexpr size_t ALIGN = 256;

size_t align_up(size_t v) { return (v + ALIGN - 1) & ~(ALIGN - 1); }

// What trec_autotune chooses for a geometry and storage: the SIRT forward's
// z chunks and the SIRT back-projection's shape.
struct Tuned {
    int nzc = 0;            // 0: not tuned
    int bp_g = 0, bp_zt = 0;  // 0: the default
};

std::mutex g_tune_mu;
std::map<std::string, Tuned> g_tuned;  // process-wide autotune results

std::string tune_key(const trec_geom &g, int storage)
{
    char b[96];
    snprintf(b, sizeof b, "%d:%d:%d:%d:%d", g.nx, g.ny, g.thickness, g.nproj, storage);
    return b;
}

}  // namespace

namespace trec {

// Event timing of launch groups (trec_profile_*).
enum ProfKind {
    PROF_PREP = 0,
    PROF_FORWARD,
    PROF_COMBINE,
    PROF_BACK,
    PROF_FINISH,
    PROF_XF_MEAN,
    PROF_XF_WARP,
    PROF_GP_PROFILE,
    PROF_GP_SEARCH,
    PROF_GP_VIEWS,
    PROF_N
};
const char *const prof_names[PROF_N] = {"prep",    "forward", "combine",    "back",      "finish",
                                        "xf_mean", "xf_warp", "gp_profile", "gp_search", "gp_views"};

struct Prof {
    bool on = false;
    struct Rec {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    double total_ms[PROF_N] = {};
    long long count[PROF_N] = {};
    hipEvent_t get()
    {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        // timing only: no system-scope fence, so recording does not write
        // back and invalidate the caches around the bracketed launches
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
        return e;
    }
};

thread_local Prof *tl_prof = nullptr;

// Brackets a launch group with events when profiling is on.
struct ProfScope {
    int kind;
    hipStream_t st;
    hipEvent_t a = nullptr;
    ProfScope(int k, hipStream_t s, bool when = true) : kind(k), st(s)
    {
        if (when && tl_prof && tl_prof->on) {
            a = tl_prof->get();
            if (a && hipEventRecord(a, st) != hipSuccess) a = nullptr;
        }
    }
    ~ProfScope()
    {
        if (!a) return;
        hipEvent_t b = tl_prof->get();
        if (b && hipEventRecord(b, st) == hipSuccess) tl_prof->pending.push_back({kind, a, b});
    }
};

}  // namespace trec

// ------------------------------------------------------------------ context

struct trec_context {
    int device = 0;
    std::string err;
    char *arena = nullptr;
    size_t arena_size = 0;
    char *staging = nullptr;  // pinned host memory for descriptor uploads
    size_t staging_size = 0;
    hipEvent_t staging_ev = nullptr;
    // The arena is reused by every call.  A call's device work waits for the
    // previous call's (arena_ev, recorded on that call's stream), so calls on
    // different streams do not overwrite descriptors or workspace that
    // earlier kernels still read.
    hipEvent_t arena_ev = nullptr;
    hipStream_t arena_stream = nullptr;  // stream of the call holding the arena
    bool arena_held = false;
    hipStream_t cap = nullptr;  // capture stream
    size_t ws_limit = 0;        // workspace budget of one call (0: the free device memory)
    hipGraphExec_t gexec = nullptr;
    std::string gkey;
    std::map<std::string, Tuned> tuned;  // per geometry and storage
    trec::Prof prof;
};

namespace trec {

void clear_error(trec_context *ctx)
{
    if (ctx) ctx->err.clear();
}

int context_device(const trec_context *ctx) { return ctx->device; }

trec_status fail(trec_context *ctx, trec_status s, const char *fmt, ...)
{
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return s;
}

}  // namespace trec

#define HIPCHK(ctx, x)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess)                                                               \
            return trec::fail((ctx), e_ == hipErrorOutOfMemory ? TREC_ERR_NOMEM : TREC_ERR_DEVICE, \
                              "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

// ------------------------------------------------------------------ geometry

namespace trec {

// TranslateAngleToCoefficients + DecorateCoefficients, main-mpi.cpp:85-183,
// linear terms: float trig of the float angle, double everything after.
void coefficients(const trec_geom &g, const float *angles, double *coef)
{
    const double alpha = -(g.pitch * PI_180);
    const double beta = g.offset * PI_180;
    const double t = -g.zshift;
    const double ca = cos(alpha), sa = sin(alpha), cb = cos(beta), sb = sin(beta);
    for (int i = 0; i < g.nproj; i++) {
        const float r = angles[i] * PI_180;
        const double a[4] = {0.0, (double)cosf(r), 0.0, (double)(-sinf(r))};
        const double b[4] = {0.0, 0.0, 1.0, 0.0};
        double na[4], nb[4];
        na[0] = a[0];
        na[1] = (a[2] * sa * sb + a[3] * ca * sb + a[1] * cb);
        na[2] = (a[2] * ca - a[3] * sa);
        na[3] = (a[3] * ca * cb + a[2] * cb * sa - a[1] * sb);
        nb[0] = b[0];
        nb[1] = (b[2] * sa * sb + b[3] * ca * sb + b[1] * cb);
        nb[2] = (b[2] * ca - b[3] * sa);
        nb[3] = (b[3] * ca * cb + b[2] * cb * sa - b[1] * sb);
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

// Origins of the reference binaries: x = X/2, z = T/2; y = T/2 in TiltRec-mpi
// (main-mpi.cpp:743-745, mrcvol.Y() is the thickness there) and Y/2 in the
// CUDA and Z binaries (main-cuda.cpp:259-261).
void origins(const trec_geom &g, double &ox, double &oy, double &oz)
{
    ox = (double)(float)(g.nx * .5);
    oz = (double)(float)(g.thickness * .5);
    oy = g.rules == TREC_RULES_MPI ? (double)(float)(g.thickness * .5) : (double)(float)(g.ny * .5);
}

// R-weighting as a circular convolution kernel of length N = X + X/10
// (filter_prj.cpp:674-736): hk[d] = (1/X) sum_k W_k exp(2 pi i d k / N) with
// W_0 = 0.2, W_k = k for k <= K = floor(0.05 N), K exp(-(k-K)^2 / (N/2)^2)
// beyond, mirrored for the negative frequencies.
std::vector<float> rweight_kernel(int X, int &N)
{
    N = X + X / 10;
    const int K = (int)floorf((float)N * 0.05f);
    const double sig = (double)((float)N * 0.5f);
    const int nb = N / 2 + 1;
    std::vector<double> W(nb);
    for (int k = 0; k < nb; ++k)
        W[k] = k == 0 ? 0.2 : (k <= K ? (double)k : (double)K * exp(-((double)(k - K) * (k - K)) / (sig * sig)));
    std::vector<double> ct(N);
    for (int m = 0; m < N; ++m) ct[m] = cos(2.0 * M_PI * m / N);
    std::vector<float> hk(N);
    for (int d = 0; d < N; ++d) {
        double acc = W[0];
        for (int k = 1; k <= (N - 1) / 2; ++k) acc += 2.0 * W[k] * ct[(int)(((long long)d * k) % N)];
        if (N % 2 == 0) acc += W[N / 2] * ((d % 2) ? -1.0 : 1.0);
        hk[d] = (float)(acc / X);
    }
    return hk;
}

// The FBP spatial filter's taps (SheppLogan / RamLak, filter_prj.cpp:137-178,
// in the same float / double operations) and N = cvGetOptimalDFTSize(X + 2 +
// taps - 1), the smallest 2^a 3^b 5^c not below it (filter_prj.cpp:277-278).
std::vector<float> spatial_taps(int mode, int X, int &N)
{
    const int L = FBP_TAPS;
    std::vector<float> h(L);
    if (mode == 0) {
        float n = -((float)L - 1.0) / 2.0;
        for (int i = 0; i < L; ++i) {
            h[i] = (float)(-2.0 / (M_PI * (4.0 * n * n - 1.0)));
            n = n + 1.0;
        }
    } else {
        int n = -(L - 1) / 2;
        for (int i = 0; i < L; ++i, ++n)
            h[i] = n == 0 ? (float)(M_PI / 4) : (std::abs(n % 2) == 1 ? (float)(-1.0 / (M_PI * n * n)) : 0.f);
    }
    for (N = X + 2 + L - 1;; ++N) {
        int r = N;
        for (int f : {2, 3, 5})
            while (r % f == 0) r /= f;
        if (r == 1) break;
    }
    return h;
}

}  // namespace trec

// ------------------------------------------------------------------ plans

namespace trec {

bool is_iterative(int k) { return k == TREC_SIRT || k == TREC_SART; }
// Methods that forward-project: they need the residual and ray weights.
bool uses_forward(int k) { return is_iterative(k) || k == TREC_ADMM; }
// FBP and WBP filter the projections, then back-project like BPT.
bool is_filtered(int k) { return k == TREC_FBP || k == TREC_WBP; }

// Device taps of a filtered method: the R-weighting's circulant kernel of
// length N (WBP), or the 251 Shepp-Logan taps with N the length of the
// reference's padded DFT (FBP).
static void filter_taps(JobPlan &p, int k, int X)
{
    if (k == TREC_WBP) p.hk_host = rweight_kernel(X, p.N);
    else if (k == TREC_FBP) p.hk_host = spatial_taps(0, X, p.N);
}

trec_status validate(trec_context *ctx, const trec_geom *g, const trec_method *m, const float *angles)
{
    if (!g || !m || !angles) return fail(ctx, TREC_ERR_INVALID, "null geometry, method or angles");
    if (g->nx < 2 || g->ny < 1 || g->nproj < 1 || g->thickness < 1)
        return fail(ctx, TREC_ERR_INVALID, "bad sizes nx=%d ny=%d nproj=%d thickness=%d", g->nx, g->ny, g->nproj,
                    g->thickness);
    if (g->nproj > 32767 || g->nx > (1 << 20) || g->ny > (1 << 20) || g->thickness > (1 << 20))
        return fail(ctx, TREC_ERR_INVALID, "sizes out of range");
    if (!std::isfinite(g->offset) || !std::isfinite(g->pitch) || !std::isfinite(g->zshift))
        return fail(ctx, TREC_ERR_INVALID, "non-finite geometry");
    if (g->rules != TREC_RULES_MPI && g->rules != TREC_RULES_CUDA) return fail(ctx, TREC_ERR_INVALID, "bad rules");
    if (m->method != TREC_BPT && m->method != TREC_FBP && m->method != TREC_WBP && m->method != TREC_SIRT &&
        m->method != TREC_SART && m->method != TREC_ADMM)
        return fail(ctx, TREC_ERR_INVALID, "unknown method %d", m->method);
    if ((is_iterative(m->method) || m->method == TREC_ADMM) && (m->iterations < 0 || m->iterations > 100000))
        return fail(ctx, TREC_ERR_INVALID, "bad iteration count %d", m->iterations);
    if (m->method == TREC_ADMM &&
        (m->cg_iterations < 0 || m->cg_iterations > 100000 || !std::isfinite(m->soft) || m->soft < 0))
        return fail(ctx, TREC_ERR_INVALID, "bad ADMM parameters cgiter=%d soft=%g", m->cg_iterations, m->soft);
    if (!std::isfinite(m->gamma)) return fail(ctx, TREC_ERR_INVALID, "non-finite gamma");
    if (m->storage < TREC_STORE_F32 || m->storage > TREC_STORE_BF16) return fail(ctx, TREC_ERR_INVALID, "bad storage");
    if (m->layout != TREC_LAYOUT_Y && m->layout != TREC_LAYOUT_Z) return fail(ctx, TREC_ERR_INVALID, "bad layout");
    for (int i = 0; i < g->nproj; ++i)
        if (!std::isfinite(angles[i])) return fail(ctx, TREC_ERR_INVALID, "non-finite tilt angle %d", i);
    return TREC_OK;
}

// The tuning record for a geometry (the context's, else the process's), or
// an empty one.
Tuned find_tuned(trec_context *ctx, const trec_geom &g, int storage)
{
    const std::string key = tune_key(g, storage);
    if (ctx) {
        auto it = ctx->tuned.find(key);
        if (it != ctx->tuned.end()) return it->second;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(key);
    return it != g_tuned.end() ? it->second : Tuned{};
}

int pick_nzc(trec_context *ctx, const trec_geom &g, int storage)
{
    if (const char *e = getenv("TREC_NZC")) {
        int v = atoi(e);
        if (v >= 1) return std::min(v, g.thickness);
    }
    const Tuned t = find_tuned(ctx, g, storage);
    if (t.nzc > 0) return std::min(t.nzc, g.thickness);
    // One z chunk (the residual written by the forward projection itself, no
    // combine pass) unless the grid would then have fewer than about eight
    // work-groups per CU; more chunks split the layers between work-groups.
    JobDesc h{};
    h.X = g.nx;
    h.nyb = (g.ny + RB - 1) / RB;
    h.nzc = 1;
    const unsigned wg1 = fwd_value_blocks(h, g.nproj);
    int nzc = 1;
    while (nzc < g.thickness && (unsigned long long)wg1 * nzc < 2048ull) nzc *= 2;
    return std::max(1, std::min(nzc, g.thickness));
}

// The widest union of voxel bands that k_forward_g stages for a tilt group
// (union_block there): over layers z, the group's tilts k and its pixel
// block at i0, from min_k floor(s_k(i0 - 1, z)) + 1 to max_k floor(s_k(i0 +
// FW_PIX - 2, z)) + 1 + ncand_k, s_k(i, z) = (i - c0_k) / a1_k + slope_k z.
// That is a maximum of affine functions of (z, i0) minus a minimum of them,
// convex, so it peaks at a corner of the ranges; floor() moves it by one.
int widest_union(const std::vector<TiltConst> &tc, const trec_geom &g)
{
    const int P = g.nproj, npb = (g.nx + FW_PIX - 1) / FW_PIX;
    int widest = 0;
    for (int th0 = 0; th0 < P; th0 += FG) {
        for (int i0 : {0, (npb - 1) * FW_PIX}) {
            for (int z : {0, g.thickness - 1}) {
                int lo = INT_MAX, hi = INT_MIN;
                for (int k = 0; k < FG; ++k) {
                    const TiltConst &t = tc[std::min(th0 + k, P - 1)];
                    const double slo = ((double)(i0 - 1) - t.c0) * t.inv_a1;
                    const double shi = ((double)(i0 + FW_PIX - 2) - t.c0) * t.inv_a1;
                    lo = std::min(lo, (int)std::floor(t.slope * z + slo) + 1);
                    hi = std::max(hi, (int)std::floor(t.slope * z + shi) + 1 + t.ncand);
                }
                widest = std::max(widest, hi - lo);
            }
        }
    }
    return widest;
}

// Fills the host side of a plan (no device memory yet).
trec_status plan_job(trec_context *ctx, JobPlan &p, const trec_geom &g, const trec_method &m, const float *angles,
                     const float *stack, const float *initial, float *volume, int force_nzc, bool force_general)
{
    p.g = g;
    p.m = m;
    p.stack = stack;
    p.initial = initial;
    p.volume = volume;
    if (m.method == TREC_BPT || is_filtered(m.method) || m.method == TREC_ADMM) p.m.storage = TREC_STORE_F32;
    p.coef.resize(8 * (size_t)g.nproj);
    coefficients(g, angles, p.coef.data());
    double ox, oy, oz;
    origins(g, ox, oy, oz);

    // Row-blocked path: pitch 0 (b = (0,0,1,0), a2 = 0 exactly) and every
    // tilt's a1 >= 0.2 (|angle + offset| <= ~78 degrees).
    bool rows = !force_general && !getenv("TREC_FORCE_GENERAL");
    for (int i = 0; i < g.nproj && rows; ++i) {
        const double *c = &p.coef[8 * (size_t)i];
        if (!(c[2] == 0.0 && c[4] == 0.0 && c[5] == 0.0 && c[6] == 1.0 && c[7] == 0.0 && c[1] >= 0.2)) rows = false;
    }
    p.path = rows ? PATH_ROWS : PATH_GENERAL;

    // numeric rules (see trec_rules).  The reference compares float sums with
    // double literals (w > 1e-7); the kernels compare floats with floats, so
    // each literal becomes the largest float not above it: for a float w,
    // w > that float exactly when w > the literal.
    auto at_most = [](double t) {
        float f = (float)t;
        if ((double)f > t) f = std::nextafter(f, 0.0f);
        return f;
    };
    float thr_upd = 0.f, thr_res = 0.f;
    if (g.rules == TREC_RULES_CUDA) {
        if (m.method == TREC_SIRT) thr_upd = at_most(10e-6);                           // basic.cu:184
        if (m.method == TREC_BPT || is_filtered(m.method)) thr_upd = at_most(1e-5);   // basic.cu:409
        if (m.method == TREC_SART) {
            thr_upd = at_most(10e-7);  // basic.cu:434
            thr_res = at_most(1e-7);   // basic.cu:170
        }
    }
    // ADMM: a tap contributes where the ray weight exceeds 10e-6 (both
    // reference binaries, main-mpiz.cpp:740-785, baiscz.cu:446-480)
    if (m.method == TREC_ADMM) thr_res = at_most(10e-6);
    const int iters = is_iterative(m.method) ? m.iterations : 0;
    if (m.method == TREC_SIRT) p.stages = iters;
    else if (m.method == TREC_SART) p.stages = iters * g.nproj;
    else if (m.method == TREC_ADMM) p.stages = 1;  // the whole CG recursion is one stage (issue_admm)
    else p.stages = 1;

    if (p.path == PATH_ROWS) {
        JobDesc &h = p.h;
        memset(&h, 0, sizeof h);
        h.X = g.nx;
        h.Y = g.ny;
        h.T = g.thickness;
        h.P = g.nproj;
        h.nyb = (g.ny + RB - 1) / RB;
        p.nyb_total = h.nyb;
        p.slab_rb = h.nyb;
        p.nslab = 1;
        h.nzc = 1;
        if (m.method == TREC_SIRT) {
            h.nzc = force_nzc > 0 ? std::min(force_nzc, g.thickness) : pick_nzc(ctx, g, p.m.storage);
            const Tuned t = find_tuned(ctx, g, p.m.storage);
            h.bp_g = t.bp_g;
            h.bp_zt = t.bp_zt;
            // TREC_BACK_SHAPE=g,zt forces one of back_shapes (tests and tuning runs)
            if (const char *e = getenv("TREC_BACK_SHAPE")) {
                int bg = 0, bz = 0;
                if (sscanf(e, "%d,%d", &bg, &bz) == 2)
                    for (const auto &sh : back_shapes(p.m.storage))
                        if (sh.first == bg && sh.second == bz) {
                            h.bp_g = bg;
                            h.bp_zt = bz;
                        }
            }
        }
        h.zc_len = (g.thickness + h.nzc - 1) / h.nzc;
        h.nzc = (g.thickness + h.zc_len - 1) / h.zc_len;
        h.storage = p.m.storage;
        h.gamma = m.gamma;
        h.thr_upd = thr_upd;
        h.thr_res = thr_res;
        h.scale = 1.0f;
        p.tc.resize(g.nproj);
        for (int i = 0; i < g.nproj; ++i) {
            const double *c = &p.coef[8 * (size_t)i];
            TiltConst &t = p.tc[i];
            t.a0 = c[0];
            t.a1 = c[1];
            t.a3 = c[3];
            t.ox = ox;
            t.oz = oz;
            t.c0 = c[0] + ox - c[1] * ox - c[3] * oz;
            t.inv_a1 = 1.0 / c[1];
            t.slope = -c[3] / c[1];
            t.a1f = (float)c[1];
            // candidates a pixel needs per layer: ceil(2 / a1); the 1e-6 keeps a
            // float-rounded 60 degrees (a1 = 0.49999997) at 4, dropping a
            // candidate whose weight is below 1.2e-7
            t.ncand = (int)std::ceil(2.0 / c[1] - 1e-6);
        }
        p.fg_small = widest_union(p.tc, g) + FW_G + 2 <= FGS_SMALL;
        filter_taps(p, m.method, g.nx);
    } else {
        GeneralJob &q = p.gj;
        q.X = g.nx;
        q.Y = g.ny;
        q.T = g.thickness;
        q.P = g.nproj;
        q.gamma = m.gamma;
        q.thr_upd = thr_upd;
        q.thr_res = thr_res;
        q.layout = m.layout;
        q.stack = stack;
        q.vol = volume;
        p.tg.resize(g.nproj);
        for (int i = 0; i < g.nproj; ++i) {
            const double *c = &p.coef[8 * (size_t)i];
            TiltGeneral &t = p.tg[i];
            for (int k = 0; k < 4; ++k) {
                t.a[k] = c[k];
                t.b[k] = c[4 + k];
            }
            t.ox = ox;
            t.oy = oy;
            t.oz = oz;
        }
        // SIRT's splat forward (a pitch with an offset the row windows do not
        // take) sums chunks of about 32 layers into partials, so that the
        // waves in flight on one XCD share their volume rows in its L2; the
        // partials take at most 4 GB.  The row windows need none, and the
        // forward takes them where rows3b_plan says so (as its launcher does).
        q.nzc_all = 1;
        bool flat = true;
        for (const TiltGeneral &t : p.tg) flat = flat && t.b[1] == 0.0;
        const char *force = getenv("TREC_GENERAL_FORWARD");
        const bool forced = force && (!strcmp(force, "splat") || !strcmp(force, "rows") ||
                                      !strcmp(force, "rows5") || !strcmp(force, "gather"));
        Rows3bPlan rp;
        const bool splat_runs = forced ? !strcmp(force, "splat")
                                       : !rows3b_plan(p.tg.data(), 0, g.nproj, g.nx, g.thickness, m.layout, rp);
        if (general_tiles(p.tg.data(), g.nproj, g.nx) && !flat && m.method == TREC_SIRT && splat_runs) {
            const size_t one = 4 * (size_t)g.nproj * g.ny * g.nx;
            q.nzc_all = (int)std::min<size_t>({8, (size_t)(g.thickness + 31) / 32, ((size_t)4 << 30) / one});
            if (q.nzc_all < 2) q.nzc_all = 1;
        }
        filter_taps(p, m.method, g.nx);
    }
    return TREC_OK;
}

// Assigns device pointers from `base` (nullptr: sizing only); returns bytes.
size_t carve(JobPlan &p, char *base)
{
    size_t off = 0;
    auto take = [&](size_t bytes) -> char * {
        char *r = base ? base + off : nullptr;
        off += align_up(bytes);
        return r;
    };
    const size_t X = p.g.nx, Y = p.g.ny, T = p.g.thickness, P = p.g.nproj;
    const int k = p.m.method;
    if (p.path == PATH_ROWS) {
        JobDesc &h = p.h;
        const size_t es = p.m.storage == TREC_STORE_F32 ? 4 : 2;
        const size_t nyb = h.nyb;
        p.d_all = (JobDesc *)take(sizeof(JobDesc) * p.nslab);
        p.d = p.d_all;
        p.d_ones = (JobDesc *)take(sizeof(JobDesc));
        h.tc = (TiltConst *)take(sizeof(TiltConst) * P);
        // the forward projection's band loads may run up to FGS cells
        // (16 bytes each) past a slice's last plane; the tail keeps even the
        // last slice's overrun inside the workspace
        h.vol = (float *)take(4 * nyb * T * X * RB + 16 * FGS);
        // half storage: the forward projection reads a half copy of the
        // volume (same tail for the band loads' overrun)
        h.hvol = (p.m.storage == TREC_STORE_F16 && is_iterative(k))
                     ? take(64 * nyb * T * (X + 1) + 16 * (FGS + FW_G))
                     : nullptr;
        h.proj = take(es * nyb * P * X * RB);
        h.res = uses_forward(k) ? take(es * nyb * P * X * RB) : nullptr;
        h.part = (k == TREC_SIRT && h.nzc > 1) ? (float *)take(4 * (size_t)h.nzc * nyb * P * X * RB) : nullptr;
        h.inv_rwt = uses_forward(k) ? (float *)take(4 * P * X) : nullptr;
        h.inv_c = k == TREC_SIRT ? (float *)take(4 * T * X) : nullptr;
        h.rwt = p.want_rwt ? (float *)take(4 * P * X) : nullptr;
        p.sumsq = p.m.storage != TREC_STORE_F32 ? (double *)take(sizeof(double) * SUMSQ_BLOCKS) : nullptr;
        p.hk = is_filtered(k) ? (float *)take(4 * p.hk_host.size()) : nullptr;
        if (k == TREC_ADMM) {
            // the CG and ADMM vectors, each a row-blocked volume, and the
            // descriptors of the three back-projection passes
            AdmmVecs &a = p.av;
            a.n = nyb * T * X * RB;
            float **vecs[7] = {&a.htb, &a.u, &a.d, &a.r, &a.p, &a.h, &a.ax};
            for (float **v : vecs) *v = (float *)take(4 * a.n + 16 * FGS);
            a.part = (double *)take(sizeof(double) * 2 * ADMM_DOT_BLOCKS);
            a.S = (double *)take(sizeof(double) * 8);
            for (int i = 0; i < 3; ++i) p.d_adm[i] = (JobDesc *)take(sizeof(JobDesc));
        }
    } else {
        GeneralJob &q = p.gj;
        p.d_tg = (TiltGeneral *)take(sizeof(TiltGeneral) * P);
        q.tg = p.d_tg;
        q.rval = uses_forward(k) || p.want_rwt ? (float *)take(4 * P * Y * X) : nullptr;
        q.rwt = uses_forward(k) || p.want_rwt ? (float *)take(4 * P * Y * X) : nullptr;
        p.stack_copy = is_filtered(k) ? (float *)take(4 * P * Y * X) : nullptr;
        const size_t npart = std::max(k == TREC_SART ? (size_t)GZ_CHUNKS * Y * X : 0,
                                      q.nzc_all > 1 ? (size_t)q.nzc_all * P * Y * X : 0);
        q.part = npart ? (float *)take(4 * npart) : nullptr;
        if (k == TREC_ADMM) {
            // the CG and ADMM vectors, each a volume in the job's layout
            AdmmVecs &a = p.av;
            a.n = X * Y * T;
            float **vecs[7] = {&a.htb, &a.u, &a.d, &a.r, &a.p, &a.h, &a.ax};
            for (float **v : vecs) *v = (float *)take(4 * a.n);
            a.part = (double *)take(sizeof(double) * 2 * ADMM_DOT_BLOCKS);
            a.S = (double *)take(sizeof(double) * 8);
        }
        p.hk = is_filtered(k) ? (float *)take(4 * p.hk_host.size()) : nullptr;
    }
    p.bytes = off;
    return off;
}

trec_status wait_arena(trec_context *ctx, hipStream_t st)
{
    HIPCHK(ctx, hipStreamWaitEvent(st, ctx->arena_ev, 0));
    ctx->arena_stream = st;
    ctx->arena_held = true;
    return TREC_OK;
}

// Makes the arena at least `bytes` long and orders `st` after the previous
// call's use of it.
trec_status ensure_arena(trec_context *ctx, size_t bytes, hipStream_t st)
{
    if (ctx->arena_size >= bytes) return wait_arena(ctx, st);
    HIPCHK(ctx, hipDeviceSynchronize());
    if (ctx->gexec) {
        (void)hipGraphExecDestroy(ctx->gexec);
        ctx->gexec = nullptr;
        ctx->gkey.clear();
    }
    if (ctx->arena) (void)hipFree(ctx->arena);
    ctx->arena = nullptr;
    ctx->arena_size = 0;
    const size_t want = align_up(bytes + bytes / 16);
    hipError_t e = hipMalloc((void **)&ctx->arena, want);
    if (e != hipSuccess) {
        ctx->arena = nullptr;
        (void)hipGetLastError();
        return fail(ctx, TREC_ERR_NOMEM, "device workspace of %zu bytes: %s", want, hipGetErrorString(e));
    }
    ctx->arena_size = want;
    return wait_arena(ctx, st);
}

// Records the end of a call's use of the arena on its stream (arena_ev).
struct ArenaGuard {
    trec_context *ctx;
    explicit ArenaGuard(trec_context *c) : ctx(c) {}
    ~ArenaGuard()
    {
        if (ctx && ctx->arena_held) {
            (void)hipEventRecord(ctx->arena_ev, ctx->arena_stream);
            ctx->arena_held = false;
        }
    }
};

// Copies the host-side descriptors of all plans to the device in one pass
// through the pinned staging buffer.
trec_status upload(trec_context *ctx, const std::vector<JobPlan *> &plans, hipStream_t st)
{
    struct Piece {
        void *dst;
        const void *src;
        size_t n;
    };
    std::vector<Piece> pcs;
    std::vector<JobDesc> ones(plans.size());
    std::vector<std::vector<JobDesc>> slabs(plans.size());
    for (size_t i = 0; i < plans.size(); ++i) {
        JobPlan &p = *plans[i];
        if (p.path == PATH_ROWS) {
            ones[i] = p.h;
            ones[i].y0 = 0;
            ones[i].nyb = 1;
            ones[i].nzc = 1;
            ones[i].zc_len = p.h.T;
            for (int sl = 0; sl < p.nslab; ++sl) {
                JobDesc hs = p.h;
                hs.y0 = sl * p.slab_rb * RB;
                hs.nyb = std::min(p.slab_rb, p.nyb_total - sl * p.slab_rb);
                slabs[i].push_back(hs);
            }
            pcs.push_back({p.d_all, slabs[i].data(), sizeof(JobDesc) * slabs[i].size()});
            pcs.push_back({p.d_ones, &ones[i], sizeof(JobDesc)});
            if (p.m.method == TREC_ADMM) {
                // 0: htb = A^T b (the stack as the image, no x term);
                // 1: ax = op(vol); 2: h = op(p).  The operator's forward
                // projection reads the stack after it has been zeroed, so its
                // residual is -R / w; out = -sum s + mu x.
                for (int k = 0; k < 3; ++k) p.h_adm[k] = p.h;
                p.h_adm[0].res = p.h.proj;
                p.h_adm[0].out = p.av.htb;
                p.h_adm[0].lin_a = 1.f;
                p.h_adm[0].lin_b = 0.f;
                p.h_adm[1].out = p.av.ax;
                p.h_adm[2].vol = p.av.p;
                p.h_adm[2].out = p.av.h;
                for (int k = 1; k < 3; ++k) {
                    p.h_adm[k].lin_a = -1.f;
                    p.h_adm[k].lin_b = p.m.gamma;
                }
                for (int k = 0; k < 3; ++k) pcs.push_back({p.d_adm[k], &p.h_adm[k], sizeof(JobDesc)});
            }
            pcs.push_back({const_cast<TiltConst *>(p.h.tc), p.tc.data(), sizeof(TiltConst) * p.tc.size()});
        } else {
            pcs.push_back({p.d_tg, p.tg.data(), sizeof(TiltGeneral) * p.tg.size()});
        }
        if (p.hk) pcs.push_back({p.hk, p.hk_host.data(), sizeof(float) * p.hk_host.size()});
    }
    size_t total = 0;
    for (auto &c : pcs) total += align_up(c.n);
    HIPCHK(ctx, hipEventSynchronize(ctx->staging_ev));
    if (ctx->staging_size < total) {
        if (ctx->staging) (void)hipHostFree(ctx->staging);
        ctx->staging = nullptr;
        ctx->staging_size = 0;
        HIPCHK(ctx, hipHostMalloc((void **)&ctx->staging, total, hipHostMallocDefault));
        ctx->staging_size = total;
    }
    size_t off = 0;
    for (auto &c : pcs) {
        memcpy(ctx->staging + off, c.src, c.n);
        HIPCHK(ctx, hipMemcpyAsync(c.dst, ctx->staging + off, c.n, hipMemcpyHostToDevice, st));
        off += align_up(c.n);
    }
    HIPCHK(ctx, hipEventRecord(ctx->staging_ev, st));
    return TREC_OK;
}

bool split_slabs(JobPlan &p, size_t limit)
{
    if (p.path != PATH_ROWS || p.m.method == TREC_ADMM) return carve(p, nullptr) <= limit;
    auto bytes_for = [&](int rb) {
        p.h.nyb = rb;
        p.slab_rb = rb;
        p.nslab = (p.nyb_total + rb - 1) / rb;
        return carve(p, nullptr);
    };
    int lo = 1, hi = p.nyb_total;
    if (bytes_for(hi) <= limit) return true;
    if (bytes_for(lo) > limit) {
        bytes_for(p.nyb_total);
        return false;
    }
    while (lo < hi) {  // largest slab whose workspace fits
        const int mid = lo + (hi - lo + 1) / 2;
        if (bytes_for(mid) <= limit) lo = mid;
        else hi = mid - 1;
    }
    bytes_for(lo);
    return true;
}

void select_slab(JobPlan &p, int s)
{
    p.h.y0 = s * p.slab_rb * RB;
    p.h.nyb = std::min(p.slab_rb, p.nyb_total - s * p.slab_rb);
    p.d = p.d_all + s;
}

// ------------------------------------------------------------------ issue

#define LCHK(x)                                  \
    do {                                         \
        hipError_t e_ = (x);                     \
        if (e_ != hipSuccess) return e_;         \
    } while (0)

// Work-groups of a value forward projection (the kernel launch_forward picks).
unsigned fwd_blocks(const JobDesc &h, int ntheta)
{
    return fwd_value_blocks(h, ntheta);
}

hipError_t issue_prep(JobPlan &p, hipStream_t st)
{
    ProfScope ps(PROF_PREP, st);
    const size_t X = p.g.nx, Y = p.g.ny, T = p.g.thickness, P = p.g.nproj;
    const int k = p.m.method;
    if (p.path == PATH_ROWS) {
        JobDesc &h = p.h;
        if (p.sumsq) {
            LCHK(launch_sumsq(p.stack, TREC_STORE_F32, P * Y * X, p.sumsq, st));
            LCHK(launch_set_scale(p.d, p.sumsq, P * Y * X, st));
        }
        LCHK(launch_block_stack(h, p.d, p.stack, st));
        if (p.initial) LCHK(launch_block_volume(h, p.d, p.initial, p.m.layout, st));
        else LCHK(launch_zero(h.vol, (size_t)h.nyb * T * X * RB, st));
        if (h.hvol) LCHK(launch_shadow(h, p.d, st));
        // rows of the float slices: 4-row planes of X cells, row stride 1, cell stride 4
        if (k == TREC_WBP)
            LCHK(launch_rweight((float *)h.proj, p.hk, p.N, (int)X, (long long)h.nyb * P * RB, 1, 4, 4,
                                (long long)X * 4, st));
        if (k == TREC_FBP)
            LCHK(launch_fbp_filter((float *)h.proj, p.hk, p.N, (int)X, (long long)h.nyb * P * RB, 1, 4, 4,
                                   (long long)X * 4, st));
        if (uses_forward(k)) {
            LaunchList L{};
            L.n = 1;
            L.off[0] = 0;
            JobDesc o = h;
            o.nyb = 1;
            o.nzc = 1;
            o.zc_len = (int)T;
            L.off[1] = forward_blocks(o, (int)P);
            L.job[0] = p.d_ones;
            L.theta0[0] = 0;
            L.ntheta[0] = 0;
            LCHK(launch_forward(L, L.off[1], -1, false, st));
        }
        if (k == TREC_SIRT) LCHK(launch_weights(h, p.d, st));
    } else {
        GeneralJob &q = p.gj;
        const size_t nv = X * Y * T;
        if (p.initial && p.initial != p.volume) LCHK(launch_copy(p.volume, p.initial, nv, st));
        else if (!p.initial) LCHK(launch_zero(p.volume, nv, st));
        if (is_filtered(k)) {
            LCHK(launch_copy(p.stack_copy, p.stack, P * Y * X, st));
            if (k == TREC_WBP)
                LCHK(launch_rweight(p.stack_copy, p.hk, p.N, (int)X, (long long)P * Y, 0, 1, 1, (long long)X, st));
            else
                LCHK(launch_fbp_filter(p.stack_copy, p.hk, p.N, (int)X, (long long)P * Y, 0, 1, 1, (long long)X,
                                       st));
        }
        if (uses_forward(k)) {
            LCHK(launch_zero(q.rwt, P * Y * X, st));
            LCHK(launch_general_forward(q, p.tg.data(), 0, (int)P, true, st));
        }
    }
    return hipSuccess;
}

hipError_t issue_finish(JobPlan &p, hipStream_t st)
{
    ProfScope ps(PROF_FINISH, st);
    if (p.path == PATH_ROWS) return launch_unblock_volume(p.h, p.d, p.volume, p.m.layout, st);
    return hipSuccess;
}

hipError_t issue_admm(JobPlan &p, hipStream_t st);
hipError_t issue_admm_general(JobPlan &p, hipStream_t st);

// One stage of a general-path job (issued on its own).
hipError_t issue_general_stage(JobPlan &p, int s, hipStream_t st)
{
    GeneralJob &q = p.gj;
    const size_t X = p.g.nx, Y = p.g.ny, P = p.g.nproj;
    switch (p.m.method) {
    case TREC_SIRT: {
        {
            ProfScope ps(PROF_FORWARD, st);
            LCHK(launch_zero(q.rval, P * Y * X, st));
            int nzc = 1;
            LCHK(launch_general_forward(q, p.tg.data(), 0, (int)P, false, st, &nzc));
            LCHK(launch_general_residual(q, 0, (int)P, false, st, nzc));
        }
        ProfScope ps(PROF_BACK, st);
        return launch_general_back(q, p.tg.data(), BACK_SIRT, 0, (int)P, q.rval, st);
    }
    case TREC_SART: {
        const int th = s % (int)P;
        {
            ProfScope ps(PROF_FORWARD, st);
            LCHK(launch_zero(q.rval + (size_t)th * Y * X, Y * X, st));
            int nzc = 1;
            LCHK(launch_general_forward(q, p.tg.data(), th, 1, false, st, &nzc));
            LCHK(launch_general_residual(q, th, 1, false, st, nzc));
        }
        ProfScope ps(PROF_BACK, st);
        return launch_general_back(q, p.tg.data(), BACK_SART, th, 1, q.rval, st);
    }
    case TREC_ADMM:
        return issue_admm_general(p, st);
    default: {
        ProfScope ps(PROF_BACK, st);
        return launch_general_back(q, p.tg.data(), BACK_BPT, 0, (int)P, is_filtered(p.m.method) ? p.stack_copy : p.stack, st);
    }
    }
}

// Collects jobs into LaunchLists (at most MAX_JOBS each) and launches them.
struct Batcher {
    struct Item {
        const JobDesc *d;
        unsigned blocks;
        short theta0, ntheta;
    };
    std::vector<Item> items;
    void add(const JobDesc *d, unsigned blocks, int theta0, int ntheta)
    {
        if (blocks) items.push_back({d, blocks, (short)theta0, (short)ntheta});
    }
    template <typename F>
    hipError_t flush(F launch)
    {
        size_t i = 0;
        while (i < items.size()) {
            LaunchList L{};
            unsigned total = 0;
            int n = 0;
            while (i < items.size() && n < MAX_JOBS) {
                L.off[n] = total;
                L.job[n] = items[i].d;
                L.theta0[n] = items[i].theta0;
                L.ntheta[n] = items[i].ntheta;
                total += items[i].blocks;
                ++n;
                ++i;
            }
            L.off[n] = total;
            L.n = n;
            LCHK(launch(L, total));
        }
        items.clear();
        return hipSuccess;
    }
};

hipError_t issue_step(const std::vector<JobPlan *> &plans, std::vector<int> &stage, hipStream_t st)
{
    Batcher fp_fused[3], fp_part[3], comb[3], back[3][3], fp_small[3];
    // SIRT back-projections by shape: (storage, tilts per pass, voxels per thread)
    std::map<std::tuple<int, int, int>, Batcher> back_sirt;
    for (size_t j = 0; j < plans.size(); ++j) {
        JobPlan &p = *plans[j];
        const int s = stage[j];
        if (s >= p.stages) continue;
        stage[j] = s + 1;
        if (p.path == PATH_GENERAL) {
            LCHK(issue_general_stage(p, s, st));
            continue;
        }
        const JobDesc &h = p.h;
        const int sto = h.storage;
        switch (p.m.method) {
        case TREC_SIRT:
            if (h.nzc == 1) {
                (p.fg_small ? fp_small[sto] : fp_fused[sto]).add(p.d, fwd_blocks(h, h.P), 0, 0);
            } else {
                fp_part[sto].add(p.d, fwd_blocks(h, h.P), 0, 0);
                comb[sto].add(p.d, combine_blocks(h, h.P), 0, 0);
            }
            back_sirt[std::make_tuple(sto, h.bp_g, h.bp_zt)].add(p.d, back_blocks(h), 0, 0);
            break;
        case TREC_SART: {
            const int th = s % h.P;
            fp_fused[sto].add(p.d, fwd_blocks(h, 1), th, 1);
            back[BACK_SART][sto].add(p.d, back_blocks(h), th, 1);
            break;
        }
        case TREC_ADMM:
            LCHK(issue_admm(p, st));
            break;
        default:
            back[BACK_BPT][sto].add(p.d, back_blocks(h), 0, 0);
            break;
        }
    }
    auto any = [](const Batcher *b, int n) {
        for (int i = 0; i < n; ++i)
            if (!b[i].items.empty()) return true;
        return false;
    };
    {
        ProfScope ps(PROF_FORWARD, st, any(fp_fused, 3) || any(fp_small, 3) || any(fp_part, 3));
        for (int sto = 0; sto < 3; ++sto)
            LCHK(fp_fused[sto].flush(
                [&](const LaunchList &L, unsigned n) { return launch_forward(L, n, sto, true, st); }));
        for (int sto = 0; sto < 3; ++sto)
            LCHK(fp_small[sto].flush(
                [&](const LaunchList &L, unsigned n) { return launch_forward(L, n, sto, true, st, true); }));
        for (int sto = 0; sto < 3; ++sto)
            LCHK(fp_part[sto].flush(
                [&](const LaunchList &L, unsigned n) { return launch_forward(L, n, sto, false, st); }));
    }
    if (!comb[0].items.empty() || !comb[1].items.empty() || !comb[2].items.empty()) {
        ProfScope ps(PROF_COMBINE, st);  // timed only when there is a combine pass
        for (int sto = 0; sto < 3; ++sto)
            LCHK(comb[sto].flush([&](const LaunchList &L, unsigned n) { return launch_combine(L, n, sto, st); }));
    }
    {
        ProfScope ps(PROF_BACK, st, !back_sirt.empty() || any(&back[0][0], 9));
        for (auto &kb : back_sirt) {
            const int sto = std::get<0>(kb.first), g = std::get<1>(kb.first), zt = std::get<2>(kb.first);
            LCHK(kb.second.flush(
                [&](const LaunchList &L, unsigned n) { return launch_back(L, n, sto, BACK_SIRT, st, g, zt); }));
        }
        for (int mode = 0; mode < 3; ++mode)
            for (int sto = 0; sto < 3; ++sto)
                LCHK(back[mode][sto].flush(
                    [&](const LaunchList &L, unsigned n) { return launch_back(L, n, sto, mode, st); }));
    }
    return hipSuccess;
}

// TiltRecZ-mpi's ADMM (main-mpiz.cpp:958-1013) on one row-blocked job, after
// issue_prep: htb = A^T b; then per outer iteration htb += (u - d) mu, the
// CG of applycg (:835-935) on op(v) = htb from the current volume, the soft
// threshold and the d update.
hipError_t issue_admm(JobPlan &p, hipStream_t st)
{
    const JobDesc &h = p.h;
    AdmmVecs a = p.av;
    a.vol = h.vol;
    a.mu = p.m.gamma;
    a.soft = p.m.soft;
    auto one = [&](int k, unsigned blocks) {
        LaunchList L{};
        L.n = 1;
        L.off[0] = 0;
        L.off[1] = blocks;
        L.job[0] = p.d_adm[k];
        L.theta0[0] = 0;
        L.ntheta[0] = 0;
        return L;
    };
    auto back = [&](int k) {
        ProfScope ps(PROF_BACK, st);
        const LaunchList L = one(k, back_blocks(p.h_adm[k]));
        return launch_back(L, L.off[1], TREC_STORE_F32, BACK_LIN, st);
    };
    auto op = [&](int k) {
        {
            ProfScope ps(PROF_FORWARD, st);
            const LaunchList L = one(k, fwd_blocks(p.h_adm[k], h.P));
            LCHK(launch_forward(L, L.off[1], TREC_STORE_F32, true, st));
        }
        return back(k);
    };
    LCHK(back(0));
    LCHK(launch_zero((float *)h.proj, (size_t)h.nyb * h.P * h.X * RB, st));
    LCHK(launch_zero(a.u, a.n, st));
    LCHK(launch_zero(a.d, a.n, st));
    for (int it = 0; it < p.m.iterations; ++it) {
        LCHK(launch_admm_vec(a, ADMM_HTB, st));
        LCHK(op(1));
        LCHK(launch_admm_vec(a, ADMM_R0, st));
        for (int cg = 0; cg < p.m.cg_iterations; ++cg) {
            LCHK(op(2));
            LCHK(launch_admm_dots(a, 0, st));
            LCHK(launch_admm_vec(a, ADMM_STEP, st));
            LCHK(op(1));
            LCHK(launch_admm_vec(a, ADMM_RES, st));
            LCHK(launch_admm_dots(a, 1, st));
            LCHK(launch_admm_vec(a, ADMM_DIR, st));
        }
        LCHK(launch_admm_vec(a, ADMM_SOFT, st));
    }
    return hipSuccess;
}

// The same recursion on a general-path job: the operator is the general
// forward projection of x, the quotient -R / w (0 where w is at most the
// threshold) and the back-projection out = -sum s + mu x.
hipError_t issue_admm_general(JobPlan &p, hipStream_t st)
{
    const GeneralJob &q0 = p.gj;
    const int P = p.g.nproj;
    AdmmVecs a = p.av;
    a.vol = q0.vol;
    a.mu = p.m.gamma;
    a.soft = p.m.soft;
    // htb = A^T b: the stack's back-projection sum, no x term
    {
        ProfScope ps(PROF_BACK, st);
        GeneralJob q = q0;
        LCHK(launch_zero(a.htb, a.n, st));
        q.vol = a.htb;
        q.out = a.htb;
        q.lin_a = 1.f;
        q.lin_b = 0.f;
        LCHK(launch_general_back(q, p.tg.data(), BACK_LIN, 0, P, p.stack, st));
    }
    auto op = [&](float *x, float *out) {
        GeneralJob q = q0;
        q.vol = x;
        q.out = out;
        q.lin_a = -1.f;
        q.lin_b = p.m.gamma;
        {
            ProfScope ps(PROF_FORWARD, st);
            LCHK(launch_zero(q.rval, (size_t)P * q.Y * q.X, st));
            LCHK(launch_general_forward(q, p.tg.data(), 0, P, false, st));
            LCHK(launch_general_residual(q, 0, P, true, st));
        }
        ProfScope ps(PROF_BACK, st);
        return launch_general_back(q, p.tg.data(), BACK_LIN, 0, P, q.rval, st);
    };
    LCHK(launch_zero(a.u, a.n, st));
    LCHK(launch_zero(a.d, a.n, st));
    for (int it = 0; it < p.m.iterations; ++it) {
        LCHK(launch_admm_vec(a, ADMM_HTB, st));
        LCHK(op(a.vol, a.ax));
        LCHK(launch_admm_vec(a, ADMM_R0, st));
        for (int cg = 0; cg < p.m.cg_iterations; ++cg) {
            LCHK(op(a.p, a.h));
            LCHK(launch_admm_dots(a, 0, st));
            LCHK(launch_admm_vec(a, ADMM_STEP, st));
            LCHK(op(a.vol, a.ax));
            LCHK(launch_admm_vec(a, ADMM_RES, st));
            LCHK(launch_admm_dots(a, 1, st));
            LCHK(launch_admm_vec(a, ADMM_DIR, st));
        }
        LCHK(launch_admm_vec(a, ADMM_SOFT, st));
    }
    return hipSuccess;
}

hipError_t issue_all(const std::vector<JobPlan *> &plans, hipStream_t st)
{
    if (plans.size() == 1 && plans[0]->nslab > 1) {
        // one job in y slabs, one after another in the same workspace
        JobPlan &p = *plans[0];
        for (int sl = 0; sl < p.nslab; ++sl) {
            select_slab(p, sl);
            LCHK(issue_prep(p, st));
            std::vector<int> stage(1, 0);
            for (int s = 0; s < p.stages; ++s) LCHK(issue_step(plans, stage, st));
            LCHK(issue_finish(p, st));
        }
        select_slab(p, 0);
        return hipSuccess;
    }
    for (auto *p : plans) LCHK(issue_prep(*p, st));
    std::vector<int> stage(plans.size(), 0);
    int S = 0;
    for (auto *p : plans) S = std::max(S, p->stages);
    for (int s = 0; s < S; ++s) LCHK(issue_step(plans, stage, st));
    for (auto *p : plans) LCHK(issue_finish(*p, st));
    return hipSuccess;
}

// 64-bit FNV-1a of a byte range.
uint64_t fnv1a(const void *data, size_t n, uint64_t h = 1469598103934665603ull)
{
    const unsigned char *b = static_cast<const unsigned char *>(data);
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

// Everything a captured schedule depends on: sizes, method, buffers, and the
// per-tilt geometry, because the host picks kernels and launch shapes from
// the coefficients (the row or general path, the general kernels' row
// windows, candidate and slot counts, the forward's LDS band).
std::string graph_key(trec_context *ctx, const std::vector<JobPlan *> &plans)
{
    std::string k;
    char b[640];
    snprintf(b, sizeof b, "%p|", (void *)ctx->arena);
    k += b;
    // the variables that force a general-path kernel choose at launch, so a
    // graph captured under one setting must not replay under another
    for (const char *v : {"TREC_GENERAL_FORWARD", "TREC_GENERAL_ROWS_LANE", "TREC_GENERAL_BACK"}) {
        const char *e = getenv(v);
        k += e ? e : "-";
        k += '|';
    }
    for (auto *pp : plans) {
        const JobPlan &p = *pp;
        const uint64_t hc = fnv1a(p.coef.data(), sizeof(double) * p.coef.size());
        snprintf(b, sizeof b,
                 "%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%a,%p,%p,%p,%d,%d,%d,%a,%a,%a,%a,%d,%016llx|", p.g.nx, p.g.ny,
                 p.g.nproj, p.g.thickness, p.g.rules, p.m.method, p.m.iterations, p.m.storage, p.m.layout, p.path,
                 (double)p.m.gamma, (const void *)p.stack, (const void *)p.initial, (void *)p.volume,
                 p.path == PATH_ROWS ? p.h.nzc : 0, (int)p.want_rwt, p.m.cg_iterations, (double)p.m.soft,
                 (double)p.g.offset, (double)p.g.pitch, (double)p.g.zshift, (int)p.fg_small,
                 (unsigned long long)hc);
        k += b;
        snprintf(b, sizeof b, "%d,%d|", p.path == PATH_ROWS ? p.h.bp_g : 0, p.path == PATH_ROWS ? p.h.bp_zt : 0);
        k += b;
    }
    return k;
}

// Workspace one call may take: the context's limit, or 90 % of what the
// device has free plus the arena the context already holds.
size_t workspace_budget(trec_context *ctx)
{
    if (ctx->ws_limit) return ctx->ws_limit;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
        (void)hipGetLastError();
        return SIZE_MAX;
    }
    return (fr + ctx->arena_size) / 10 * 9;
}

// Allocates, uploads and issues a list of planned jobs on `st`.  A single
// row-path job whose workspace exceeds the budget runs in y slabs.
trec_status run_jobs(trec_context *ctx, std::vector<JobPlan> &plan_store, hipStream_t st, bool use_graph)
{
    std::vector<JobPlan *> plans;
    for (auto &p : plan_store) plans.push_back(&p);
    if (plans.size() == 1) {
        const size_t budget = workspace_budget(ctx);
        if (carve(*plans[0], nullptr) > budget && !split_slabs(*plans[0], budget))
            return fail(ctx, TREC_ERR_NOMEM, "workspace of one row block exceeds the budget of %zu bytes", budget);
        if (plans[0]->nslab > 1) use_graph = false;
    }
    size_t total = 0;
    for (auto *p : plans) total += carve(*p, nullptr);
    trec_status s = ensure_arena(ctx, total, st);
    if (s != TREC_OK) return s;
    size_t off = 0;
    for (auto *p : plans) {
        carve(*p, ctx->arena + off);
        off += p->bytes;
    }
    s = upload(ctx, plans, st);
    if (s != TREC_OK) return s;
    if (getenv("TREC_NO_GRAPH") || ctx->prof.on) use_graph = false;
    if (!use_graph) {
        tl_prof = &ctx->prof;
        hipError_t e = issue_all(plans, st);
        tl_prof = nullptr;
        HIPCHK(ctx, e);
        return TREC_OK;
    }
    const std::string key = graph_key(ctx, plans);
    if (!ctx->gexec || key != ctx->gkey) {
        if (ctx->gexec) {
            HIPCHK(ctx, hipGraphExecDestroy(ctx->gexec));
            ctx->gexec = nullptr;
            ctx->gkey.clear();
        }
        hipGraph_t graph = nullptr;
        HIPCHK(ctx, hipStreamBeginCapture(ctx->cap, hipStreamCaptureModeThreadLocal));
        hipError_t e = issue_all(plans, ctx->cap);
        hipError_t e2 = hipStreamEndCapture(ctx->cap, &graph);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(ctx, TREC_ERR_DEVICE, "graph capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
        }
        e = hipGraphInstantiate(&ctx->gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            ctx->gexec = nullptr;
            return fail(ctx, TREC_ERR_DEVICE, "graph instantiate: %s", hipGetErrorString(e));
        }
        ctx->gkey = key;
    }
    HIPCHK(ctx, hipGraphLaunch(ctx->gexec, st));
    return TREC_OK;
}

size_t request_bytes(const trec_geom &g, const trec_method &m, const float *angles, int *nzc)
{
    if (nzc) *nzc = 0;
    JobPlan p;
    if (validate(nullptr, &g, &m, angles) != TREC_OK) return 0;
    if (plan_job(nullptr, p, g, m, angles, nullptr, nullptr, nullptr) != TREC_OK) return 0;
    if (nzc && p.path == PATH_ROWS) *nzc = p.h.nzc;
    const size_t X = g.nx, Y = g.ny, T = g.thickness, P = g.nproj;
    return carve(p, nullptr) + align_up(4 * P * Y * X) + align_up(4 * X * Y * T);
}

size_t request_bytes_bound(const trec_geom &g, const trec_method &m)
{
    std::vector<float> ang(std::max(g.nproj, 1), 0.f);
    const size_t rows = request_bytes(g, m, ang.data(), nullptr);
    if (rows == 0) return 0;
    JobPlan p;
    if (plan_job(nullptr, p, g, m, ang.data(), nullptr, nullptr, nullptr, 0, true) != TREC_OK) return 0;
    const size_t X = g.nx, Y = g.ny, T = g.thickness, P = g.nproj;
    return std::max(rows, carve(p, nullptr) + align_up(4 * P * Y * X) + align_up(4 * X * Y * T));
}

}  // namespace trec


// ------------------------------------------------------------------ .xf application

namespace trec {

// OpenCV's 1-D interpolation kernels at fraction x (imgwarp.cpp
// interpolateCubic / interpolateLanczos4), evaluated in the same float and
// double operations on the host, so the tables the device reads are the
// ones warpAffine builds.
static void xf_cubic(float x, float *c)
{
#pragma clang fp contract(off)
    const float A = -0.75f;
    c[0] = ((A * (x + 1) - 5 * A) * (x + 1) + 8 * A) * (x + 1) - 4 * A;
    c[1] = ((A + 2) * x - (A + 3)) * x * x + 1;
    c[2] = ((A + 2) * (1 - x) - (A + 3)) * (1 - x) * (1 - x) + 1;
    c[3] = 1.f - c[0] - c[1] - c[2];
}

// sin(y_k) / y_k^2 for the 8 taps through the angle-sum identity with one
// sin/cos pair, normalised to sum 1; a tap at distance 0 weighs 1e30 before
// normalisation.
static void xf_lanczos4(float x, float *c)
{
#pragma clang fp contract(off)
    static const double s45 = 0.70710678118654752440084436210485;
    static const double cs[8][2] = {{1, 0}, {-s45, -s45}, {0, 1}, {s45, -s45},
                                    {-1, 0}, {s45, s45}, {0, -1}, {-s45, s45}};
    const double pi = 3.1415926535897932384626433832795;
    float sum = 0;
    const double y0 = -(double)(x + 3) * pi * 0.25, s0 = std::sin(y0), c0 = std::cos(y0);
    for (int k = 0; k < 8; ++k) {
        const float d = x + 3 - k;
        if (std::fabs(d) >= 1e-6f) {
            const double y = -(double)d * pi * 0.25;
            c[k] = (float)((cs[k][0] * s0 + cs[k][1] * c0) / (y * y));
        } else {
            c[k] = 1e30f;
        }
        sum += c[k];
    }
    sum = 1.f / sum;
    for (int k = 0; k < 8; ++k) c[k] *= sum;
}

// mrcstack's forward matrix in float (main.cpp:56-60), then warpAffine's
// inversion in double.
static XfConst xf_inverse(const float *xf, int W, int H)
{
#pragma clang fp contract(off)
    const float a11 = xf[0], a12 = xf[1], a21 = xf[2], a22 = xf[3], dx = xf[4], dy = xf[5];
    const float cx = W * .5, cy = H * .5;
    const float f[6] = {a11, a12, cx + dx - a11 * cx - a12 * cy, a21, a22, cy + dy - a21 * cx - a22 * cy};
    XfConst r;
    double *M = r.m;
    for (int k = 0; k < 6; ++k) M[k] = f[k];
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
    return r;
}

}  // namespace trec

// ------------------------------------------------------------------ C ABI

extern "C" {

const char *trec_status_string(trec_status s)
{
    switch (s) {
    case TREC_OK: return "ok";
    case TREC_ERR_INVALID: return "invalid argument";
    case TREC_ERR_UNSUPPORTED: return "unsupported";
    case TREC_ERR_NOMEM: return "out of memory";
    case TREC_ERR_DEVICE: return "device error";
    case TREC_ERR_IO: return "i/o error";
    case TREC_ERR_BUSY: return "busy";
    case TREC_ERR_NOT_READY: return "not ready";
    }
    return "unknown status";
}

const char *trec_version(void) { return "tiltrec-hip 1.0 gfx950"; }

trec_status trec_context_create(int device, trec_context **out)
{
    if (!out) return TREC_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return TREC_ERR_DEVICE;
    }
    if (device < 0 || device >= n) return TREC_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return TREC_ERR_DEVICE;
    trec_context *c = new (std::nothrow) trec_context;
    if (!c) return TREC_ERR_NOMEM;
    c->device = device;
    // the no-allocation entry points' device array: its address on this
    // device, taken once per device with the device current
    if (static_scratch_init(device) != hipSuccess) {
        (void)hipGetLastError();
        delete c;
        return TREC_ERR_DEVICE;
    }
    if (hipStreamCreateWithFlags(&c->cap, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->staging_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->arena_ev, hipEventDisableTiming) != hipSuccess) {
        trec_context_destroy(c);
        return TREC_ERR_DEVICE;
    }
    *out = c;
    return TREC_OK;
}

void trec_context_destroy(trec_context *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
    if (ctx->arena) (void)hipFree(ctx->arena);
    if (ctx->staging) (void)hipHostFree(ctx->staging);
    if (ctx->staging_ev) (void)hipEventDestroy(ctx->staging_ev);
    if (ctx->arena_ev) (void)hipEventDestroy(ctx->arena_ev);
    if (ctx->cap) (void)hipStreamDestroy(ctx->cap);
    for (auto &r : ctx->prof.pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : ctx->prof.pool) (void)hipEventDestroy(e);
    delete ctx;
}

trec_status trec_context_trim(trec_context *ctx)
{
    if (!ctx) return TREC_ERR_INVALID;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipDeviceSynchronize());
    if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
    ctx->gexec = nullptr;
    ctx->gkey.clear();
    if (ctx->arena) (void)hipFree(ctx->arena);
    ctx->arena = nullptr;
    ctx->arena_size = 0;
    return TREC_OK;
}

const char *trec_last_error(const trec_context *ctx) { return ctx ? ctx->err.c_str() : ""; }

trec_status trec_coefficients(const trec_geom *geom, const float *angles_deg, double *coef_out)
{
    if (!geom || !angles_deg || !coef_out || geom->nproj < 1) return TREC_ERR_INVALID;
    coefficients(*geom, angles_deg, coef_out);
    return TREC_OK;
}

trec_status trec_workspace_size(const trec_geom *geom, const trec_method *method, size_t *bytes)
{
    if (!geom || !method || !bytes) return TREC_ERR_INVALID;
    std::vector<float> ang(std::max(geom->nproj, 1), 0.f);
    trec_status s = validate(nullptr, geom, method, ang.data());
    if (s != TREC_OK) return s;
    JobPlan p;
    s = plan_job(nullptr, p, *geom, *method, ang.data(), nullptr, nullptr, nullptr);
    if (s != TREC_OK) return s;
    *bytes = carve(p, nullptr);
    return TREC_OK;
}

trec_status trec_reconstruct(trec_context *ctx, const trec_geom *geom, const float *angles_deg, const float *stack,
                             const float *initial, const trec_method *method, float *volume, void *stream)
{
    if (!ctx) return TREC_ERR_INVALID;
    ctx->err.clear();
    ArenaGuard arena_guard(ctx);
    trec_status s = validate(ctx, geom, method, angles_deg);
    if (s != TREC_OK) return s;
    if (!stack || !volume) return fail(ctx, TREC_ERR_INVALID, "null stack or volume");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    std::vector<JobPlan> plans(1);
    s = plan_job(ctx, plans[0], *geom, *method, angles_deg, stack, initial, volume);
    if (s != TREC_OK) return s;
    // a single job is issued directly unless it is made of many short stages
    const bool graph = plans[0].stages > 64;
    return run_jobs(ctx, plans, (hipStream_t)stream, graph);
}

trec_status trec_reconstruct_batch(trec_context *ctx, int njobs, const trec_job *jobs, void *stream)
{
    if (!ctx) return TREC_ERR_INVALID;
    ctx->err.clear();
    ArenaGuard arena_guard(ctx);
    if (njobs < 0 || (njobs > 0 && !jobs)) return fail(ctx, TREC_ERR_INVALID, "bad job list");
    if (njobs == 0) return TREC_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    std::vector<JobPlan> plans(njobs);
    for (int i = 0; i < njobs; ++i) {
        const trec_job &j = jobs[i];
        trec_status s = validate(ctx, &j.geom, &j.method, j.angles_deg);
        if (s != TREC_OK) return s;
        if (!j.stack || !j.volume) return fail(ctx, TREC_ERR_INVALID, "job %d: null stack or volume", i);
        s = plan_job(ctx, plans[i], j.geom, j.method, j.angles_deg, j.stack, j.initial, j.volume);
        if (s != TREC_OK) return s;
    }
    return run_jobs(ctx, plans, (hipStream_t)stream, true);
}

trec_status trec_filter_rweight(trec_context *ctx, float *rows, int nx, long long nrows, void *stream)
{
    if (!ctx) return TREC_ERR_INVALID;
    ctx->err.clear();
    ArenaGuard arena_guard(ctx);
    if (!rows || nx < 2 || nrows < 0) return fail(ctx, TREC_ERR_INVALID, "bad filter arguments");
    if (nrows == 0) return TREC_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int N = 0;
    std::vector<float> hk = rweight_kernel(nx, N);
    hipStream_t st = (hipStream_t)stream;
    trec_status s = ensure_arena(ctx, align_up(sizeof(float) * N), st);
    if (s != TREC_OK) return s;
    HIPCHK(ctx, hipEventSynchronize(ctx->staging_ev));
    if (ctx->staging_size < sizeof(float) * N) {
        if (ctx->staging) (void)hipHostFree(ctx->staging);
        ctx->staging = nullptr;
        ctx->staging_size = 0;
        HIPCHK(ctx, hipHostMalloc((void **)&ctx->staging, sizeof(float) * N, hipHostMallocDefault));
        ctx->staging_size = sizeof(float) * N;
    }
    memcpy(ctx->staging, hk.data(), sizeof(float) * N);
    HIPCHK(ctx, hipMemcpyAsync(ctx->arena, ctx->staging, sizeof(float) * N, hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipEventRecord(ctx->staging_ev, st));
    HIPCHK(ctx, launch_rweight(rows, (const float *)ctx->arena, N, nx, nrows, 0, 1, 1, nx, st));
    return TREC_OK;
}

trec_status trec_apply_xf(trec_context *ctx, int nx, int ny, int nsec, const float *xf_host, int mode,
                          const float *in, float *out, void *stream)
{
    if (!ctx) return TREC_ERR_INVALID;
    ctx->err.clear();
    ArenaGuard arena_guard(ctx);
    if (nx < 1 || ny < 1 || nsec < 1 || nsec > 65535)
        return fail(ctx, TREC_ERR_INVALID, "bad stack size %d x %d x %d", nx, ny, nsec);
    if ((size_t)nx * ny > ((size_t)1 << 29))
        return fail(ctx, TREC_ERR_INVALID, "a section of %d x %d pixels exceeds 2^29", nx, ny);
    if ((size_t)((nx + XF_TW - 1) / XF_TW) * ((ny + XF_TH - 1) / XF_TH) * nsec >= ((size_t)1 << 26))
        return fail(ctx, TREC_ERR_INVALID, "stack of %d sections of %d x %d is too large", nsec, nx, ny);
    if (!xf_host || !in || !out) return fail(ctx, TREC_ERR_INVALID, "null argument");
    const size_t n = (size_t)nx * ny;
    if (out < in + n * nsec && in < out + n * nsec)
        return fail(ctx, TREC_ERR_INVALID, "input and output stacks overlap");
    std::vector<XfConst> xc(nsec);
    // Row and column terms are below 2^30 when every source coordinate is
    // below 2^20 pixels, so their int sum cannot wrap.
    const double lim = (double)(1 << 20);
    for (int i = 0; i < nsec; ++i) {
        for (int k = 0; k < 6; ++k)
            if (!std::isfinite(xf_host[6 * i + k]))
                return fail(ctx, TREC_ERR_INVALID, "section %d: non-finite .xf value", i);
        xc[i] = xf_inverse(xf_host + 6 * i, nx, ny);
        const double *m = xc[i].m;
        const double X = nx - 1, Y = ny - 1;
        const double t[6] = {std::fabs(m[2]), std::fabs(m[1] * Y + m[2]), std::fabs(m[0] * X),
                             std::fabs(m[5]), std::fabs(m[4] * Y + m[5]), std::fabs(m[3] * X)};
        for (double v : t)
            if (!(v < lim))
                return fail(ctx, TREC_ERR_INVALID, "section %d: the .xf maps pixels 2^20 or more pixels away", i);
    }
    const int lanczos = mode == 2;
    const int K = lanczos ? 8 : 4;
    // LDS window of a tile: its source footprint spans |m0| (TW-1) + |m1| (TH-1)
    // columns and |m3| (TW-1) + |m4| (TH-1) rows, plus the 1/32-pixel rounding
    // and the kernel's K taps.  Inner windows are loaded with a row pitch of
    // 72 or 80 floats in whole 256-float steps; the kernel checks each tile
    // against the reserve and takes a slower path where it does not fit.
    double rw_max = 0, rh_max = 0, rwh_max = 0;
    for (int i = 0; i < nsec; ++i) {
        const double *m = xc[i].m;
        const double rw = std::floor(std::fabs(m[0]) * (XF_TW - 1) + std::fabs(m[1]) * (XF_TH - 1) + 2) + K;
        const double rh = std::floor(std::fabs(m[3]) * (XF_TW - 1) + std::fabs(m[4]) * (XF_TH - 1) + 2) + K;
        rw_max = std::max(rw_max, rw);
        rh_max = std::max(rh_max, rh);
        rwh_max = std::max(rwh_max, rw * rh);
    }
    const int pitch = rw_max <= 72 ? 72 : 80;
    const double win = std::max(std::ceil(pitch * rh_max / 256) * 256, rwh_max);
    const int win_cap = (int)std::min<double>(win, XF_LDS_MAX);
    float tab[XF_TAB * XF_TP] = {};
    for (int i = 0; i < XF_TAB; ++i) {
        const float x = i * (1.f / XF_TAB);
        if (lanczos) xf_lanczos4(x, tab + XF_TP * i);
        else xf_cubic(x, tab + XF_TP * i);
    }
    // partial sums per section: at most 2048 work-groups in all (8 per CU),
    // so the sum pass runs in a single round
    const int nb = (int)std::max<size_t>(1, std::min<size_t>(2048 / nsec, (n + 4095) / 4096));

    const size_t b_xc = align_up(sizeof(XfConst) * nsec), b_tab = align_up(sizeof(float) * XF_TAB * XF_TP);
    const size_t b_rows = align_up(sizeof(int2) * ny * nsec), b_cols = align_up(sizeof(int2) * nx * nsec);
    const size_t b_part = align_up(sizeof(double) * nb * nsec), b_cval = align_up(sizeof(float) * nsec);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    trec_status s = ensure_arena(ctx, b_xc + b_tab + b_rows + b_cols + b_part + b_cval, (hipStream_t)stream);
    if (s != TREC_OK) return s;
    XfLaunch a{};
    a.W = nx;
    a.H = ny;
    a.P = nsec;
    a.nb = nb;
    a.lanczos = lanczos;
    a.win_cap = win_cap;
    a.pitch = pitch;
    char *base = ctx->arena;
    a.xc = (const XfConst *)base;
    a.tab = (const float *)(base + b_xc);
    a.rows = (int2 *)(base + b_xc + b_tab);
    a.cols = (int2 *)(base + b_xc + b_tab + b_rows);
    a.part = (double *)(base + b_xc + b_tab + b_rows + b_cols);
    a.cval = (float *)(base + b_xc + b_tab + b_rows + b_cols + b_part);
    a.src = in;
    a.dst = out;

    const size_t up = b_xc + sizeof(float) * XF_TAB * XF_TP;
    HIPCHK(ctx, hipEventSynchronize(ctx->staging_ev));
    if (ctx->staging_size < up) {
        if (ctx->staging) (void)hipHostFree(ctx->staging);
        ctx->staging = nullptr;
        ctx->staging_size = 0;
        HIPCHK(ctx, hipHostMalloc((void **)&ctx->staging, up, hipHostMallocDefault));
        ctx->staging_size = up;
    }
    memcpy(ctx->staging, xc.data(), sizeof(XfConst) * nsec);
    memcpy(ctx->staging + b_xc, tab, sizeof(float) * XF_TAB * XF_TP);
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(ctx, hipMemcpyAsync(base, ctx->staging, up, hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipEventRecord(ctx->staging_ev, st));
    tl_prof = &ctx->prof;
    hipError_t e;
    {
        ProfScope ps(PROF_XF_MEAN, st);
        e = launch_xf_terms(a, st);
        if (e == hipSuccess) e = launch_xf_mean(a, st);
    }
    if (e == hipSuccess) {
        ProfScope ps(PROF_XF_WARP, st);
        e = launch_xf_warp(a, st);
    }
    tl_prof = nullptr;
    HIPCHK(ctx, e);
    return TREC_OK;
}

trec_status trec_geopar(trec_context *ctx, const float *volume, int nx, int ny, int nz, int axis, int flags,
                        float *views, float *chips, float *result_host, int *exhaustive_host, void *stream)
{
    if (!ctx) return TREC_ERR_INVALID;
    ctx->err.clear();
    ArenaGuard arena_guard(ctx);
    if (!volume) return fail(ctx, TREC_ERR_INVALID, "null volume");
    if (nx < 1 || ny < 1 || nz < 1) return fail(ctx, TREC_ERR_INVALID, "bad volume size %d x %d x %d", nx, ny, nz);
    if (axis != TREC_GP_VIEW_X && axis != TREC_GP_VIEW_Y) return fail(ctx, TREC_ERR_INVALID, "bad view axis %d", axis);
    if (ny > GP_MAXN) return fail(ctx, TREC_ERR_INVALID, "%d rows per view exceed %d", ny, GP_MAXN);
    if (axis == TREC_GP_VIEW_Y && nz > 65535)
        return fail(ctx, TREC_ERR_INVALID, "%d sections exceed 65535 for y views", nz);
    const int nviews = axis == TREC_GP_VIEW_X ? nx : nz, cols = axis == TREC_GP_VIEW_X ? nz : nx;
    const int nchips = cols > GP_WIDTH ? (cols - GP_WIDTH - 1) / GP_STEP + 1 : 0;
    if ((size_t)nviews * nchips > (size_t)INT_MAX)
        return fail(ctx, TREC_ERR_INVALID, "%d views of %d chips are too many", nviews, nchips);
    // partial min/max: about 2048 work-groups over the volume
    int splits;
    if (axis == TREC_GP_VIEW_X) {
        const size_t tiles = (size_t)(nx + WG - 1) / WG;
        splits = (int)std::max<size_t>(1, std::min<size_t>({2048 / tiles, 512, (size_t)nz * ny}));
    } else {
        splits = (int)std::max<size_t>(1, std::min<size_t>({2048 / (size_t)nz, 512, ((size_t)ny * nx + WG - 1) / WG}));
    }
    const size_t b_part = align_up(sizeof(float) * splits * nviews), b_ab = align_up(sizeof(float2) * nviews);
    const size_t b_prof = align_up(sizeof(float) * (size_t)nviews * nchips * ny);
    const size_t b_chips = chips ? 0 : align_up(sizeof(float4) * (size_t)nviews * std::max(nchips, 1));
    const size_t b_views = views ? 0 : align_up(sizeof(float) * 3 * (size_t)nviews);
    const size_t b_small = align_up(sizeof(float) * 4 + sizeof(int) * 4);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    trec_status s = ensure_arena(ctx, 2 * b_part + b_ab + b_prof + b_chips + b_views + b_small, (hipStream_t)stream);
    if (s != TREC_OK) return s;
    char *base = ctx->arena;
    GpLaunch a{};
    a.vol = volume;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.axis = axis;
    a.nviews = nviews;
    a.cols = cols;
    a.nchips = nchips;
    a.exhaustive = (flags & TREC_GP_EXHAUSTIVE) ? 1 : 0;
    a.splits = splits;
    a.pmin = (float *)base;
    a.pmax = (float *)(base + b_part);
    a.ab = (float2 *)(base + 2 * b_part);
    a.prof = (float *)(base + 2 * b_part + b_ab);
    char *p = base + 2 * b_part + b_ab + b_prof;
    a.chips = chips ? (float4 *)chips : (float4 *)p;
    p += b_chips;
    a.views = views ? views : (float *)p;
    p += b_views;
    a.avg = (float *)p;
    a.flags = (int *)(p + sizeof(float) * 4);
    hipStream_t st = (hipStream_t)stream;
    tl_prof = &ctx->prof;
    hipError_t e = launch_zero((float *)a.flags, 4, st);
    if (e == hipSuccess) {
        ProfScope ps(PROF_GP_PROFILE, st);
        e = launch_gp_scale(a, st);
        if (e == hipSuccess) e = launch_gp_profiles(a, st);
    }
    if (e == hipSuccess) {
        ProfScope ps(PROF_GP_SEARCH, st);
        e = launch_gp_search(a, st);
    }
    if (e == hipSuccess) {
        ProfScope ps(PROF_GP_VIEWS, st);
        e = launch_gp_views(a, st);
    }
    tl_prof = nullptr;
    HIPCHK(ctx, e);
    if (result_host || exhaustive_host) {
        float hb[4];
        int hf[4];
        HIPCHK(ctx, hipMemcpyAsync(hb, a.avg, sizeof hb, hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipMemcpyAsync(hf, a.flags, sizeof hf, hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipStreamSynchronize(st));
        if (hf[0]) return fail(ctx, TREC_ERR_INVALID, "the volume holds non-finite values");
        if (result_host) memcpy(result_host, hb, 3 * sizeof(float));
        if (exhaustive_host) *exhaustive_host = hf[1];
    }
    return TREC_OK;
}

trec_status trec_convert(trec_context *ctx, const void *in, int in_type, void *out, int out_type, size_t n,
                         void *stream)
{
    if (!ctx) return TREC_ERR_INVALID;
    ctx->err.clear();
    if (!in || !out) return fail(ctx, TREC_ERR_INVALID, "null buffer");
    if (in_type < TREC_DT_I8 || in_type > TREC_DT_F32) return fail(ctx, TREC_ERR_INVALID, "bad input type %d", in_type);
    if (out_type != TREC_DT_F32 && out_type != TREC_DT_F16)
        return fail(ctx, TREC_ERR_INVALID, "output type %d is not float or half", out_type);
    if (n == 0) return TREC_OK;
    static const size_t esz[] = {1, 1, 1, 2, 2, 2, 4};
    const char *a = (const char *)in, *b = (const char *)out;
    if (a < b + n * esz[out_type] && b < a + n * esz[in_type])
        return fail(ctx, TREC_ERR_INVALID, "input and output overlap");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, launch_convert(in, in_type, out, out_type, n, (hipStream_t)stream));
    return TREC_OK;
}

trec_status trec_float_to_byte(trec_context *ctx, const float *in, size_t n, unsigned char *out,
                               float *stats_host, void *stream)
{
    if (!ctx) return TREC_ERR_INVALID;
    ctx->err.clear();
    ArenaGuard arena_guard(ctx);
    if (!in || !out || n == 0) return fail(ctx, TREC_ERR_INVALID, "bad conversion arguments");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    trec_status s = ensure_arena(ctx, align_up(sizeof(double) * 2048) + align_up(2 * sizeof(float)), (hipStream_t)stream);
    if (s != TREC_OK) return s;
    double *part = (double *)ctx->arena;
    float *stats = (float *)(ctx->arena + align_up(sizeof(double) * 2048));
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(ctx, launch_to_byte(in, n, part, stats, out, st));
    if (stats_host) {
        HIPCHK(ctx, hipMemcpyAsync(stats_host, stats, 2 * sizeof(float), hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipStreamSynchronize(st));
    }
    return TREC_OK;
}

trec_status trec_filter_fbp(trec_context *ctx, float *rows, int nx, long long nrows, int mode, void *stream)
{
    if (!ctx) return TREC_ERR_INVALID;
    ctx->err.clear();
    ArenaGuard arena_guard(ctx);
    if (!rows || nx < 1 || nrows < 0 || (mode != 0 && mode != 1))
        return fail(ctx, TREC_ERR_INVALID, "bad filter arguments");
    if (nrows == 0) return TREC_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int N = 0;
    std::vector<float> h = spatial_taps(mode, nx, N);
    const size_t bytes = sizeof(float) * h.size();
    trec_status s = ensure_arena(ctx, align_up(bytes), (hipStream_t)stream);
    if (s != TREC_OK) return s;
    HIPCHK(ctx, hipEventSynchronize(ctx->staging_ev));
    if (ctx->staging_size < bytes) {
        if (ctx->staging) (void)hipHostFree(ctx->staging);
        ctx->staging = nullptr;
        ctx->staging_size = 0;
        HIPCHK(ctx, hipHostMalloc((void **)&ctx->staging, bytes, hipHostMallocDefault));
        ctx->staging_size = bytes;
    }
    memcpy(ctx->staging, h.data(), bytes);
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(ctx, hipMemcpyAsync(ctx->arena, ctx->staging, bytes, hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipEventRecord(ctx->staging_ev, st));
    HIPCHK(ctx, launch_fbp_filter(rows, (const float *)ctx->arena, N, nx, nrows, 0, 1, 1, nx, st));
    return TREC_OK;
}

trec_status trec_forward_project(trec_context *ctx, const trec_geom *geom, const float *angles_deg,
                                 const float *volume, float *rval, float *rwt, void *stream)
{
    if (!ctx) return TREC_ERR_INVALID;
    ctx->err.clear();
    ArenaGuard arena_guard(ctx);
    trec_method m{TREC_SIRT, 0, 1.0f, TREC_STORE_F32, TREC_LAYOUT_Y};
    trec_status s = validate(ctx, geom, &m, angles_deg);
    if (s != TREC_OK) return s;
    if (!volume || !rval) return fail(ctx, TREC_ERR_INVALID, "null volume or rval");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    std::vector<JobPlan> plans(1);
    JobPlan &p = plans[0];
    p.want_rwt = true;
    s = plan_job(ctx, p, *geom, m, angles_deg, volume, nullptr, nullptr);
    if (s != TREC_OK) return s;
    const size_t X = geom->nx, Y = geom->ny, T = geom->thickness, P = geom->nproj;
    if (p.path == PATH_ROWS) {
        p.h.nzc = 1;
        p.h.zc_len = (int)T;
        p.m.method = TREC_SIRT;
        // the partial buffer of one z chunk holds the raw projection
        size_t bytes = carve(p, nullptr) + align_up(4 * (size_t)p.h.nyb * P * X * RB);
        s = ensure_arena(ctx, bytes, st);
        if (s != TREC_OK) return s;
        carve(p, ctx->arena);
        p.h.part = (float *)(ctx->arena + p.bytes);
        s = upload(ctx, {&p}, st);
        if (s != TREC_OK) return s;
        JobDesc &h = p.h;
        HIPCHK(ctx, launch_block_volume(h, p.d, volume, TREC_LAYOUT_Y, st));
        LaunchList L{};
        L.n = 1;
        L.job[0] = p.d;
        L.off[0] = 0;
        L.off[1] = fwd_blocks(h, (int)P);
        HIPCHK(ctx, launch_forward(L, L.off[1], TREC_STORE_F32, false, st));
        HIPCHK(ctx, launch_unblock_proj(h.part, (int)P, (int)X, (int)Y, h.nyb, rval, st));
        if (rwt) {
            JobDesc o = h;
            o.nyb = 1;
            L.job[0] = p.d_ones;
            L.off[1] = forward_blocks(o, (int)P);
            HIPCHK(ctx, launch_forward(L, L.off[1], -1, false, st));
            HIPCHK(ctx, launch_broadcast_rows(h.rwt, (int)P, (int)X, (int)Y, rwt, st));
        }
        return TREC_OK;
    }
    s = ensure_arena(ctx, carve(p, nullptr), st);
    if (s != TREC_OK) return s;
    carve(p, ctx->arena);
    s = upload(ctx, {&p}, st);
    if (s != TREC_OK) return s;
    GeneralJob q = p.gj;
    q.vol = const_cast<float *>(volume);
    q.layout = TREC_LAYOUT_Y;
    q.rval = rval;
    HIPCHK(ctx, launch_zero(rval, P * Y * X, st));
    HIPCHK(ctx, launch_general_forward(q, p.tg.data(), 0, (int)P, false, st));
    if (rwt) {
        q.rwt = rwt;
        HIPCHK(ctx, launch_zero(rwt, P * Y * X, st));
        HIPCHK(ctx, launch_general_forward(q, p.tg.data(), 0, (int)P, true, st));
    }
    return TREC_OK;
}

trec_status trec_backproject(trec_context *ctx, const trec_geom *geom, const float *angles_deg, const float *stack,
                             float *volume, void *stream)
{
    trec_method m{TREC_BPT, 0, 1.0f, TREC_STORE_F32, TREC_LAYOUT_Y};
    return trec_reconstruct(ctx, geom, angles_deg, stack, volume, &m, volume, stream);
}

trec_status trec_sirt_iterate(trec_context *ctx, const trec_geom *geom, const float *angles_deg, const float *stack,
                              float gamma, float *volume, void *stream)
{
    trec_method m{TREC_SIRT, 1, gamma, TREC_STORE_F32, TREC_LAYOUT_Y};
    return trec_reconstruct(ctx, geom, angles_deg, stack, volume, &m, volume, stream);
}

trec_status trec_set_workspace_limit(trec_context *ctx, size_t bytes)
{
    if (!ctx) return TREC_ERR_INVALID;
    ctx->ws_limit = bytes;
    return TREC_OK;
}

trec_status trec_plan(trec_context *ctx, const trec_geom *geom, const float *angles_deg, const trec_method *method,
                      trec_plan_info *info)
{
    if (!ctx || !info) return TREC_ERR_INVALID;
    ctx->err.clear();
    trec_status s = validate(ctx, geom, method, angles_deg);
    if (s != TREC_OK) return s;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    JobPlan p;
    s = plan_job(ctx, p, *geom, *method, angles_deg, nullptr, nullptr, nullptr);
    if (s != TREC_OK) return s;
    const size_t budget = workspace_budget(ctx);
    if (carve(p, nullptr) > budget && !split_slabs(p, budget))
        return fail(ctx, TREC_ERR_NOMEM, "workspace of one row block exceeds the budget of %zu bytes", budget);
    memset(info, 0, sizeof *info);
    info->path = p.path;
    info->workspace = carve(p, nullptr);
    info->slabs = p.nslab;
    info->slab_rows = p.path == PATH_ROWS ? std::min(p.slab_rb * RB, geom->ny) : geom->ny;
    if (p.path == PATH_ROWS) {
        info->zchunks = p.h.nzc;
        if (uses_forward(p.m.method))
            info->band_cells = (p.m.method == TREC_SIRT && p.h.nzc == 1 && p.fg_small) ? FGS_SMALL : FGS;
        info->back_tilts = p.h.bp_g ? p.h.bp_g : back_shapes(p.m.storage)[0].first;
        info->back_voxels_z = p.h.bp_zt ? p.h.bp_zt : back_shapes(p.m.storage)[0].second;
    }
    return TREC_OK;
}

trec_status trec_request_footprint(const trec_geom *geom, const trec_method *method, size_t *bytes)
{
    if (!geom || !method || !bytes) return TREC_ERR_INVALID;
    std::vector<float> ang(std::max(geom->nproj, 1), 0.f);
    trec_status s = validate(nullptr, geom, method, ang.data());
    if (s != TREC_OK) return s;
    *bytes = request_bytes_bound(*geom, *method);
    return TREC_OK;
}

trec_status trec_profile_enable(trec_context *ctx, int on)
{
    if (!ctx) return TREC_ERR_INVALID;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    for (auto &r : ctx->prof.pending) {
        (void)hipEventSynchronize(r.b);
        ctx->prof.pool.push_back(r.a);
        ctx->prof.pool.push_back(r.b);
    }
    ctx->prof.pending.clear();
    for (int k = 0; k < PROF_N; ++k) {
        ctx->prof.total_ms[k] = 0;
        ctx->prof.count[k] = 0;
    }
    ctx->prof.on = on != 0;
    return TREC_OK;
}

trec_status trec_profile_read(trec_context *ctx, const char *kernel, double *total_ms, long long *launches)
{
    if (!ctx || !kernel) return TREC_ERR_INVALID;
    int k = 0;
    while (k < PROF_N && strcmp(kernel, prof_names[k])) ++k;
    if (k == PROF_N) return fail(ctx, TREC_ERR_INVALID, "unknown kernel name %s", kernel);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    for (auto &r : ctx->prof.pending) {
        HIPCHK(ctx, hipEventSynchronize(r.b));
        float ms = 0.f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, r.a, r.b));
        ctx->prof.total_ms[r.kind] += ms;
        ctx->prof.count[r.kind] += 1;
        ctx->prof.pool.push_back(r.a);
        ctx->prof.pool.push_back(r.b);
    }
    ctx->prof.pending.clear();
    if (total_ms) *total_ms = ctx->prof.total_ms[k];
    if (launches) *launches = ctx->prof.count[k];
    return TREC_OK;
}

trec_status trec_autotune(trec_context *ctx, const trec_geom *geom, const trec_method *method)
{
    if (!ctx) return TREC_ERR_INVALID;
    ctx->err.clear();
    ArenaGuard arena_guard(ctx);
    if (!geom || !method) return fail(ctx, TREC_ERR_INVALID, "null geometry or method");
    std::vector<float> ang(std::max(geom->nproj, 1));
    for (int i = 0; i < geom->nproj; ++i) ang[i] = geom->nproj > 1 ? -60.f + 120.f * i / (geom->nproj - 1) : 0.f;
    trec_status s = validate(ctx, geom, method, ang.data());
    if (s != TREC_OK) return s;
    if (method->method != TREC_SIRT) return TREC_OK;  // only SIRT has tunable shapes
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const size_t X = geom->nx, Y = geom->ny, T = geom->thickness, P = geom->nproj;
    float *stack = nullptr, *vol = nullptr;
    HIPCHK(ctx, hipMalloc((void **)&stack, 4 * P * Y * X));
    hipError_t e = hipMalloc((void **)&vol, 4 * X * Y * T);
    if (e != hipSuccess) {
        (void)hipFree(stack);
        return fail(ctx, TREC_ERR_NOMEM, "autotune buffers: %s", hipGetErrorString(e));
    }
    hipStream_t st = ctx->cap;
    if (launch_zero(stack, P * Y * X, st) != hipSuccess || launch_zero(vol, X * Y * T, st) != hipSuccess) {
        (void)hipFree(stack);
        (void)hipFree(vol);
        return fail(ctx, TREC_ERR_DEVICE, "autotune: zero fill failed");
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    constexpr int REPS = 6;
    trec_status res = TREC_OK;
    bool rows_path = true;
    // the best of REPS whole SIRT iterations (after one warm-up) with z
    // chunks nzc and back-projection shape (g, zt); 0 for the shape is the
    // default
    auto measure = [&](int nzc, int g, int zt, float &ms) -> bool {
        JobPlan p;
        trec_method m = *method;
        m.iterations = 1 + REPS;
        if ((res = plan_job(ctx, p, *geom, m, ang.data(), stack, nullptr, vol, nzc)) != TREC_OK) return false;
        if (p.path != PATH_ROWS) {
            rows_path = false;
            return false;
        }
        p.h.bp_g = g;
        p.h.bp_zt = zt;
        if ((res = ensure_arena(ctx, carve(p, nullptr), st)) != TREC_OK) return false;
        carve(p, ctx->arena);
        if ((res = upload(ctx, {&p}, st)) != TREC_OK) return false;
        std::vector<JobPlan *> pl{&p};
        std::vector<int> stage{0};
        if (issue_prep(p, st) != hipSuccess || issue_step(pl, stage, st) != hipSuccess) {
            res = fail(ctx, TREC_ERR_DEVICE, "autotune launch failed");
            return false;
        }
        ms = 1e30f;
        for (int rep = 0; rep < REPS; ++rep) {
            float t = 0.f;
            if (hipEventRecord(e0, st) != hipSuccess || issue_step(pl, stage, st) != hipSuccess ||
                hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                res = fail(ctx, TREC_ERR_DEVICE, "autotune run failed");
                return false;
            }
            (void)hipEventElapsedTime(&t, e0, e1);
            ms = std::min(ms, t);
        }
        return true;
    };
    // Candidates are tried in order of preference (the first is the
    // default).  A later one is kept only when it beats the best so far by
    // more than `margin`: a step varies by a percent or two between runs, so
    // this keeps the choice stable when candidates are close.
    Tuned best;
    float best_ms = 1e30f;
    const float margin = 0.02f;
    const int nzcs[] = {1, 2, 4, 8, 16, 32};
    for (int nzc : nzcs) {
        if (nzc > (int)T) break;
        float ms;
        if (!measure(nzc, 0, 0, ms)) break;
        if (ms < best_ms * (1.f - margin)) {
            best_ms = ms;
            best.nzc = nzc;
        }
    }
    if (res == TREC_OK && rows_path && best.nzc > 0) {
        const auto &shapes = back_shapes(method->storage);
        for (size_t k = 1; k < shapes.size(); ++k) {
            float ms;
            if (!measure(best.nzc, shapes[k].first, shapes[k].second, ms)) break;
            if (ms < best_ms * (1.f - margin)) {
                best_ms = ms;
                best.bp_g = shapes[k].first;
                best.bp_zt = shapes[k].second;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(st);
    (void)hipFree(stack);
    (void)hipFree(vol);
    if (res != TREC_OK) return res;
    if (best.nzc == 0) return TREC_OK;  // the general path: nothing to tune
    const std::string key = tune_key(*geom, method->storage);
    ctx->tuned[key] = best;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[key] = best;
    }
    return TREC_OK;
}

}  // extern "C"

