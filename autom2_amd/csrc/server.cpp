// server.cpp -- the serving loop of libtiltrec_hip.so (trec_server_*).
//
// AuTom2 starts one TiltRec process per tilt series; a server instead keeps
// one process and one device arena and reconstructs series as they arrive.
// Two host threads share the work:
//
//   the loop thread repeats a step:
//     1. admit queued requests (first come, first served) while fewer than
//        max_active are in flight and the request's footprint fits the
//        arena; an admitted request's stack is handed to the copier;
//     2. prepare the requests whose stack arrived (plans, descriptors,
//        layout kernels) and add them to the active set;
//     3. advance every active request by one stage (a SIRT iteration, a
//        SART tilt, a BPT/WBP pass, a whole ADMM recursion): the stages of
//        all active requests are merged into one launch per kernel kind;
//     4. for requests whose last stage ran, queue the output kernels and
//        hand the volume to the copier.
//   two copier threads, one for stacks in and one for volumes out, move
//   data through two pinned staging buffers each on streams of their own,
//   so large copies overlap the reconstructions of the other requests and
//   each other: the host-side memcpy of one chunk runs while the DMA of the
//   next is in flight.  When a volume is out the request's arena block is
//   freed and the waiters are woken.
//
// A request that arrives while others run joins at the next step, so the
// device stays busy with a changing mix.  The kernels compute each job from
// its own data in a fixed order, so a request's volume is bit for bit the
// same whether it ran alone or in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "range_alloc.h"
#include "trec_host.h"

using namespace trec;

namespace {

constexpr size_t ALIGN = 256;
size_t align_up(size_t v) { return (v + ALIGN - 1) & ~(ALIGN - 1); }

// staging chunk of the copier (two of them, pinned)
constexpr size_t CHUNK = 32u << 20;

// memcpy of a staging chunk split over a few host threads (one thread
// copies about 10 GB/s, well below the DMA rate)
void par_memcpy(char *dst, const char *src, size_t n)
{
    constexpr int NT = 8;
    if (n < ((size_t)4 << 20)) {
        memcpy(dst, src, n);
        return;
    }
    const size_t part = ((n + NT - 1) / NT + 4095) & ~(size_t)4095;  // part * NT >= n
    std::thread th[NT - 1];
    for (int t = 1; t < NT; ++t) {
        const size_t o = std::min(n, part * t), e = std::min(n, part * (t + 1));
        th[t - 1] = std::thread([=] { memcpy(dst + o, src + o, e - o); });
    }
    memcpy(dst, src, std::min(n, part));
    for (auto &t : th) t.join();
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over one device allocation (range_alloc.h keeps the
// offsets).
class Arena {
  public:
    bool init(size_t bytes)
    {
        if (hipMalloc((void **)&base_, bytes) != hipSuccess) {
            (void)hipGetLastError();
            base_ = nullptr;
            return false;
        }
        ranges_.reset(bytes);
        return true;
    }
    ~Arena()
    {
        if (base_) (void)hipFree(base_);
    }
    char *alloc(size_t bytes)
    {
        const size_t off = ranges_.alloc(bytes);
        return off == RangeAllocator::NONE ? nullptr : base_ + off;
    }
    void release(char *p) { ranges_.release((size_t)(p - base_)); }
    size_t size() const { return ranges_.size(); }
    bool fits(size_t bytes) const { return ranges_.fits(bytes); }

  private:
    char *base_ = nullptr;
    RangeAllocator ranges_;
};

struct Request {
    uint64_t id = 0;
    trec_geom g{};
    trec_method m{};
    std::vector<float> angles;
    const void *stack_host = nullptr;
    void *volume_out = nullptr;
    bool stack_half = false, volume_half = false;
    size_t footprint = 0;
    int nzc = 0;  // z chunks the footprint was planned with
    double submit_s = 0, admit_s = 0, done_s = 0, staged_s = 0, computed_s = 0;
    trec_status status = TREC_OK;
    bool done = false;
    char *block = nullptr;  // arena block: stack | volume | workspace
    hipEvent_t out_ready = nullptr;  // the volume is final on the device
    JobPlan plan;

    size_t stack_elems() const { return (size_t)g.nproj * g.ny * g.nx; }
    size_t volume_elems() const { return (size_t)g.nx * g.ny * g.thickness; }
    size_t stack_bytes() const { return 4 * stack_elems(); }    // on the device, float
    size_t volume_bytes() const { return 4 * volume_elems(); }  // on the device, float
    // bytes crossing PCIe
    size_t stack_io_bytes() const { return (stack_half ? 2 : 4) * stack_elems(); }
    size_t volume_io_bytes() const { return (volume_half ? 2 : 4) * volume_elems(); }
    // half staging for half input or output (after the workspace)
    size_t half_bytes() const
    {
        return 2 * std::max(stack_half ? stack_elems() : 0, volume_half ? volume_elems() : 0);
    }
    float *dstack() const { return (float *)block; }
    float *dvolume() const { return (float *)(block + align_up(stack_bytes())); }
    char *workspace() const { return block + align_up(stack_bytes()) + align_up(volume_bytes()); }
    size_t workspace_room() const { return footprint - align_up(stack_bytes()) - align_up(volume_bytes()) - align_up(half_bytes()); }
    void *dhalf() const { return block + footprint - align_up(half_bytes()); }
    void *dstack_in() const { return stack_half ? dhalf() : (void *)dstack(); }
    const void *dvolume_out() const { return volume_half ? dhalf() : (const void *)dvolume(); }
};

// One direction of host <-> device traffic.
struct Copier {
    hipStream_t stream = nullptr;
    char *pinned[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::deque<Request *> jobs;
    std::thread th;

    bool init()
    {
        if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return false;
        for (int b = 0; b < 2; ++b)
            if (hipHostMalloc((void **)&pinned[b], CHUNK, hipHostMallocDefault) != hipSuccess ||
                hipEventCreateWithFlags(&ev[b], hipEventDisableTiming) != hipSuccess)
                return false;
        return true;
    }
    void release()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (int b = 0; b < 2; ++b) {
            if (pinned[b]) (void)hipHostFree(pinned[b]);
            if (ev[b]) (void)hipEventDestroy(ev[b]);
            pinned[b] = nullptr;
            ev[b] = nullptr;
        }
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};

}  // namespace

struct trec_server {
    int device = 0;
    int max_active = 8;
    trec_context *ctx = nullptr;
    hipStream_t stream = nullptr;  // reconstructions
    Copier cin, cout;              // stacks in, volumes out
    Arena arena;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_copy;
    bool stop = false;          // no more requests: the loop ends when all are done
    bool copier_stop = false;   // the loop has ended: the copier ends when idle
    uint64_t next_id = 1;
    int in_flight = 0;                  // admitted and not finished
    std::deque<Request *> queue;        // waiting for admission
    std::deque<Request *> ready;        // stack on the device, not yet prepared
    std::map<uint64_t, std::unique_ptr<Request>> reqs;
    std::thread loop;

    void run();
    void run_copier(Copier &c, bool in);
    trec_status prepare(Request &r);
    trec_status copy_in(Copier &c, Request &r);
    trec_status copy_out(Copier &c, Request &r);
    void finish(Request &r, trec_status st);  // with mu held
};

// Stack: host -> pinned chunk (memcpy) -> device (DMA).  The memcpy of chunk
// k + 1 runs while chunk k's DMA is in flight.
trec_status trec_server::copy_in(Copier &c, Request &r)
{
    const char *src = reinterpret_cast<const char *>(r.stack_host);
    char *dst = reinterpret_cast<char *>(r.dstack_in());
    const size_t n = r.stack_io_bytes();
    int b = 0;
    for (size_t off = 0; off < n; off += CHUNK, b ^= 1) {
        const size_t len = std::min(CHUNK, n - off);
        if (hipEventSynchronize(c.ev[b]) != hipSuccess) return TREC_ERR_DEVICE;  // buffer b is free
        par_memcpy(c.pinned[b], src + off, len);
        if (hipMemcpyAsync(dst + off, c.pinned[b], len, hipMemcpyHostToDevice, c.stream) != hipSuccess ||
            hipEventRecord(c.ev[b], c.stream) != hipSuccess)
            return TREC_ERR_DEVICE;
    }
    return hipStreamSynchronize(c.stream) == hipSuccess ? TREC_OK : TREC_ERR_DEVICE;
}

// Volume: device -> pinned chunk (DMA) -> host (memcpy), chunk k + 1's DMA
// overlapping chunk k's memcpy.
trec_status trec_server::copy_out(Copier &c, Request &r)
{
    const char *src = reinterpret_cast<const char *>(r.dvolume_out());
    char *dst = reinterpret_cast<char *>(r.volume_out);
    const size_t n = r.volume_io_bytes();
    if (hipStreamWaitEvent(c.stream, r.out_ready, 0) != hipSuccess) return TREC_ERR_DEVICE;
    const size_t nchunk = (n + CHUNK - 1) / CHUNK;
    auto issue = [&](size_t k) {
        const size_t off = k * CHUNK, len = std::min(CHUNK, n - off);
        return hipMemcpyAsync(c.pinned[k & 1], src + off, len, hipMemcpyDeviceToHost, c.stream) == hipSuccess &&
               hipEventRecord(c.ev[k & 1], c.stream) == hipSuccess;
    };
    if (nchunk && !issue(0)) return TREC_ERR_DEVICE;
    for (size_t k = 0; k < nchunk; ++k) {
        if (k + 1 < nchunk && !issue(k + 1)) return TREC_ERR_DEVICE;
        if (hipEventSynchronize(c.ev[k & 1]) != hipSuccess) return TREC_ERR_DEVICE;
        if (k == 0) r.computed_s = now_s();
        const size_t off = k * CHUNK, len = std::min(CHUNK, n - off);
        par_memcpy(dst + off, c.pinned[k & 1], len);
    }
    return TREC_OK;
}

void trec_server::finish(Request &r, trec_status st)
{
    if (r.status == TREC_OK) r.status = st;
    if (r.block) arena.release(r.block);
    r.block = nullptr;
    if (r.out_ready) (void)hipEventDestroy(r.out_ready);
    r.out_ready = nullptr;
    r.done = true;
    r.done_s = now_s();
    --in_flight;
    cv_done.notify_all();
    cv_work.notify_all();
}

void trec_server::run_copier(Copier &c, bool in)
{
    (void)hipSetDevice(device);
    for (;;) {
        Request *job;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_copy.wait(lk, [&] { return copier_stop || !c.jobs.empty(); });
            if (c.jobs.empty()) return;
            job = c.jobs.front();
            c.jobs.pop_front();
        }
        Request &r = *job;
        if (in) {
            const trec_status st = copy_in(c, r);
            r.staged_s = now_s();
            std::lock_guard<std::mutex> lk(mu);
            if (st == TREC_OK) {
                ready.push_back(&r);
                cv_work.notify_all();
            } else {
                finish(r, st);
            }
        } else {
            const trec_status st = copy_out(c, r);
            std::lock_guard<std::mutex> lk(mu);
            finish(r, st);
        }
    }
}

// Plans an arrived request and queues its preparation kernels.
trec_status trec_server::prepare(Request &r)
{
    trec_status s = plan_job(ctx, r.plan, r.g, r.m, r.angles.data(), r.dstack(), nullptr, r.dvolume(), r.nzc);
    if (s != TREC_OK) return s;
    // the block was sized at submit from the same angles and z chunks; the
    // workspace must not run into the half staging or the next block
    if (carve(r.plan, nullptr) > r.workspace_room())
        return fail(ctx, TREC_ERR_INVALID, "request %llu: workspace exceeds its admitted footprint",
                    (unsigned long long)r.id);
    carve(r.plan, r.workspace());
    // a half stack is widened to float on the device (copy_in has finished)
    if (r.stack_half &&
        launch_convert(r.dhalf(), TREC_DT_F16, r.dstack(), TREC_DT_F32, r.stack_elems(), stream) != hipSuccess)
        return TREC_ERR_DEVICE;
    s = upload(ctx, {&r.plan}, stream);
    if (s != TREC_OK) return s;
    if (issue_prep(r.plan, stream) != hipSuccess) return TREC_ERR_DEVICE;
    return TREC_OK;
}

void trec_server::run()
{
    (void)hipSetDevice(device);
    std::vector<Request *> active;
    std::vector<int> stage;
    // Two steps in flight: after issuing step k the loop waits for step k - 1
    // only, so admission, preparation and the issue of step k + 1 happen on
    // the host while step k runs on the device.
    hipEvent_t step_ev[2] = {nullptr, nullptr};
    for (auto &ev : step_ev)
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) ev = nullptr;
    int cur = 0;
    bool prev_pending = false;
    for (;;) {
        std::vector<Request *> arrived;
        {
            std::unique_lock<std::mutex> lk(mu);
            auto admissible = [&] {
                return !queue.empty() && in_flight < max_active && arena.fits(queue.front()->footprint);
            };
            if (active.empty())
                cv_work.wait(lk, [&] { return (stop && queue.empty() && in_flight == 0) || !ready.empty() || admissible(); });
            if (stop && queue.empty() && in_flight == 0 && active.empty()) {
                for (auto ev : step_ev)
                    if (ev) (void)hipEventDestroy(ev);
                return;
            }
            // first come, first served: stop at the first request that does not fit yet
            while (!queue.empty() && in_flight < max_active) {
                Request *r = queue.front();
                r->block = arena.alloc(r->footprint);
                if (!r->block) break;
                queue.pop_front();
                r->admit_s = now_s();
                ++in_flight;
                if (hipEventCreateWithFlags(&r->out_ready, hipEventDisableTiming) != hipSuccess) {
                    r->out_ready = nullptr;
                    finish(*r, TREC_ERR_DEVICE);
                    continue;
                }
                cin.jobs.push_back(r);
                cv_copy.notify_all();
            }
            while (!ready.empty()) {
                arrived.push_back(ready.front());
                ready.pop_front();
            }
        }
        for (Request *r : arrived) {
            const trec_status st = prepare(*r);
            if (st == TREC_OK) {
                active.push_back(r);
                stage.push_back(0);
            } else {
                (void)hipStreamSynchronize(stream);
                std::lock_guard<std::mutex> lk(mu);
                finish(*r, st);
            }
        }
        if (active.empty()) continue;
        std::vector<JobPlan *> plans;
        for (Request *r : active) plans.push_back(&r->plan);
        hipError_t e = issue_step(plans, stage, stream);
        std::vector<size_t> finished;
        for (size_t k = 0; k < active.size() && e == hipSuccess; ++k) {
            if (stage[k] < active[k]->plan.stages) continue;
            Request &r = *active[k];
            e = issue_finish(r.plan, stream);
            if (e == hipSuccess && r.volume_half)
                e = launch_convert(r.dvolume(), TREC_DT_F32, r.dhalf(), TREC_DT_F16, r.volume_elems(), stream);
            if (e == hipSuccess) e = hipEventRecord(r.out_ready, stream);
            finished.push_back(k);
        }
        if (e == hipSuccess) {
            if (!step_ev[0] || !step_ev[1]) {
                e = hipStreamSynchronize(stream);  // no events: one step at a time
            } else {
                e = hipEventRecord(step_ev[cur], stream);
                if (e == hipSuccess && prev_pending) e = hipEventSynchronize(step_ev[cur ^ 1]);
                prev_pending = true;
                cur ^= 1;
            }
        }
        if (e != hipSuccess) {
            // a device failure ends every active request with an error
            (void)hipGetLastError();
            (void)hipStreamSynchronize(stream);
            prev_pending = false;
            std::lock_guard<std::mutex> lk(mu);
            for (Request *r : active) finish(*r, TREC_ERR_DEVICE);
            active.clear();
            stage.clear();
            continue;
        }
        if (finished.empty()) continue;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (size_t k : finished) cout.jobs.push_back(active[k]);
            cv_copy.notify_all();
        }
        std::vector<Request *> keep_r;
        std::vector<int> keep_s;
        size_t f = 0;
        for (size_t k = 0; k < active.size(); ++k) {
            if (f < finished.size() && finished[f] == k) {
                ++f;
                continue;
            }
            keep_r.push_back(active[k]);
            keep_s.push_back(stage[k]);
        }
        active.swap(keep_r);
        stage.swap(keep_s);
    }
}

extern "C" {

trec_status trec_server_create(int device, size_t hbm_budget, int max_active, trec_server **out)
{
    if (!out || max_active < 1 || max_active > MAX_JOBS) return TREC_ERR_INVALID;
    *out = nullptr;
    std::unique_ptr<trec_server> s(new (std::nothrow) trec_server);
    if (!s) return TREC_ERR_NOMEM;
    trec_status st = trec_context_create(device, &s->ctx);
    if (st != TREC_OK) return st;
    s->device = device;
    s->max_active = max_active;
    auto fail_out = [&](trec_status code) {
        s->cin.release();
        s->cout.release();
        if (s->stream) (void)hipStreamDestroy(s->stream);
        trec_context_destroy(s->ctx);
        return code;
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) return fail_out(TREC_ERR_DEVICE);
    if (!s->cin.init() || !s->cout.init()) return fail_out(TREC_ERR_NOMEM);
    if (hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return fail_out(TREC_ERR_DEVICE);
        hbm_budget = fr / 10 * 9;
    }
    if (!s->arena.init(align_up(hbm_budget))) return fail_out(TREC_ERR_NOMEM);
    trec_server *raw = s.release();
    raw->cin.th = std::thread([raw] { raw->run_copier(raw->cin, true); });
    raw->cout.th = std::thread([raw] { raw->run_copier(raw->cout, false); });
    raw->loop = std::thread([raw] { raw->run(); });
    *out = raw;
    return TREC_OK;
}

trec_status trec_server_submit(trec_server *srv, const trec_request *req, uint64_t *id_out)
{
    if (!srv || !req || !id_out) return TREC_ERR_INVALID;
    if (!req->stack || !req->volume_out) return TREC_ERR_INVALID;
    trec_status s = validate(nullptr, &req->geom, &req->method, req->angles_deg);
    if (s != TREC_OK) return s;
    auto r = std::make_unique<Request>();
    r->g = req->geom;
    r->m = req->method;
    r->angles.assign(req->angles_deg, req->angles_deg + req->geom.nproj);
    r->stack_host = req->stack;
    r->volume_out = req->volume_out;
    if ((req->stack_half != 0 && req->stack_half != 1) || (req->volume_half != 0 && req->volume_half != 1))
        return TREC_ERR_INVALID;
    r->stack_half = req->stack_half == 1;
    r->volume_half = req->volume_half == 1;
    r->footprint = request_bytes(req->geom, req->method, req->angles_deg, &r->nzc);
    if (r->footprint == 0) return TREC_ERR_INVALID;
    r->footprint += align_up(r->half_bytes());
    if (align_up(r->footprint) > srv->arena.size()) return TREC_ERR_BUSY;  // can never fit
    r->submit_s = now_s();
    std::lock_guard<std::mutex> lk(srv->mu);
    r->id = srv->next_id++;
    *id_out = r->id;
    srv->queue.push_back(r.get());
    srv->reqs[r->id] = std::move(r);
    srv->cv_work.notify_all();
    return TREC_OK;
}

trec_status trec_server_poll(trec_server *srv, uint64_t id, int *done)
{
    if (!srv || !done) return TREC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return TREC_ERR_INVALID;
    *done = it->second->done ? 1 : 0;
    return TREC_OK;
}

trec_status trec_server_wait(trec_server *srv, uint64_t id, trec_result *res)
{
    if (!srv) return TREC_ERR_INVALID;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return TREC_ERR_INVALID;
    Request *r = it->second.get();
    srv->cv_done.wait(lk, [&] { return r->done; });
    if (res) {
        res->status = r->status;
        res->submit_s = r->submit_s;
        res->admit_s = r->admit_s;
        res->done_s = r->done_s;
        res->staged_s = r->staged_s;
        res->computed_s = r->computed_s;
        res->footprint = r->footprint;
    }
    const trec_status st = r->status;
    srv->reqs.erase(it);
    return st;
}

void trec_server_destroy(trec_server *srv)
{
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
        srv->cv_work.notify_all();
    }
    if (srv->loop.joinable()) srv->loop.join();
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->copier_stop = true;
        srv->cv_copy.notify_all();
    }
    if (srv->cin.th.joinable()) srv->cin.th.join();
    if (srv->cout.th.joinable()) srv->cout.th.join();
    (void)hipStreamSynchronize(srv->stream);
    srv->cin.release();
    srv->cout.release();
    (void)hipStreamDestroy(srv->stream);
    trec_context_destroy(srv->ctx);
    delete srv;
}

}  // extern "C"
