// server.cpp -- the serving loop of libtiltrec_hip.so (trec_server_*).
//
// AuTom2 starts one TiltRec process per tilt series; a server instead keeps
// one process and one device arena and reconstructs series as they arrive.
// Its loop thread repeats a step:
//   1. admit queued requests (first come, first served) while the active
//      set is below max_active and the request's footprint fits the arena;
//      an admitted request's stack is copied in and its plan prepared;
//   2. advance every active request by one stage (a SIRT iteration, a SART
//      tilt, a BPT/WBP pass): the stages of all active requests are merged
//      into one launch per kernel kind;
//   3. finish the requests whose last stage ran: copy the volume out, free
//      their arena block, wake the waiters.
// A request that arrives while others run joins at the next step, so the
// device stays busy with a changing mix.  The kernels compute each job from
// its own data in a fixed order, so a request's volume is bit for bit the
// same whether it ran alone or in any mix.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "trec_host.h"

using namespace trec;

namespace {

constexpr size_t ALIGN = 256;
size_t align_up(size_t v) { return (v + ALIGN - 1) & ~(ALIGN - 1); }

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over one device allocation.
class Arena {
  public:
    bool init(size_t bytes)
    {
        if (hipMalloc((void **)&base_, bytes) != hipSuccess) {
            (void)hipGetLastError();
            base_ = nullptr;
            return false;
        }
        size_ = bytes;
        free_[0] = bytes;
        return true;
    }
    ~Arena()
    {
        if (base_) (void)hipFree(base_);
    }
    char *alloc(size_t bytes)
    {
        bytes = align_up(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const size_t off = it->first, len = it->second;
                free_.erase(it);
                if (len > bytes) free_[off + bytes] = len - bytes;
                used_[off] = bytes;
                return base_ + off;
            }
        }
        return nullptr;
    }
    void release(char *p)
    {
        const size_t off = (size_t)(p - base_);
        auto u = used_.find(off);
        if (u == used_.end()) return;
        size_t len = u->second;
        used_.erase(u);
        auto next = free_.lower_bound(off);
        if (next != free_.end() && next->first == off + len) {
            len += next->second;
            free_.erase(next);
        }
        auto it = free_.emplace(off, len).first;
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == off) {
                prev->second += len;
                free_.erase(it);
            }
        }
    }
    size_t size() const { return size_; }

  private:
    char *base_ = nullptr;
    size_t size_ = 0;
    std::map<size_t, size_t> free_, used_;
};

struct Request {
    uint64_t id = 0;
    trec_geom g{};
    trec_method m{};
    std::vector<float> angles;
    const float *stack_host = nullptr;
    float *volume_out = nullptr;
    size_t footprint = 0;
    double submit_s = 0, admit_s = 0, done_s = 0;
    trec_status status = TREC_OK;
    bool done = false;
    char *block = nullptr;  // arena block: stack | volume | workspace
    JobPlan plan;
};

}  // namespace

struct trec_server {
    int device = 0;
    int max_active = 8;
    trec_context *ctx = nullptr;
    hipStream_t stream = nullptr;
    Arena arena;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    uint64_t next_id = 1;
    std::deque<Request *> queue;
    std::map<uint64_t, std::unique_ptr<Request>> reqs;
    std::thread loop;

    void run();
    trec_status admit(Request &r);
};

trec_status trec_server::admit(Request &r)
{
    const size_t X = r.g.nx, Y = r.g.ny, T = r.g.thickness, P = r.g.nproj;
    float *dstack = (float *)r.block;
    float *dvol = (float *)(r.block + align_up(4 * P * Y * X));
    char *ws = r.block + align_up(4 * P * Y * X) + align_up(4 * X * Y * T);
    trec_status s = plan_job(ctx, r.plan, r.g, r.m, r.angles.data(), dstack, nullptr, dvol);
    if (s != TREC_OK) return s;
    carve(r.plan, ws);
    if (hipMemcpyAsync(dstack, r.stack_host, 4 * P * Y * X, hipMemcpyHostToDevice, stream) != hipSuccess)
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
    for (;;) {
        std::vector<Request *> admitted;
        {
            std::unique_lock<std::mutex> lk(mu);
            if (active.empty()) cv_work.wait(lk, [&] { return stop || !queue.empty(); });
            if (stop && queue.empty() && active.empty()) return;
            while (!queue.empty() && (int)active.size() + (int)admitted.size() < max_active) {
                Request *r = queue.front();
                r->block = arena.alloc(r->footprint);
                if (!r->block) break;  // first come, first served: wait for memory
                queue.pop_front();
                r->admit_s = now_s();
                admitted.push_back(r);
            }
        }
        for (Request *r : admitted) {
            r->status = admit(*r);
            if (r->status == TREC_OK) {
                active.push_back(r);
                stage.push_back(0);
            } else {
                (void)hipStreamSynchronize(stream);
                std::lock_guard<std::mutex> lk(mu);
                arena.release(r->block);
                r->done = true;
                r->done_s = now_s();
                cv_done.notify_all();
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
            if (e == hipSuccess)
                e = hipMemcpyAsync(r.volume_out, r.plan.volume, 4 * (size_t)r.g.nx * r.g.ny * r.g.thickness,
                                   hipMemcpyDeviceToHost, stream);
            finished.push_back(k);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) {
            // a device failure ends every active request with an error
            (void)hipGetLastError();
            finished.clear();
            for (size_t k = 0; k < active.size(); ++k) {
                active[k]->status = TREC_ERR_DEVICE;
                finished.push_back(k);
            }
        }
        if (finished.empty()) continue;
        {
            std::lock_guard<std::mutex> lk(mu);
            const double t = now_s();
            for (size_t k : finished) {
                Request &r = *active[k];
                arena.release(r.block);
                r.done = true;
                r.done_s = t;
            }
            cv_done.notify_all();
            cv_work.notify_all();
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
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        trec_context_destroy(s->ctx);
        return TREC_ERR_DEVICE;
    }
    if (hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            (void)hipStreamDestroy(s->stream);
            trec_context_destroy(s->ctx);
            return TREC_ERR_DEVICE;
        }
        hbm_budget = fr / 10 * 9;
    }
    if (!s->arena.init(align_up(hbm_budget))) {
        (void)hipStreamDestroy(s->stream);
        trec_context_destroy(s->ctx);
        return TREC_ERR_NOMEM;
    }
    trec_server *raw = s.release();
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
    r->footprint = request_bytes(req->geom, req->method);
    if (r->footprint == 0) return TREC_ERR_INVALID;
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
    (void)hipStreamSynchronize(srv->stream);
    (void)hipStreamDestroy(srv->stream);
    trec_context_destroy(srv->ctx);
    delete srv;
}

}  // extern "C"
