// Serving form: a request loop that runs community-detection jobs (generate a seeded graph
// on the device, then the whole phase loop of main.cpp:242-513) on several slots at once.
// Each slot is a worker thread with its own context, hence its own HIP stream, so the jobs
// of different slots overlap on the device.  The oldest queued request is admitted first,
// as soon as a slot is free and its HBM reservation fits beside the reservations of the
// running jobs; a job cannot allocate beyond its reservation (its context budget), so the
// running jobs together never exceed the server's budget.
//
// Small requests are batched: a slot that admits a small plain-Louvain request also takes
// the other queued small requests with the same run parameters, oldest first, wherever
// they wait in the queue (up to a batch size, and while their reservation fits), and runs
// them as one batched job (vt_batch_*), whose per-graph results equal vt_run's on each
// graph alone.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "vt_internal.hpp"

namespace {

struct Job {
    int64_t ticket = 0;
    bool claimed = false;  // a vt_server_wait is blocked on it
    vt_graph* g = nullptr;
    vt_request req{};
    uint64_t reserve = 0;
    bool done = false;
    vt_run_result res{};
    vt_request_info info{};
    std::string err;
};

}  // namespace

struct vt_server {
    int device = 0;
    uint64_t budget = 0;
    std::vector<vt_ctx*> ctxs;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job*> queue;
    std::map<int64_t, std::unique_ptr<Job>> jobs;
    int64_t next_ticket = 1;
    uint64_t reserved = 0;
    int32_t max_batch = 32;          // requests per batched job (1: no batching)
    int64_t batch_max_n = 1 << 16;   // largest graph a batched job takes
    bool stop = false;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace vt {
namespace {

// Expected directed entries of a request's graph: RGG with radius sqrt(2 ln n / (pi n)) has
// mean degree about 2 ln n (less near the border), plus the extra-edge share; BA has 2 n m.
double expected_entries(const vt_request& r)
{
    const double n = (double)r.n;
    if (r.kind == VT_REQ_BA) return 2.0 * n * r.param;
    return n > 1 ? 2.0 * n * std::log(n) * (1.0 + r.param / 100.0) : 0.0;
}

// Plain Louvain with no coloring or early termination, small enough to batch.
bool batchable(const vt_server* s, const vt_request& r)
{
    const vt_run_params& p = r.run;
    return s->max_batch > 1 && r.n <= s->batch_max_n && p.coloring == 0 && p.vertex_order == 0 && p.et_type == 0 &&
           !(p.one_phase && p.threshold_cycling);
}

// Run parameters a batch shares (the rest are unused for batchable requests).
bool same_run(const vt_run_params& a, const vt_run_params& b)
{
    return a.threshold == b.threshold && a.threshold_cycling == b.threshold_cycling && a.one_phase == b.one_phase &&
           a.flags == b.flags;
}

int gen_graph(vt_ctx* ctx, Job* j)
{
    const int rc = j->req.kind == VT_REQ_BA ? vt_gen_ba(ctx, j->req.seed, j->req.n, (int64_t)j->req.param, &j->g)
                                            : vt_gen_rgg(ctx, j->req.seed, j->req.n, j->req.param, &j->g);
    if (rc == VT_OK) vt_graph_info(j->g, &j->info.nv, &j->info.ne);
    return rc;
}

// Several requests as one batched job: every graph generated on the slot, the ones with
// edges run together through vt_batch_run (an edgeless graph goes through vt_run alone).
void run_group(vt_server* s, int slot, const std::vector<Job*>& grp, uint64_t reserve)
{
    vt_ctx* ctx = s->ctxs[slot];
    ctx->budget = reserve;
    ctx->peak = ctx->in_use;
    std::vector<Job*> members;
    for (Job* j : grp) {
        const int rc = gen_graph(ctx, j);
        if (rc != VT_OK) {
            j->info.status = rc;
            j->err = vt_last_error(ctx);
        } else if (j->info.ne > 0) {
            members.push_back(j);
        } else {
            j->info.status = vt_run(ctx, j->g, &j->req.run, &j->res);
            if (j->info.status != VT_OK) j->err = vt_last_error(ctx);
        }
    }
    if (!members.empty()) {
        std::vector<vt_graph*> gs;
        for (Job* j : members) gs.push_back(j->g);
        std::vector<vt_run_result> res(members.size());
        vt_batch* b = nullptr;
        int rc = vt_batch_create(ctx, (int32_t)gs.size(), gs.data(), &b);
        if (rc == VT_OK) rc = vt_batch_run(ctx, b, &members[0]->req.run, res.data(), nullptr);
        const std::string err = rc == VT_OK ? std::string() : std::string(vt_last_error(ctx));
        if (b) vt_batch_free(ctx, b);
        for (size_t i = 0; i < members.size(); i++) {
            members[i]->info.status = rc;
            members[i]->err = err;
            if (rc == VT_OK) members[i]->res = res[i];
        }
    }
    for (Job* j : grp) {
        if (j->g) vt_graph_free(ctx, j->g);
        j->g = nullptr;
    }
    release_scratch(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    for (Job* j : grp) {
        j->info.peak_bytes = ctx->peak;  // the batched job's high-water mark
        j->info.batch = (int32_t)grp.size();
    }
    ctx->budget = 0;
}

void run_job(vt_server* s, int slot, Job* j)
{
    vt_ctx* ctx = s->ctxs[slot];
    ctx->budget = j->reserve;
    ctx->peak = ctx->in_use;
    vt_graph* g = nullptr;
    int rc = j->req.kind == VT_REQ_BA ? vt_gen_ba(ctx, j->req.seed, j->req.n, (int64_t)j->req.param, &g)
                                      : vt_gen_rgg(ctx, j->req.seed, j->req.n, j->req.param, &g);
    if (rc == VT_OK) {
        vt_graph_info(g, &j->info.nv, &j->info.ne);
        rc = vt_run(ctx, g, &j->req.run, &j->res);
    }
    if (rc != VT_OK) j->err = vt_last_error(ctx);
    if (g) vt_graph_free(ctx, g);
    // nothing of this job stays on the slot: its scratch goes back before the next admission
    release_scratch(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    j->info.status = rc;
    j->info.peak_bytes = ctx->peak;
    j->info.batch = 1;
    ctx->budget = 0;
}

void worker(vt_server* s, int slot)
{
    (void)hipSetDevice(s->device);
    for (;;) {
        std::vector<Job*> grp;
        uint64_t reserve = 0;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] {
                return (s->stop && s->queue.empty()) ||
                       (!s->queue.empty() && s->reserved + s->queue.front()->reserve <= s->budget);
            });
            if (s->queue.empty()) return;  // stopping and drained
            Job* j = s->queue.front();
            s->queue.pop_front();
            grp.push_back(j);
            reserve = j->reserve;
            if (batchable(s, j->req)) {
                // the queued requests that match, oldest first, while they fit: a batched job
                // holds its members' graphs and their union, so it reserves twice their sum.
                // Requests that do not match keep their places, so a large request ahead of
                // small ones no longer holds them back from this job.
                uint64_t sum = j->reserve;
                for (auto it = s->queue.begin(); it != s->queue.end() && (int32_t)grp.size() < s->max_batch;) {
                    Job* k = *it;
                    if (!batchable(s, k->req) || !same_run(k->req.run, j->req.run)) {
                        ++it;
                        continue;
                    }
                    if (s->reserved + 2 * (sum + k->reserve) > s->budget) break;
                    it = s->queue.erase(it);
                    grp.push_back(k);
                    sum += k->reserve;
                }
                if (grp.size() > 1) reserve = 2 * sum;
            }
            s->reserved += reserve;
            const double t = s->now();
            for (Job* k : grp) {
                k->info.start_s = t;
                k->info.slot = slot;
                k->info.reserved_bytes = reserve;
            }
        }
        s->cv_work.notify_all();  // the next request may fit too
        if (grp.size() == 1) run_job(s, slot, grp[0]);
        else run_group(s, slot, grp, reserve);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->reserved -= reserve;
            const double t = s->now();
            for (Job* k : grp) {
                k->info.end_s = t;
                k->done = true;
            }
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

}  // namespace
}  // namespace vt

using namespace vt;

namespace {
// Errors of the vt_server_* calls, per calling thread: submit and wait may run on several
// host threads at once.
thread_local std::string tl_server_err;
}  // namespace

extern "C" {

uint64_t vt_request_reservation(const vt_request* r)
{
    if (!r || r->n < 1) return 0;
    // CSR (16 B per entry + 8 per vertex) held by the caller graph and a level graph, the
    // rebuild's key/weight double buffers (32 B per entry) and sort scratch, generator
    // buffers, the phase state (about 200 B per vertex), with a 25% margin on the entry count.
    const double e = 1.25 * expected_entries(*r) + 1024.0;
    return (uint64_t)(112.0 * e + 320.0 * (double)r->n) + (64ull << 20);
}

int vt_server_create(int device, uint64_t hbm_budget_bytes, int32_t slots, vt_server** out)
{
    if (!out || slots < 1 || slots > 16 || hbm_budget_bytes == 0) return VT_ERR_ARG;
    *out = nullptr;
    vt_server* s = new vt_server();
    s->device = device;
    s->budget = hbm_budget_bytes;
    for (int i = 0; i < slots; i++) {
        vt_ctx* c = nullptr;
        const int rc = vt_ctx_create(device, 0, &c);
        if (rc != VT_OK) {
            for (vt_ctx* x : s->ctxs) vt_ctx_destroy(x);
            delete s;
            return rc;
        }
        s->ctxs.push_back(c);
    }
    for (int i = 0; i < slots; i++) s->workers.emplace_back(worker, s, i);
    *out = s;
    return VT_OK;
}

int vt_server_set_batching(vt_server* s, int32_t max_batch, int64_t max_n)
{
    if (!s || max_batch < 1 || max_batch > 4096 || max_n < 1) return VT_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->max_batch = max_batch;
    s->batch_max_n = max_n;
    return VT_OK;
}

int vt_server_submit(vt_server* s, const vt_request* req, int64_t* ticket)
{
    if (!s || !req || !ticket) return VT_ERR_ARG;
    if (req->n < 1 || (req->kind != VT_REQ_RGG && req->kind != VT_REQ_BA) ||
        (req->kind == VT_REQ_BA && req->param < 1.0)) {
        tl_server_err = "vt_server_submit: bad request";
        return VT_ERR_ARG;
    }
    auto j = std::make_unique<Job>();
    j->req = *req;
    j->reserve = vt_request_reservation(req);
    if (j->reserve > s->budget) {
        tl_server_err = "vt_server_submit: request needs " + std::to_string(j->reserve) + " bytes, more than the budget";
        return VT_ERR_BUDGET;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) {
            tl_server_err = "vt_server_submit: the server is stopping";
            return VT_ERR_STATE;
        }
        j->ticket = s->next_ticket++;
        j->info.submit_s = s->now();
        j->info.reserved_bytes = j->reserve;
        *ticket = j->ticket;
        s->queue.push_back(j.get());
        s->jobs[j->ticket] = std::move(j);
    }
    s->cv_work.notify_all();
    return VT_OK;
}

int vt_server_wait(vt_server* s, int64_t ticket, vt_run_result* res, vt_request_info* info)
{
    if (!s) return VT_ERR_ARG;
    std::unique_ptr<Job> j;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->jobs.find(ticket);
        if (it == s->jobs.end()) {
            tl_server_err = "vt_server_wait: unknown or already collected ticket";
            return VT_ERR_ARG;
        }
        Job* p = it->second.get();
        if (p->claimed) {
            tl_server_err = "vt_server_wait: another thread is already waiting on this ticket";
            return VT_ERR_STATE;
        }
        p->claimed = true;
        s->cv_done.wait(lk, [&] { return p->done; });
        // std::map iterators stay valid while other tickets come and go, and only the
        // thread that claimed this ticket erases it
        j = std::move(it->second);
        s->jobs.erase(it);
    }
    if (res) *res = j->res;
    if (info) *info = j->info;
    if (j->info.status != VT_OK) tl_server_err = j->err;
    return j->info.status;
}

const char* vt_server_last_error(const vt_server* s)
{
    return s ? tl_server_err.c_str() : "null server";
}

int vt_server_destroy(vt_server* s)
{
    if (!s) return VT_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    for (vt_ctx* c : s->ctxs) vt_ctx_destroy(c);
    delete s;
    return VT_OK;
}

}  // extern "C"
