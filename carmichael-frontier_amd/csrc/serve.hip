// serve.hip -- a device-resident serving loop for modulus solves (continuous batching).
//
// The reference solves instances one after another from Python (gpu/run_mstar_family.py loops over
// the family's rows and radii and calls the driver's per-instance routine, which launches, waits
// and reads back). Here requests are submitted as they arrive and served by `workers` threads, each
// with its own context (HIP stream and device pool) on the one GPU:
//   * a free worker takes every request that has arrived (up to max_batch) and solves them as ONE
//     batch (one key space, one set of launches per radius), so requests that arrive while others
//     are on the GPU join the next batch instead of waiting behind it;
//   * before its solve a batch is priced (cf_batch_footprint) and admitted only while the priced
//     footprints of the batches on the device stay within the memory budget; a batch larger than
//     the whole budget runs alone;
//   * the workers' streams run concurrently, so one batch's kernels overlap another's.
// A request's result (records per side, every pair) is the one it gets when solved alone: batches
// keep instances in disjoint key ranges.
#include <chrono>
#include <cstdio>
#include <condition_variable>
#include <deque>
#include <set>
#include <thread>

#include "cf_internal.h"

namespace {

double now_us(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
}

struct Request {
    cf_instance inst;
    std::vector<u64> B, INS;   // owned copies of the instance's prime lists
    u64 tag;
    double t_submit;
    u64 key_M;                 // the modulus of its keys: a batch's key moduli must sum below 2^64 - 1
};

struct Served {
    u64 tag = 0;
    int status = 0;
    std::string error;
    int batch_size = 0;
    u64 nD = 0, nI = 0;
    std::vector<u64> pd, pi;
    double t_submit = 0, t_start = 0, t_done = 0;
};

}  // namespace

struct cf_server {
    int device = 0;
    int max_batch = 1;
    u64 budget = 0;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_budget;
    std::deque<Request> pending;
    std::deque<Served> done;
    std::map<u64, Served> held;        // results handed out, arrays alive until cf_server_release
    std::set<u64> tags;                // tags submitted and not yet released (unique)
    u64 in_use = 0;                    // priced footprint of the batches on the device
    int busy = 0;                      // batches on the device
    int outstanding = 0;               // submitted and not yet handed out
    bool stop = false;
    std::vector<std::thread> workers;
};

namespace {

void serve_batch(cf_server *s, cf_ctx *ctx, std::vector<Request> &reqs) {
    const double t_start = now_us(s->t0);
    std::vector<cf_instance> insts(reqs.size());
    for (size_t k = 0; k < reqs.size(); ++k) insts[k] = reqs[k].inst;
    std::vector<Served> out(reqs.size());
    int rc = CF_OK;
    std::string err;
    cf_batch *b = nullptr;
    u64 price = 0;
    bool admitted = false;
    rc = cf_batch_create(ctx, insts.data(), (int)insts.size(), &b);
    if (!rc) rc = cf_batch_footprint(b, &price, nullptr, nullptr);
    if (!rc) {
        // admission by footprint: wait until the batch fits beside the others (alone if it must)
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_budget.wait(lk, [&] { return s->busy == 0 || s->in_use + price <= s->budget; });
        s->in_use += price;
        s->busy += 1;
        admitted = true;
    }
    u64 *pd = nullptr, *pi = nullptr, *pj = nullptr, np = 0;
    std::vector<u64> counts(2 * reqs.size(), 0);
    if (!rc) rc = cf_batch_solve(b, 0, &pd, &pi, &pj, &np, counts.data(), nullptr);
    std::vector<u64> hd(np), hi(np), hj(np);
    if (!rc && np) {
        try {
            d2h(ctx, hd.data(), pd, np);
            d2h(ctx, hi.data(), pi, np);
            d2h(ctx, hj.data(), pj, np);
        } catch (const CfError &e) {
            rc = e.code;
            err = e.msg;
        }
    }
    if (rc && err.empty()) err = cf_last_error();
    if (pd) cf_free(ctx, pd);
    if (pi) cf_free(ctx, pi);
    if (pj) cf_free(ctx, pj);
    if (b) cf_batch_destroy(b);
    if (admitted) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->in_use -= price;
        s->busy -= 1;
    }
    s->cv_budget.notify_all();
    const double t_done = now_us(s->t0);
    for (size_t k = 0; k < reqs.size(); ++k) {
        Served &r = out[k];
        r.tag = reqs[k].tag;
        r.status = rc;
        r.error = err;
        r.batch_size = (int)reqs.size();
        r.t_submit = reqs[k].t_submit;
        r.t_start = t_start;
        r.t_done = t_done;
        if (!rc) {
            r.nD = counts[2 * k];
            r.nI = counts[2 * k + 1];
        }
    }
    for (u64 q = 0; q < hd.size(); ++q) {
        Served &r = out[hj[q]];
        r.pd.push_back(hd[q]);
        r.pi.push_back(hi[q]);
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (auto &r : out) s->done.push_back(std::move(r));
    }
    s->cv_done.notify_all();
}

void worker_main(cf_server *s) {
    cf_ctx *ctx = nullptr;
    const int rc = cf_ctx_create(s->device, &ctx);
    for (;;) {
        std::vector<Request> take;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
            if (s->pending.empty()) break;   // stopping and drained
            // one batch is one key space: its requests' key moduli must sum below 2^64 - 1
            unsigned __int128 keys = 0;
            while (!s->pending.empty() && (int)take.size() < s->max_batch) {
                keys += s->pending.front().key_M;
                if (!take.empty() && keys >= (unsigned __int128)~0ULL) break;
                take.push_back(std::move(s->pending.front()));
                s->pending.pop_front();
            }
        }
        if (rc) {   // no context on this worker: fail its requests rather than hang them
            std::lock_guard<std::mutex> lk(s->mu);
            for (auto &q : take) {
                Served r;
                r.tag = q.tag;
                r.status = rc;
                r.error = "worker context: " + std::string(cf_last_error());
                s->done.push_back(std::move(r));
            }
            s->cv_done.notify_all();
            continue;
        }
        serve_batch(s, ctx, take);
    }
    if (ctx) cf_ctx_destroy(ctx);
}

}  // namespace

extern "C" {

int cf_server_create(int device, int workers, int max_batch, uint64_t mem_budget, cf_server **out) {
    API_BEGIN
    CF_REQUIRE(out, "NULL argument");
    CF_REQUIRE(workers >= 1 && workers <= 16, "need 1 <= workers <= 16");
    CF_REQUIRE(max_batch >= 1, "need max_batch >= 1");
    std::unique_ptr<cf_server> s(new cf_server());
    s->device = device;
    s->max_batch = max_batch;
    s->t0 = std::chrono::steady_clock::now();
    if (mem_budget) {
        s->budget = mem_budget;
    } else {
        CF_HIP(hipSetDevice(device));
        size_t free_b = 0, total_b = 0;
        CF_HIP(hipMemGetInfo(&free_b, &total_b));
        s->budget = free_b / 2;
    }
    for (int w = 0; w < workers; ++w) s->workers.emplace_back(worker_main, s.get());
    *out = s.release();
    API_END
}

int cf_server_submit(cf_server *s, const cf_instance *inst, uint64_t tag) {
    API_BEGIN
    CF_REQUIRE(s && inst && inst->B && inst->INS, "NULL argument");
    const KeyMod km = key_mod(*inst);
    (void)spec_of_d(*inst, 0, &km);   // a malformed request fails here, not inside a batch of others
    (void)spec_of_i(*inst, 0, &km);
    Request q;
    q.key_M = km.key_M;
    q.inst = *inst;
    q.B.assign(inst->B, inst->B + inst->nB);
    q.INS.assign(inst->INS, inst->INS + inst->nI);
    q.inst.B = q.B.data();
    q.inst.INS = q.INS.data();
    q.tag = tag;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        CF_REQUIRE(!s->stop, "the server is stopping");
        // a held result is found by its tag: a second request under a tag in use would replace it
        // and free the arrays its caller still reads
        CF_REQUIRE(!s->tags.count(tag), "tag already in use by a request that was not released");
        s->tags.insert(tag);
        q.t_submit = now_us(s->t0);
        s->pending.push_back(std::move(q));
        s->pending.back().inst.B = s->pending.back().B.data();
        s->pending.back().inst.INS = s->pending.back().INS.data();
        s->outstanding += 1;
    }
    s->cv_work.notify_one();
    API_END
}

int cf_server_next(cf_server *s, int timeout_ms, cf_served *out) {
    API_BEGIN
    CF_REQUIRE(s && out, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    const bool got = s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms < 0 ? 0 : timeout_ms),
                                         [&] { return !s->done.empty(); });
    if (!got) return 0;
    Served r = std::move(s->done.front());
    s->done.pop_front();
    s->outstanding -= 1;
    Served &h = s->held[r.tag] = std::move(r);
    *out = cf_served{};
    out->tag = h.tag;
    out->status = h.status;
    out->batch_size = h.batch_size;
    out->nD = h.nD;
    out->nI = h.nI;
    out->npairs = h.pd.size();
    out->pairs_d = h.pd.data();
    out->pairs_i = h.pi.data();
    out->t_submit_us = h.t_submit;
    out->t_start_us = h.t_start;
    out->t_done_us = h.t_done;
    std::snprintf(out->error, sizeof(out->error), "%s", h.error.c_str());
    return 1;
    API_END
}

int cf_server_release(cf_server *s, uint64_t tag) {
    API_BEGIN
    CF_REQUIRE(s, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->held.erase(tag)) {
        s->tags.erase(tag);
    } else {
        // releasing a tag whose result was not handed out yet would let the tag be reused while
        // its request is still queued or on the device
        CF_REQUIRE(!s->tags.count(tag), "the tag's result has not been handed out by cf_server_next");
    }
    API_END
}

int cf_server_pending(cf_server *s, uint64_t *queued, uint64_t *outstanding) {
    API_BEGIN
    CF_REQUIRE(s, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (queued) *queued = s->pending.size();
    if (outstanding) *outstanding = (u64)s->outstanding;
    API_END
}

int cf_server_destroy(cf_server *s) {
    API_BEGIN
    if (!s) return CF_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    delete s;
    API_END
}

}  // extern "C"
