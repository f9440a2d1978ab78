// Request loop over HBM-resident geometries: continuous batching of SAL-gradient requests.
//
// Each registered geometry keeps its tree, interaction list, mixed-precision plan and cell areas in
// device memory.  A request carries R fields (host arrays, caller's point order) for one geometry.
// The server runs `slots` evaluations at once, each on its own stream with its own workspace and
// worker thread, so requests that arrive while others are on the GPU overlap with them.  When a
// slot frees up it takes the oldest queued request and merges the queued requests of the same
// geometry behind it into one multi-field evaluation (up to max_fields fields), as long as the
// device memory it needs fits the budget: held bytes of all slots plus this slot's growth must stay
// within hbm_budget, and the growth must fit the device's free memory minus hbm_reserve.  A request
// that cannot fit even with every other slot empty is rejected with SAL_ERR_LIMIT.
//
// A field's result does not depend on the fields it is batched with: every kernel computes each
// field's column with the same operation order whatever the other columns hold.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>

#include "common.h"
#include "plan.h"

namespace sal {
namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

enum ReqState { RQ_QUEUED = 0, RQ_RUNNING = 1, RQ_DONE = 2 };

struct Req {
    int gid = 0, R = 0;
    const double* ssh = nullptr;
    double *gx = nullptr, *gy = nullptr;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int state = RQ_QUEUED, status = SAL_OK;
    int64_t batch = -1;
    int batch_fields = 0;
    std::string err;
};

struct Geo {
    sal_geom g = nullptr;
    sal_eval_opts o{};
    DBuf<double> area;
};

struct Slot {
    hipStream_t s = nullptr;
    sal_work w = nullptr;
    DBuf<double> ssh, gx, gy;
    bool busy = false;
    int64_t held = 0;  // device bytes of the slot's buffers (or reserved for the batch it is starting)
    std::thread th;
    int64_t io_bytes() const { return (int64_t)(ssh.bytes() + gx.bytes() + gy.bytes()); }
};

}  // namespace
}  // namespace sal

struct sal_server_s {
    sal_server_opts opts{};
    std::vector<std::unique_ptr<sal::Geo>> geos;
    std::vector<std::unique_ptr<sal::Slot>> slots;
    std::deque<sal::Req> reqs;  // by id (a deque: references stay valid as requests are added)
    std::deque<int64_t> queue;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    int64_t budget = 0;
    double t0 = 0;
    int64_t n_batches = 0, n_done = 0, n_rejected = 0, n_merged = 0, peak_held = 0;
    int busy_now = 0, busy_peak = 0;
    int alive = 0;  // worker threads still running

    int64_t held_total() const {
        int64_t h = 0;
        for (auto& s : slots) h += s->held;
        return h;
    }
    void worker(sal::Slot* sl);
    bool pick(sal::Slot* sl, std::vector<int64_t>& batch, int& F);
    void run(sal::Slot* sl, const sal::Geo& G, const std::vector<int64_t>& batch, const std::vector<sal::Req>& part,
             int F);
};

namespace sal {
namespace {

// Bytes a slot holds after growing its buffers for F fields of geometry G.
int64_t slot_after(sal_server s, const Slot* sl, const Geo& G, int F) {
    const size_t nf = (size_t)G.g->n * F;
    int64_t io = 0;
    for (const DBuf<double>* b : {&sl->ssh, &sl->gx, &sl->gy}) io += 8 * (int64_t)std::max(b->n, nf);
    (void)s;
    return io + work_bytes_after(sl->w, G.g, G.o, F);
}

}  // namespace
}  // namespace sal

using namespace sal;

// Under the lock: choose the batch for idle slot `sl`, or return false to wait.
bool sal_server_s::pick(Slot* sl, std::vector<int64_t>& batch, int& F) {
    while (!queue.empty()) {
        const int64_t h = queue.front();
        Req& hr = reqs[h];
        const Geo& G = *geos[hr.gid];
        // candidate batch: the head plus queued requests of its geometry, FIFO, within max_fields
        std::vector<int64_t> cand{h};
        int f = hr.R;
        if (opts.batching) {
            for (size_t i = 1; i < queue.size(); ++i) {
                const Req& r = reqs[queue[i]];
                if (r.gid == hr.gid && f + r.R <= opts.max_fields) {
                    cand.push_back(queue[i]);
                    f += r.R;
                }
            }
        }
        const int64_t others = held_total() - sl->held;
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) { (void)hipGetLastError(); free_b = 0; }
        auto fits = [&](int fields) {
            const int64_t after = slot_after(this, sl, G, fields);
            const int64_t growth = after - sl->held;
            return others + after <= budget && growth <= (int64_t)free_b - opts.hbm_reserve;
        };
        while (cand.size() > 1 && !fits(f)) {  // drop the youngest merged requests first
            f -= reqs[cand.back()].R;
            cand.pop_back();
        }
        if (fits(f)) {
            for (int64_t id : cand) queue.erase(std::find(queue.begin(), queue.end(), id));
            batch = cand;
            F = f;
            sl->held = std::max(sl->held, slot_after(this, sl, G, f));
            return true;
        }
        // the head alone does not fit now: can it ever?  (every other slot emptied, this slot fresh)
        const int64_t alone = 8 * 3 * (int64_t)G.g->n * hr.R + mixed_work_need(G.g, G.o, hr.R);
        if (alone > budget) {
            queue.pop_front();
            hr.state = RQ_DONE;
            hr.status = SAL_ERR_LIMIT;
            hr.err = "request needs " + std::to_string(alone) + " device bytes, over the server budget of " +
                     std::to_string(budget);
            hr.t_done = now_s() - t0;
            ++n_rejected;
            cv_done.notify_all();
            continue;
        }
        // free idle slots' cached buffers (this one's too) and retry once; otherwise wait for a slot
        bool released = false;
        for (auto& o : slots) {
            if (o.get() != sl && o->busy) continue;
            if (o->held == 0) continue;
            o->ssh.release(); o->gx.release(); o->gy.release();
            work_release(o->w);
            o->held = 0;
            released = true;
        }
        if (released) continue;
        if (busy_now == 0) {  // nothing will free up: the device itself lacks the memory
            queue.pop_front();
            hr.state = RQ_DONE;
            hr.status = SAL_ERR_NOMEM;
            hr.err = "request does not fit the device's free memory";
            hr.t_done = now_s() - t0;
            ++n_rejected;
            cv_done.notify_all();
            continue;
        }
        return false;
    }
    return false;
}

// Outside the lock: `part` holds copies of the batch's requests.
void sal_server_s::run(Slot* sl, const Geo& G, const std::vector<int64_t>& batch, const std::vector<Req>& part,
                       int F) {
    const int64_t n = G.g->n;
    int status = SAL_OK;
    std::string err;
    try {
        sl->ssh.ensure((size_t)n * F);
        sl->gx.ensure((size_t)n * F);
        sl->gy.ensure((size_t)n * F);
        int64_t off = 0;
        for (const Req& r : part) {
            SAL_HIP(hipMemcpyAsync(sl->ssh.p + off * n, r.ssh, sizeof(double) * n * r.R, hipMemcpyHostToDevice, sl->s));
            off += r.R;
        }
        eval_mixed_work(G.g, G.o, sl->w, G.area.p, sl->ssh.p, F, sl->gx.p, sl->gy.p, 0, sl->s);
        off = 0;
        for (const Req& r : part) {
            SAL_HIP(hipMemcpyAsync(r.gx, sl->gx.p + off * n, sizeof(double) * n * r.R, hipMemcpyDeviceToHost, sl->s));
            SAL_HIP(hipMemcpyAsync(r.gy, sl->gy.p + off * n, sizeof(double) * n * r.R, hipMemcpyDeviceToHost, sl->s));
            off += r.R;
        }
        SAL_HIP(hipStreamSynchronize(sl->s));
    } catch (const Error& e) {
        status = e.code;
        err = e.msg;
        (void)hipStreamSynchronize(sl->s);
    } catch (const std::bad_alloc&) {
        status = SAL_ERR_NOMEM;
        err = "host allocation failed";
    }
    std::lock_guard<std::mutex> lk(mu);
    const double t = now_s() - t0;
    for (int64_t id : batch) {
        Req& r = reqs[id];
        r.state = RQ_DONE;
        r.status = status;
        r.err = err;
        r.t_done = t;
    }
    sl->held = sl->io_bytes() + work_held(sl->w);
    sl->busy = false;
    --busy_now;
    n_done += (int64_t)batch.size();
    cv_done.notify_all();
    cv_work.notify_all();
}

void sal_server_s::worker(Slot* sl) {
    SAL_HIP(hipSetDevice(opts.device));
    for (;;) {
        std::vector<int64_t> batch;
        std::vector<Req> part;
        const Geo* G = nullptr;
        int F = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || pick(sl, batch, F); });
            if (batch.empty()) return;  // stopping with nothing picked
            const double t = now_s() - t0;
            const int64_t b = n_batches++;
            if (batch.size() > 1) n_merged += (int64_t)batch.size();
            for (int64_t id : batch) {
                Req& r = reqs[id];
                r.state = RQ_RUNNING;
                r.t_admit = t;
                r.batch = b;
                r.batch_fields = F;
                part.push_back(r);
            }
            G = geos[part[0].gid].get();
            sl->busy = true;
            busy_peak = std::max(busy_peak, ++busy_now);
            peak_held = std::max(peak_held, held_total());
        }
        run(sl, *G, batch, part, F);
    }
}

extern "C" {

int sal_server_opts_default(sal_server_opts* o) {
    return guard([&] {
        SAL_REQUIRE(o, SAL_ERR_ARG, "null pointer");
        o->slots = 3;
        o->max_fields = 32;
        o->batching = 1;
        o->graph = 0;
        o->device = 0;
        o->hbm_budget = 0;
        o->hbm_reserve = (int64_t)4 << 30;
    });
}

int sal_server_create(const sal_server_opts* o, sal_server* out) {
    return guard([&] {
        SAL_REQUIRE(o && out, SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(o->slots >= 1 && o->slots <= 8, SAL_ERR_ARG, "slots must be in [1, 8]");
        SAL_REQUIRE(o->max_fields >= 1 && o->max_fields <= 256, SAL_ERR_ARG, "max_fields must be in [1, 256]");
        SAL_REQUIRE(o->hbm_budget >= 0 && o->hbm_reserve >= 0, SAL_ERR_ARG, "negative memory budget");
        SAL_HIP(hipSetDevice(o->device));
        auto s = std::make_unique<sal_server_s>();
        s->opts = *o;
        size_t free_b = 0, total_b = 0;
        SAL_HIP(hipMemGetInfo(&free_b, &total_b));
        const int64_t avail = std::max<int64_t>(0, (int64_t)free_b - o->hbm_reserve);
        s->budget = o->hbm_budget > 0 ? std::min<int64_t>(o->hbm_budget, avail) : avail;
        s->t0 = now_s();
        for (int i = 0; i < o->slots; ++i) {
            auto sl = std::make_unique<Slot>();
            SAL_HIP(hipStreamCreateWithFlags(&sl->s, hipStreamNonBlocking));
            sl->w = work_new();
            s->slots.push_back(std::move(sl));
        }
        s->alive = (int)s->slots.size();
        for (auto& sl : s->slots) sl->th = std::thread([srv = s.get(), p = sl.get()] {
            try { srv->worker(p); } catch (...) {}
            std::lock_guard<std::mutex> lk(srv->mu);
            --srv->alive;
            srv->cv_done.notify_all();
        });
        *out = s.release();
    });
}

int sal_server_add_geometry(sal_server s, sal_geom g, const double* area, const sal_eval_opts* o, int* gid) {
    return guard([&] {
        SAL_REQUIRE(s && g && area && o && gid, SAL_ERR_ARG, "null pointer");
        SAL_REQUIRE(o->precision == SAL_PREC_MIXED, SAL_ERR_ARG, "the server evaluates in SAL_PREC_MIXED");
        auto G = std::make_unique<Geo>();
        G->g = g;
        G->o = *o;
        G->o.graph = s->opts.graph;
        G->area.alloc((size_t)g->n);
        h2d(G->area.p, area, (size_t)g->n, g->stream);
        SAL_HIP(hipStreamSynchronize(g->stream));
        (void)mixed_work_need(g, G->o, 1);  // builds the plan and work lists now, not under load
        std::lock_guard<std::mutex> lk(s->mu);
        *gid = (int)s->geos.size();
        s->geos.push_back(std::move(G));
    });
}

int sal_server_submit(sal_server s, int gid, int R, const double* ssh, double* gx, double* gy, int64_t* req) {
    return guard([&] {
        SAL_REQUIRE(s && ssh && gx && gy && req, SAL_ERR_ARG, "null pointer");
        std::lock_guard<std::mutex> lk(s->mu);
        SAL_REQUIRE(gid >= 0 && gid < (int)s->geos.size(), SAL_ERR_ARG, "unknown geometry");
        SAL_REQUIRE(R >= 1 && R <= s->opts.max_fields, SAL_ERR_ARG, "R must be in [1, max_fields]");
        SAL_REQUIRE(!s->stop, SAL_ERR_STATE, "server is stopping");
        Req r;
        r.gid = gid;
        r.R = R;
        r.ssh = ssh;
        r.gx = gx;
        r.gy = gy;
        r.t_submit = now_s() - s->t0;
        *req = (int64_t)s->reqs.size();
        s->reqs.push_back(std::move(r));
        s->queue.push_back(*req);
        s->cv_work.notify_all();
    });
}

int sal_server_wait(sal_server s, int64_t req, double timeout_s, int* done) {
    return guard([&] {
        SAL_REQUIRE(s && done, SAL_ERR_ARG, "null pointer");
        std::unique_lock<std::mutex> lk(s->mu);
        SAL_REQUIRE(req >= 0 && req < (int64_t)s->reqs.size(), SAL_ERR_ARG, "unknown request");
        auto ready = [&] { return s->reqs[req].state == RQ_DONE; };
        if (timeout_s < 0) s->cv_done.wait(lk, ready);
        else s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready);
        *done = ready() ? 1 : 0;
        if (*done && s->reqs[req].status != SAL_OK)
            throw Error{s->reqs[req].status, "request " + std::to_string(req) + ": " + s->reqs[req].err};
    });
}

int sal_server_request_info(sal_server s, int64_t req, double t[3], int64_t* batch, int* batch_fields) {
    return guard([&] {
        SAL_REQUIRE(s && t && batch && batch_fields, SAL_ERR_ARG, "null pointer");
        std::lock_guard<std::mutex> lk(s->mu);
        SAL_REQUIRE(req >= 0 && req < (int64_t)s->reqs.size(), SAL_ERR_ARG, "unknown request");
        const Req& r = s->reqs[req];
        t[0] = r.t_submit;
        t[1] = r.t_admit;
        t[2] = r.t_done;
        *batch = r.batch;
        *batch_fields = r.batch_fields;
    });
}

int sal_server_stats(sal_server s, int64_t st[8]) {
    return guard([&] {
        SAL_REQUIRE(s && st, SAL_ERR_ARG, "null pointer");
        std::lock_guard<std::mutex> lk(s->mu);
        st[0] = (int64_t)s->reqs.size();
        st[1] = s->n_done;
        st[2] = s->n_rejected;
        st[3] = s->n_batches;
        st[4] = s->n_merged;
        st[5] = s->busy_peak;
        st[6] = s->peak_held;
        st[7] = s->budget;
    });
}

int sal_server_destroy(sal_server s) {
    return guard([&] {
        if (!s) return;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            // finish queued work first
            s->cv_done.wait(lk, [&] { return (s->queue.empty() && s->busy_now == 0) || s->alive == 0; });
            s->stop = true;
            s->cv_work.notify_all();
        }
        for (auto& sl : s->slots)
            if (sl->th.joinable()) sl->th.join();
        for (auto& sl : s->slots) {
            work_delete(sl->w);
            if (sl->s) (void)hipStreamDestroy(sl->s);
        }
        delete s;
    });
}

}  // extern "C"
