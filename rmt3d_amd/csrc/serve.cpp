// Serving loop for independent fluid-structure solves.
//
// Requests (a configuration, its objects, a step count) are queued in arrival
// order and admitted first-in first-out against a device-memory budget: the
// head of the queue starts when its estimated footprint fits beside the
// contexts already running.  Each admitted request gets its own context
// (fields resident in HBM for its whole life) and its own non-blocking
// stream, driven by one worker thread of a fixed pool, so the kernels of
// different requests overlap on the GPU while each request's own stage order
// is exactly the single-solve order (its results are the same alone and in
// the mix).  The final fields stay on the host until the caller releases the
// request.  The reference has no serving layer; this is the unit its
// stepper-per-process model runs many of.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "fluid3d.h"

namespace {

struct Request {
    int id = 0;
    rmt3d_config cfg{};
    std::vector<rmt3d_object> objs;
    int prec = RMT3D_PREC_FP64;
    int steps = 0;
    double est = 0, held = 0;
    int status = RMT3D_REQ_QUEUED;
    int rc = 0, steps_done = 0;
    double t_submit = 0, t_admit = -1, t_done = -1, sim_time = 0;
    long long cell_steps = 0;
    std::vector<double> u, v, w, p;
    std::vector<uint32_t> c;
    std::string err;
};

}  // namespace

struct rmt3d_server {
    double budget = 0;
    int max_active = 1;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request *> queue;
    std::map<int, Request *> all;
    double used = 0, peak_used = 0;
    int active = 0, peak_active = 0;
    bool stop = false;
    int next_id = 1;
    std::vector<std::thread> workers;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    void worker();
    void run(Request *r);
};

void rmt3d_server::worker() {
    for (;;) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return (stop && queue.empty()) || (!queue.empty() && used + queue.front()->est <= budget); });
            if (queue.empty()) return;
            r = queue.front();
            queue.pop_front();
            r->held = r->est;
            used += r->held;
            active++;
            peak_used = std::max(peak_used, used);
            peak_active = std::max(peak_active, active);
            r->status = RMT3D_REQ_RUNNING;
            r->t_admit = now();
        }
        run(r);
        {
            std::lock_guard<std::mutex> lk(mu);
            used -= r->held;
            active--;
            r->t_done = now();
            r->status = (r->rc == 0) ? RMT3D_REQ_DONE : RMT3D_REQ_FAILED;
        }
        cv.notify_all();
    }
}

void rmt3d_server::run(Request *r) {
    rmt::Sim *S = new (std::nothrow) rmt::Sim;
    if (!S) { r->rc = RMT3D_ERR_NOMEM; return; }
    try {
        int rc = S->create(&r->cfg, r->objs.data(), r->prec);
        {
            // account for what the context really holds
            std::lock_guard<std::mutex> lk(mu);
            used += (double)S->dev_bytes - r->held;
            r->held = (double)S->dev_bytes;
            peak_used = std::max(peak_used, used);
        }
        if (rc) { r->rc = rc; r->err = S->err; delete S; return; }
        int e = S->initialize();
        if (e < 0) { r->rc = e; r->err = S->err; delete S; return; }
        double dt;
        int total;
        rc = S->prepare_run(e, &dt, &total);
        if (rc) { r->rc = rc; r->err = S->err; delete S; return; }
        long long cells = (long long)r->cfg.sys_size[0] * r->cfg.sys_size[1] * r->cfg.sys_size[2];
        for (int s = 0; s < r->steps; s++) {
            rc = S->step(0);
            if (rc) { r->rc = rc; r->err = S->err; break; }
            std::lock_guard<std::mutex> lk(mu);
            r->steps_done++;
            r->cell_steps += cells;
        }
        size_t Ng = (size_t)S->g.Ng;
        r->u.resize(Ng); r->v.resize(Ng); r->w.resize(Ng); r->p.resize(Ng); r->c.resize(Ng);
        rmt3d_state st{};
        st.u = r->u.data(); st.v = r->v.data(); st.w = r->w.data(); st.p = r->p.data(); st.rm_c = r->c.data();
        int drc = S->download(&st);
        if (drc && !r->rc) r->rc = drc;
        r->sim_time = S->time;
    } catch (...) {
        r->rc = RMT3D_ERR_FATAL;
    }
    delete S;
}

extern "C" {

int rmt3d_estimate_bytes(const rmt3d_config *cfg, const rmt3d_object *objs, int precision, double *bytes) {
    if (!cfg || !bytes) return RMT3D_ERR_ARG;
    if (cfg->n_obj > 0 && !objs && cfg->sim_type != RMT3D_SIM_FTEST) return RMT3D_ERR_ARG;
    *bytes = rmt::Sim::estimate_bytes(*cfg, objs, precision);
    return 0;
}

int rmt3d_server_create(double hbm_budget_bytes, int max_active, rmt3d_server **out) {
    if (!out || hbm_budget_bytes <= 0 || max_active < 1 || max_active > 16) return RMT3D_ERR_ARG;
    rmt3d_server *S = new (std::nothrow) rmt3d_server;
    if (!S) return RMT3D_ERR_NOMEM;
    S->budget = hbm_budget_bytes;
    S->max_active = max_active;
    try {
        for (int i = 0; i < max_active; i++) S->workers.emplace_back([S] { S->worker(); });
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->stop = true;
        }
        S->cv.notify_all();
        for (auto &t : S->workers) t.join();
        delete S;
        return RMT3D_ERR_FATAL;
    }
    *out = S;
    return 0;
}

int rmt3d_server_submit(rmt3d_server *S, const rmt3d_config *cfg, const rmt3d_object *objs, int precision,
                        int steps, int *req) {
    if (!S || !cfg || !req || steps < 0) return RMT3D_ERR_ARG;
    if (precision != RMT3D_PREC_FP64 && precision != RMT3D_PREC_MIXED) return RMT3D_ERR_ARG;
    int n_obj = (cfg->sim_type == RMT3D_SIM_FTEST) ? 0 : cfg->n_obj;
    if (n_obj > 0 && !objs) return RMT3D_ERR_ARG;
    Request *r = new (std::nothrow) Request;
    if (!r) return RMT3D_ERR_NOMEM;
    r->cfg = *cfg;
    if (n_obj > 0) r->objs.assign(objs, objs + n_obj);
    r->prec = precision;
    r->steps = steps;
    r->est = rmt::Sim::estimate_bytes(*cfg, r->objs.data(), precision);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (r->est > S->budget || S->stop) {
            delete r;
            return RMT3D_ERR_ARG;
        }
        r->id = S->next_id++;
        r->t_submit = S->now();
        S->all[r->id] = r;
        S->queue.push_back(r);
        *req = r->id;
    }
    S->cv.notify_all();
    return 0;
}

// rc, sim_time and the result vectors are written by the worker while the
// request runs and published by the status change under the lock; they are
// read only once the request has ended
static void fill_info(const Request *r, rmt3d_request_info *info) {
    bool ended = r->status == RMT3D_REQ_DONE || r->status == RMT3D_REQ_FAILED;
    info->status = r->status;
    info->rc = ended ? r->rc : 0;
    info->steps_done = r->steps_done;
    info->cell_steps = r->cell_steps;
    info->t_submit = r->t_submit;
    info->t_admit = r->t_admit;
    info->t_done = r->t_done;
    info->device_bytes = r->held > 0 ? r->held : r->est;   // actual once admitted
    info->sim_time = ended ? r->sim_time : 0;
}

int rmt3d_server_wait(rmt3d_server *S, int req, double timeout_s, rmt3d_request_info *info) {
    if (!S) return RMT3D_ERR_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(req);
    if (it == S->all.end()) return RMT3D_ERR_ARG;
    Request *r = it->second;
    auto done = [&] { return r->status == RMT3D_REQ_DONE || r->status == RMT3D_REQ_FAILED; };
    bool ok;
    if (timeout_s < 0) {
        S->cv.wait(lk, done);
        ok = true;
    } else {
        ok = S->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
    }
    if (info) fill_info(r, info);
    return ok ? 0 : 1;
}

int rmt3d_server_info(rmt3d_server *S, int req, rmt3d_request_info *info) {
    if (!S || !info) return RMT3D_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(req);
    if (it == S->all.end()) return RMT3D_ERR_ARG;
    fill_info(it->second, info);
    return 0;
}

const char *rmt3d_server_error(rmt3d_server *S, int req) {
    if (!S) return "";
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(req);
    if (it == S->all.end()) return "";
    Request *r = it->second;
    bool ended = r->status == RMT3D_REQ_DONE || r->status == RMT3D_REQ_FAILED;
    return ended ? r->err.c_str() : "";
}

int rmt3d_server_result(rmt3d_server *S, int req, rmt3d_state *st) {
    if (!S || !st) return RMT3D_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(req);
    if (it == S->all.end()) return RMT3D_ERR_ARG;
    Request *r = it->second;
    if (r->status != RMT3D_REQ_DONE && r->status != RMT3D_REQ_FAILED) return RMT3D_ERR_ARG;
    if (r->u.empty()) return RMT3D_ERR_ARG;
    size_t n = r->u.size();
    if (st->u) std::copy(r->u.begin(), r->u.end(), st->u);
    if (st->v) std::copy(r->v.begin(), r->v.end(), st->v);
    if (st->w) std::copy(r->w.begin(), r->w.end(), st->w);
    if (st->p) std::copy(r->p.begin(), r->p.end(), st->p);
    if (st->rm_c) std::copy(r->c.begin(), r->c.end(), st->rm_c);
    (void)n;
    return 0;
}

int rmt3d_server_release(rmt3d_server *S, int req) {
    if (!S) return RMT3D_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(req);
    if (it == S->all.end()) return RMT3D_ERR_ARG;
    Request *r = it->second;
    if (r->status != RMT3D_REQ_DONE && r->status != RMT3D_REQ_FAILED) return RMT3D_ERR_ARG;
    S->all.erase(it);
    delete r;
    return 0;
}

int rmt3d_server_stats(rmt3d_server *S, double *peak_bytes, int *peak_active, double *now) {
    if (!S) return RMT3D_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    if (peak_bytes) *peak_bytes = S->peak_used;
    if (peak_active) *peak_active = S->peak_active;
    if (now) *now = S->now();
    return 0;
}

void rmt3d_server_destroy(rmt3d_server *S) {
    if (!S) return;
    {
        // requests still queued are cancelled; running ones finish
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        for (Request *r : S->queue) {
            r->status = RMT3D_REQ_FAILED;
            r->rc = RMT3D_ERR_ARG;
            r->err = "cancelled: server destroyed";
        }
        S->queue.clear();
    }
    S->cv.notify_all();
    for (auto &t : S->workers) t.join();
    for (auto &kv : S->all) delete kv.second;
    delete S;
}

}  // extern "C"
