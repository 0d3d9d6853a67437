// server.cpp -- continuous-batching request loop over the simultaneous fit.
//
// No counterpart in the reference (which fits one image per call): requests
// (image tiles of uneven size and star count) arrive at any time and are
// admitted in arrival order as soon as a worker slot is free and their HBM
// estimate fits the server's budget.  Each slot owns a cp_ctx (its own HIP
// stream and cached device blocks), so admitted fits overlap on the device,
// and device buffers stay resident between requests.  A fit's result does not
// depend on what runs beside it: every request uses only its own buffers and
// stream, and all reductions are in a fixed order.
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/crowdphot_hip.h"

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Job {
    cp_request req;
    cp_psf_desc psf;
    cp_sim_opts opts;
    uint64_t bytes = 0;
    int state = 0;   // 0 queued, 1 running, 2 done
    int rc = CP_OK;
    int slot = -1;
    bool waited = false;   // a thread is blocked in cp_server_wait on it
    std::string msg;
    double t_submit = 0, t_start = 0, t_end = 0;
    uint64_t charged = 0, peak = 0, peak_resident = 0;
};

}  // namespace

struct cp_server {
    int device = 0;
    uint64_t budget = 0;
    std::vector<cp_ctx*> ctxs;
    // per slot: running, and the bytes it holds against the budget (the
    // running request's estimate, or an idle slot's cached blocks)
    std::vector<uint8_t> busy;
    std::vector<uint64_t> charge;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> queue;
    std::unordered_map<uint64_t, Job> jobs;
    uint64_t next_ticket = 1;
    int n_waiters = 0;
    bool stop = false;
    std::string err;
};

extern "C" uint64_t cp_request_bytes(int32_t ny, int32_t nx, int32_t n_stars, double fit_rad, double model_rad,
                                     int32_t precision, int32_t device_inputs) {
    if (ny <= 0 || nx <= 0 || n_stars < 0 || !(fit_rad > 0)) return 0;
    // an upper bound on the device blocks one simultaneous fit holds at once
    // (its largest phase is the LSQR trials), by buffer family
    const double T = precision == CP_PREC_FP64 ? 8.0 : 4.0;
    const double P = (double)ny * nx;
    const double R = std::ceil(fit_rad), S2 = (2 * R + 1) * (2 * R + 1);
    const double mR = model_rad > 0 ? std::ceil(std::max(model_rad, fit_rad)) : 10.0 * R;
    const double n = n_stars, E = n * S2;
    double b = P * (6 * T + 1 + 4);                           // data, data_work, w, 2 residuals, u; union mask;
                                                              // float32 residual staging
    b += P * 8;                                               // row-major inputs transposed on the device
    if (!device_inputs) b += P * 8 + n * 24;                 // staged image, weights, catalogue
    b += E * 4 * T;                                           // stamp Jacobian rows
    // star-pair rows of the normal matrix (per entry: neighbour, the (neighbour,
    // slot) word and a 3 x 3 float64 block; per upper entry a 16-byte record;
    // build-time scratch): pairs per star at eight times the mean density of
    // anchors within 2R, and at least 0.3 S^2
    const double K = std::min(n, std::max(0.3 * S2, 1 + 8 * n * (4 * R + 1) * (4 * R + 1) / P));
    b += n * K * 100;
    const double tiles = std::pow((2 * mR + 1) / 64.0 + 1.0, 2);
    b += n * tiles * 16 + 4 * P / 4096 * 24;                  // model-box tile lists, live copies, scans
    b += n * 1100 + (8 << 20);                                // per-star records and vectors, partials, results
    return (uint64_t)(b * 1.15);
}

namespace {

// Bytes the other slots hold against the budget.
uint64_t held_by_others(const cp_server* s, int slot) {
    uint64_t o = 0;
    for (int j = 0; j < (int)s->ctxs.size(); ++j)
        if (j != slot) o += s->charge[j];
    return o;
}

uint64_t cached_of(cp_ctx* c) {
    cp_mem_stats m{};
    return cp_mem_stats_get(c, &m) == CP_OK ? m.cached : 0;
}

// Whether the request at the head of the queue may start on the idle `slot`
// (called with the server lock held), and what it is charged: `need` plus the
// blocks the slot keeps cached.  Cached blocks are released only when the
// request would not fit beside them: first the other idle slots', then the
// slot's own.
bool admissible(cp_server* s, int slot, uint64_t need, uint64_t* charge) {
    cp_ctx* own = s->ctxs[slot];
    if (!s->budget) {
        *charge = need;
        return true;
    }
    auto fits = [&] { return held_by_others(s, slot) + s->charge[slot] + need <= s->budget; };
    if (!fits()) {
        for (int j = 0; j < (int)s->ctxs.size(); ++j)
            if (j != slot && !s->busy[j] && s->charge[j]) {
                (void)cp_release_cached(s->ctxs[j]);
                s->charge[j] = cached_of(s->ctxs[j]);
            }
    }
    if (!fits() && s->charge[slot]) {
        (void)cp_release_cached(own);
        s->charge[slot] = cached_of(own);
    }
    if (!fits()) return false;
    *charge = s->charge[slot] + need;
    return true;
}

}  // namespace

static void worker_main(cp_server* s, int slot) {
    cp_ctx* ctx = s->ctxs[slot];
    for (;;) {
        uint64_t t;
        Job* job;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            uint64_t charge = 0;
            s->cv_work.wait(lk, [&] {
                if (s->stop) return true;
                if (s->queue.empty()) return false;
                return admissible(s, slot, s->jobs[s->queue.front()].bytes, &charge);
            });
            if (s->stop) return;
            t = s->queue.front();
            s->queue.pop_front();
            job = &s->jobs[t];   // element references survive rehashing
            job->state = 1;
            job->slot = slot;
            job->t_start = now_s();
            job->charged = charge;
            s->busy[slot] = 1;
            s->charge[slot] = charge;
            // the slot's blocks (reused cached ones included) stay within its charge
            (void)cp_set_budget(ctx, s->budget ? charge : 0);
            // the next queued request may fit beside this one
            s->cv_work.notify_one();
        }
        const cp_request& r = job->req;
        int rc;
        if (r.device_inputs)
            rc = cp_fit_simultaneous_device(ctx, &job->psf, r.image, r.inv_var, r.ny, r.nx, r.y0, r.x0, r.flux0,
                                            r.n_stars, r.fit_rad, &job->opts, r.out);
        else
            rc = cp_fit_simultaneous(ctx, &job->psf, r.image, r.inv_var, r.ny, r.nx, r.y0, r.x0, r.flux0,
                                     r.n_stars, r.fit_rad, &job->opts, r.out);
        cp_mem_stats m{};
        (void)cp_mem_stats_get(ctx, &m);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            job->rc = rc;
            if (rc != CP_OK) job->msg = cp_last_error(ctx);
            job->t_end = now_s();
            job->peak = m.peak;
            job->peak_resident = m.peak_resident;
            job->state = 2;
            s->busy[slot] = 0;
            s->charge[slot] = s->budget ? m.cached : 0;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

extern "C" {

int cp_server_create(cp_server** out, int32_t device, uint64_t hbm_budget, int32_t n_slots) {
    if (!out) return CP_EARG;
    *out = nullptr;
    if (n_slots < 1 || n_slots > 64) return CP_EARG;
    cp_server* s = new cp_server;
    s->device = device;
    s->budget = hbm_budget;
    for (int i = 0; i < n_slots; ++i) {
        cp_ctx* c = nullptr;
        int rc = cp_create(&c, device, 0);
        if (rc != CP_OK) {
            for (auto* x : s->ctxs) cp_destroy(x);
            delete s;
            return rc;
        }
        s->ctxs.push_back(c);
    }
    s->busy.assign(n_slots, 0);
    s->charge.assign(n_slots, 0);
    for (int i = 0; i < n_slots; ++i) s->workers.emplace_back(worker_main, s, i);
    *out = s;
    return CP_OK;
}

int cp_server_submit(cp_server* s, const cp_request* r, uint64_t* ticket) {
    if (!s) return CP_ESTATE;
    std::lock_guard<std::mutex> lk(s->mu);
    if (!r || !ticket || !r->psf || !r->opts || !r->out || !r->image) {
        s->err = "cp_server_submit: NULL request, ticket, psf, opts, out or image";
        return CP_EARG;
    }
    const uint64_t b = cp_request_bytes(r->ny, r->nx, r->n_stars, r->fit_rad, r->opts->model_rad,
                                        r->opts->precision, r->device_inputs);
    if (b == 0) {
        s->err = "cp_server_submit: bad image size, star count or fit_rad";
        return CP_EARG;
    }
    if (s->budget && b > s->budget) {
        s->err = "cp_server_submit: request needs " + std::to_string(b) + " bytes, more than the HBM budget " +
                 std::to_string(s->budget);
        return CP_ENOMEM;
    }
    const uint64_t t = s->next_ticket++;
    Job& j = s->jobs[t];
    j.req = *r;
    j.psf = *r->psf;
    j.opts = *r->opts;
    j.req.psf = nullptr;
    j.req.opts = nullptr;
    j.bytes = b;
    j.t_submit = now_s();
    s->queue.push_back(t);
    *ticket = t;
    s->cv_work.notify_one();
    return CP_OK;
}

int cp_server_poll(cp_server* s, uint64_t ticket, int32_t* done) {
    if (!s) return CP_ESTATE;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end() || !done) {
        s->err = "cp_server_poll: unknown ticket";
        return CP_EARG;
    }
    *done = it->second.state == 2;
    return CP_OK;
}

int cp_server_wait(cp_server* s, uint64_t ticket, cp_request_stat* st, char* msg, int32_t msg_cap) {
    if (!s) return CP_ESTATE;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) {
        s->err = "cp_server_wait: unknown ticket";
        return CP_EARG;
    }
    Job& j = it->second;   // a reference, not the iterator: submits may rehash the map
    if (j.waited) {
        s->err = "cp_server_wait: another thread is already waiting on this ticket";
        return CP_ESTATE;
    }
    j.waited = true;
    ++s->n_waiters;
    s->cv_done.wait(lk, [&] { return j.state == 2; });
    if (st) {
        st->rc = j.rc;
        st->slot = j.slot;
        st->t_submit = j.t_submit;
        st->t_start = j.t_start;
        st->t_end = j.t_end;
        st->hbm_bytes = j.bytes;
        st->hbm_charged = j.charged;
        st->hbm_peak = j.peak;
        st->hbm_resident_peak = j.peak_resident;
    }
    if (msg && msg_cap > 0) {
        std::strncpy(msg, j.msg.c_str(), (size_t)msg_cap - 1);
        msg[msg_cap - 1] = 0;
    }
    const int rc = j.rc;
    s->jobs.erase(ticket);
    --s->n_waiters;
    if (s->stop) s->cv_done.notify_all();   // cp_server_destroy waits for the last waiter
    return rc;
}

const char* cp_server_last_error(const cp_server* s) { return s ? s->err.c_str() : "NULL server"; }

void cp_server_destroy(cp_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        // queued requests never run: complete them so their waiters return
        for (uint64_t t : s->queue) {
            Job& j = s->jobs[t];
            j.state = 2;
            j.rc = CP_ESTATE;
            j.msg = "the server was destroyed before this request ran";
            j.t_start = j.t_end = now_s();
        }
        s->queue.clear();
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    for (auto& w : s->workers) w.join();   // running requests finish and notify
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [&] { return s->n_waiters == 0; });
    }
    for (auto* c : s->ctxs) cp_destroy(c);
    delete s;
}

}  // extern "C"
