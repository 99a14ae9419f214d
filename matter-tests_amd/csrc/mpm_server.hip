// mpm_server.hip — the served form: a request loop with continuous batching over the batch API.
//
// Requests (scene + steps) arrive at given times. The main thread runs the device: at every chunk
// boundary it admits arrived requests in arrival order while their estimated device footprint fits the
// HBM budget beside the resident scenes (strict FIFO: a request that does not fit holds back the ones
// behind it), adds prepared scenes to the batch, steps the batch by one chunk, and retires scenes whose
// steps are done. A preparation thread creates the admitted scenes' contexts and uploads their particles
// while the main thread keeps the batch stepping, so set-up overlaps device work.
#include "mpm_context.hpp"
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

enum ReqState { RQ_QUEUED, RQ_PREPARING, RQ_READY, RQ_RESIDENT, RQ_DONE, RQ_FAILED };

struct Request {
    uint64_t id = 0;
    mpm_config cfg{};
    std::vector<mpm_plate> plates;
    std::vector<mpm_object> objects;
    uint32_t n = 0, nsteps = 0;
    double arrival = 0.0;
    // caller's arrays (copies); empty = not given
    std::vector<double> f[10];
    uint64_t est_bytes = 0;
    // life cycle
    ReqState state = RQ_QUEUED;
    mpm_context* ctx = nullptr;
    double t_admit = -1.0, t_resident = -1.0, t_done = -1.0;
    uint32_t status = 0;
    mpm_step_info info{};
    std::vector<double> out[10];  // final state (keep_results)
};

// Device bytes of a context for n particles whose bounding box spans `span` nodes per axis, sized as the
// engine sizes it (two particle layouts, sort arrays, grid and staging with the headroom of one chunk).
uint64_t estimate_bytes(const mpm_config& c, uint32_t n, const uint32_t* span, uint32_t chunk) {
    const int D = c.dim;
    const bool split = c.precision == MPM_PRECISION_SPLIT;
    const size_t tf = split ? sizeof(float) : sizeof(double);
    const size_t nd = D + D * D + 6, nt = D + D * D + (D == 3 ? 6 : 3);
    const uint64_t pbc = c.pbc ? 1 : 0;  // periodic: key, perm, the cell bytes and the block arrays also hold the copies
    uint64_t bytes = 2 * (nd * sizeof(double) + nt * tf) * ((uint64_t)n + 1024) + (uint64_t)(16 + 8 * pbc + 2 + 2 * pbc) * n + 4;
    uint64_t nodes = 1, blocks = 1;
    const uint32_t E = D == 3 ? 4 : 8;
    for (int d = 0; d < D; ++d) {
        const uint64_t a = span[d] + 8 + 2 * (chunk + 4);
        nodes *= a;
        blocks *= (a + E - 1) / E;
    }
    const uint32_t W = c.spline_degree == 2 ? 3 : 4, T = E + W - 1;
    const uint64_t NT = D == 3 ? (uint64_t)T * T * T : (uint64_t)T * T;
    bytes += (uint64_t)(1 + 2 * D + 1) * nodes * tf + blocks * (12 * (1 + pbc) + 130 + mpm::kMaxNV * NT * tf);
    return bytes;
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t lo = (size_t)pos;
    const size_t hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (pos - (double)lo) * (v[hi] - v[lo]);
}

// Creates and fills the context of a request (preparation thread).
int prepare(Request& r, int device) {
    if (hipSetDevice(device) != hipSuccess) return MPM_ERR_HIP;
    mpm_context* c = nullptr;
    int rc = mpm_create(&r.cfg, &c);
    if (rc) return rc;
    for (const mpm_plate& p : r.plates) if ((rc = mpm_add_plate(c, &p))) { mpm_destroy(c); return rc; }
    for (const mpm_object& o : r.objects) if ((rc = mpm_add_object(c, &o))) { mpm_destroy(c); return rc; }
    mpm_particle_arrays a{};
    double** slot[10] = {&a.x, &a.v, &a.F, &a.Bmat, &a.eps_pl_dev, &a.eps_pl_vol, &a.eps_pl_vol_pradhana, &a.delta_gamma,
                         &a.viscosity, &a.muI};
    for (int k = 0; k < 10; ++k) *slot[k] = r.f[k].empty() ? nullptr : r.f[k].data();
    if ((rc = mpm_set_particles(c, r.n, &a))) { mpm_destroy(c); return rc; }
    r.ctx = c;  // the inputs stay until the request is done: a run stopped by its time limit re-queues it
    return 0;
}

// A request that has reached RQ_DONE or RQ_FAILED no longer needs its input copies.
void release_inputs(Request& r) { for (auto& v : r.f) std::vector<double>().swap(v); }

}  // namespace

struct mpm_server {
    mpm_server_options opt{};
    int device = 0;
    std::vector<std::unique_ptr<Request>> reqs;
};

extern "C" {

int mpm_server_create(const mpm_server_options* opt, mpm_server** out) {
    if (!out) return MPM_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MPM_ERR_HIP;
    mpm_server* s = new (std::nothrow) mpm_server();
    if (!s) return MPM_ERR_OUT_OF_MEMORY;
    if (opt) s->opt = *opt;
    if (!s->opt.max_scenes) s->opt.max_scenes = 256;
    if (!s->opt.chunk_steps) s->opt.chunk_steps = 8;
    if (!s->opt.hbm_budget_bytes) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) { delete s; return MPM_ERR_HIP; }
        s->opt.hbm_budget_bytes = (uint64_t)(0.8 * (double)free_b);
    }
    (void)hipGetDevice(&s->device);
    *out = s;
    return 0;
}

void mpm_server_destroy(mpm_server* s) {
    if (!s) return;
    for (auto& r : s->reqs) if (r->ctx) { mpm_destroy(r->ctx); r->ctx = nullptr; }
    delete s;
}

int mpm_server_submit(mpm_server* s, const mpm_request_desc* q, uint64_t* id) {
    if (!s || !q || !q->cfg || !q->particles || !q->particles->x || q->n == 0) return MPM_ERR_INVALID_ARGUMENT;
    if ((q->n_plates && !q->plates) || (q->n_objects && !q->objects)) return MPM_ERR_INVALID_ARGUMENT;
    const int D = q->cfg->dim;
    if ((D != 2 && D != 3) || !(q->cfg->dx > 0)) return MPM_ERR_INVALID_ARGUMENT;
    std::unique_ptr<Request> r(new (std::nothrow) Request());
    if (!r) return MPM_ERR_OUT_OF_MEMORY;
    r->id = s->reqs.size();
    r->cfg = *q->cfg;
    r->plates.assign(q->plates, q->plates + q->n_plates);
    r->objects.assign(q->objects, q->objects + q->n_objects);
    r->n = q->n; r->nsteps = q->nsteps; r->arrival = q->arrival_s;
    const mpm_particle_arrays& a = *q->particles;
    const double* src[10] = {a.x, a.v, a.F, a.Bmat, a.eps_pl_dev, a.eps_pl_vol, a.eps_pl_vol_pradhana, a.delta_gamma,
                             a.viscosity, a.muI};
    const size_t per[10] = {(size_t)D, (size_t)D, (size_t)D * D, (size_t)D * D, 1, 1, 1, 1, 1, 1};
    for (int k = 0; k < 10; ++k) if (src[k]) r->f[k].assign(src[k], src[k] + per[k] * q->n);
    uint32_t span[3] = {0, 0, 0};
    for (int d = 0; d < D && q->cfg->pbc; ++d) {  // the fixed grid of a periodic run (remeshFixed(4))
        const double L = d == 0 ? q->cfg->Lx : (d == 1 ? q->cfg->Ly : q->cfg->Lz);
        const double sp = L / q->cfg->dx + (d == 1 ? 7.0 : 11.0);
        if (!(sp < 4.0e6) || !(L > 0)) return MPM_ERR_INVALID_ARGUMENT;
        span[d] = (uint32_t)sp;
    }
    for (int d = 0; d < D && !q->cfg->pbc; ++d) {
        double mn = a.x[d], mx = a.x[d];
        for (uint32_t p = 1; p < q->n; ++p) { const double v = a.x[(size_t)p * D + d]; mn = std::min(mn, v); mx = std::max(mx, v); }
        const double sp = (mx - mn) / q->cfg->dx;
        if (!(sp < 4.0e6)) return MPM_ERR_INVALID_ARGUMENT;
        span[d] = (uint32_t)sp;
    }
    r->est_bytes = estimate_bytes(r->cfg, r->n, span, s->opt.chunk_steps);
    if (id) *id = r->id;
    s->reqs.push_back(std::move(r));
    return 0;
}

int mpm_server_run(mpm_server* s, double time_limit_s, mpm_server_stats* stats) {
    if (!s) return MPM_ERR_INVALID_ARGUMENT;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (hipSetDevice(s->device) != hipSuccess) return MPM_ERR_HIP;
    mpm_batch_options bo{s->opt.chunk_steps, s->opt.use_graph};
    mpm_batch* batch = nullptr;
    int rc = mpm_batch_create(&bo, &batch);
    if (rc) return rc;

    // arrival order of the requests not yet served
    std::vector<Request*> order;
    for (auto& r : s->reqs) if (r->state == RQ_QUEUED) order.push_back(r.get());
    std::stable_sort(order.begin(), order.end(), [](const Request* a, const Request* b) { return a->arrival < b->arrival; });

    std::mutex mu;
    std::mutex hip_calls;  // held by the preparation thread around its HIP calls; no graph capture meanwhile
    mpm::batch_set_capture_lock(batch, &hip_calls);
    std::condition_variable cv_prep, cv_main;
    std::deque<Request*> to_prepare, prepared;
    bool quit = false;
    const Clock::time_point t0 = Clock::now();
    auto now_s = [&]() { return std::chrono::duration<double>(Clock::now() - t0).count(); };

    std::thread worker([&]() {
        for (;;) {
            Request* r = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_prep.wait(lk, [&]() { return quit || !to_prepare.empty(); });
                if (to_prepare.empty()) return;
                r = to_prepare.front();
                to_prepare.pop_front();
            }
            int prc;
            {
                std::lock_guard<std::mutex> hk(hip_calls);
                prc = prepare(*r, s->device);
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                if (prc) { r->status |= (uint32_t)prc; r->state = RQ_FAILED; }
                else r->state = RQ_READY;
                prepared.push_back(r);
            }
            cv_main.notify_one();
        }
    });

    std::vector<Request*> resident;
    std::vector<double> step_ms, req_ms, queue_ms;
    uint64_t particle_steps = 0, peak_bytes = 0, pending_bytes = 0;
    uint32_t peak_resident = 0, completed = 0, failed = 0, in_flight = 0;
    size_t next = 0;  // first request of `order` not yet admitted
    uint32_t status = 0;
    const size_t total = order.size();
    double t_last_done = 0.0;
    while (completed + failed < total) {
        const double t = now_s();
        if (time_limit_s > 0 && t > time_limit_s) break;
        // admission: arrived requests, in arrival order, while the footprint fits
        uint64_t resident_bytes = 0;
        for (Request* r : resident) resident_bytes += mpm_device_bytes(r->ctx);
        while (next < total && order[next]->arrival <= t && resident.size() + in_flight < s->opt.max_scenes &&
               resident_bytes + pending_bytes + order[next]->est_bytes <= s->opt.hbm_budget_bytes) {
            Request* r = order[next++];
            r->t_admit = t;
            pending_bytes += r->est_bytes;
            ++in_flight;
            std::lock_guard<std::mutex> lk(mu);
            r->state = RQ_PREPARING;
            to_prepare.push_back(r);
            cv_prep.notify_one();
        }
        // a request larger than the whole budget can never be admitted
        if (next < total && resident.empty() && in_flight == 0 && order[next]->arrival <= t &&
            order[next]->est_bytes > s->opt.hbm_budget_bytes) {
            Request* r = order[next++];
            r->state = RQ_FAILED; r->status |= MPM_ERR_OUT_OF_MEMORY; ++failed;
            release_inputs(*r);
            continue;
        }
        // prepared scenes join the batch
        std::deque<Request*> ready;
        {
            std::lock_guard<std::mutex> lk(mu);
            ready.swap(prepared);
        }
        for (Request* r : ready) {
            --in_flight;
            pending_bytes -= std::min(pending_bytes, r->est_bytes);
            if (r->state == RQ_FAILED) { ++failed; release_inputs(*r); continue; }
            const int arc = mpm_batch_add(batch, r->ctx, r->nsteps);
            if (arc) {
                r->status |= (uint32_t)arc; r->state = RQ_FAILED; mpm_destroy(r->ctx); r->ctx = nullptr; ++failed;
                release_inputs(*r);
                continue;
            }
            r->state = RQ_RESIDENT;
            r->t_resident = now_s();
            queue_ms.push_back(1e3 * (r->t_resident - r->arrival));
            resident.push_back(r);
        }
        if (resident.empty()) {  // nothing on the device: wait for a prepared scene or the next arrival
            std::unique_lock<std::mutex> lk(mu);
            double wait = 0.002;
            if (in_flight == 0 && next < total) wait = std::max(0.0, order[next]->arrival - now_s());
            cv_main.wait_for(lk, std::chrono::duration<double>(std::min(wait, 0.05)), [&]() { return !prepared.empty(); });
            continue;
        }
        peak_resident = std::max<uint32_t>(peak_resident, (uint32_t)resident.size());
        uint64_t rb = 0;
        for (Request* r : resident) rb += mpm_device_bytes(r->ctx);
        peak_bytes = std::max(peak_bytes, rb + pending_bytes);
        // one chunk of the batch
        const double tc0 = now_s();
        mpm_batch_info bi{};
        const int src = mpm_batch_step(batch, s->opt.chunk_steps, 0, &bi);
        const double tc1 = now_s();
        if (src & (MPM_ERR_HIP | MPM_ERR_OUT_OF_MEMORY)) { status |= (uint32_t)src; break; }
        particle_steps += bi.particle_steps;
        uint32_t max_done = 0;
        std::vector<uint32_t> did(resident.size());
        for (size_t i = 0; i < resident.size(); ++i) {
            mpm_step_info inf{};
            uint32_t left = 0;
            mpm_scene_info(resident[i]->ctx, &inf, &left);
            did[i] = inf.steps_done;
            max_done = std::max(max_done, inf.steps_done);
        }
        if (max_done) {
            const double per_step = 1e3 * (tc1 - tc0) / max_done;
            for (size_t i = 0; i < resident.size(); ++i) step_ms.insert(step_ms.end(), did[i], per_step);
        }
        // retire scenes whose budget is spent (or that stopped on an error)
        std::vector<Request*> keep;
        for (size_t i = 0; i < resident.size(); ++i) {
            Request* r = resident[i];
            mpm_step_info inf{};
            uint32_t left = 0;
            mpm_scene_info(r->ctx, &inf, &left);
            r->info.steps_done += inf.steps_done;
            r->info.dt = inf.dt; r->info.time = inf.time; r->info.step = inf.step; r->info.frame = inf.frame;
            r->info.plastic_count = inf.plastic_count; r->info.frame_ended = inf.frame_ended; r->info.finished = inf.finished;
            r->info.status |= inf.status; r->status |= inf.status;
            if (left > 0) { keep.push_back(r); continue; }
            (void)mpm_batch_remove(batch, r->ctx);
            r->t_done = now_s();
            t_last_done = r->t_done;
            const bool bad = (r->status & MPM_ERR_EXIT_MASK) != 0;
            if (s->opt.keep_results && !bad) {
                const int D = r->cfg.dim;
                const size_t per[10] = {(size_t)D, (size_t)D, (size_t)D * D, (size_t)D * D, 1, 1, 1, 1, 1, 1};
                for (int k = 0; k < 10; ++k) r->out[k].assign(per[k] * r->n, 0.0);
                mpm_particle_arrays o{r->out[0].data(), r->out[1].data(), r->out[2].data(), r->out[3].data(), r->out[4].data(),
                                      r->out[5].data(), r->out[6].data(), r->out[7].data(), r->out[8].data(), r->out[9].data()};
                r->status |= (uint32_t)mpm_get_particles(r->ctx, &o);
            }
            mpm_destroy(r->ctx);
            r->ctx = nullptr;
            release_inputs(*r);
            if (bad) { r->state = RQ_FAILED; ++failed; }
            else { r->state = RQ_DONE; ++completed; req_ms.push_back(1e3 * (r->t_done - r->arrival)); }
        }
        resident.swap(keep);
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        quit = true;
        to_prepare.clear();
    }
    cv_prep.notify_all();
    worker.join();
    // Requests interrupted by the time limit (admitted and waiting for preparation, prepared, or resident) go
    // back to the queue with their inputs and start from the beginning in a later run; a request whose
    // preparation failed stays failed.
    for (Request* r : prepared) if (r->ctx) { mpm_destroy(r->ctx); r->ctx = nullptr; }
    for (Request* r : resident) { (void)mpm_batch_remove(batch, r->ctx); mpm_destroy(r->ctx); r->ctx = nullptr; }
    for (Request* r : order) {
        if (r->state != RQ_PREPARING && r->state != RQ_READY && r->state != RQ_RESIDENT) continue;
        r->state = RQ_QUEUED; r->status = 0; r->info = mpm_step_info{};
        r->t_admit = r->t_resident = r->t_done = -1.0;
    }
    mpm_batch_destroy(batch);
    if (stats) {
        stats->submitted = (uint32_t)total; stats->completed = completed; stats->failed = failed;
        stats->peak_resident = peak_resident; stats->peak_resident_bytes = peak_bytes;
        stats->particle_steps = particle_steps;
        stats->wall_s = t_last_done > 0 ? t_last_done : now_s();
        stats->throughput = stats->wall_s > 0 ? (double)particle_steps / stats->wall_s : 0.0;
        stats->step_ms_p50 = percentile(step_ms, 0.5); stats->step_ms_p99 = percentile(step_ms, 0.99);
        stats->request_ms_p50 = percentile(req_ms, 0.5); stats->request_ms_p99 = percentile(req_ms, 0.99);
        stats->queue_ms_p50 = percentile(queue_ms, 0.5); stats->queue_ms_p99 = percentile(queue_ms, 0.99);
    }
    return (int)(status & MPM_ERR_EXIT_MASK);
}

int mpm_server_result(mpm_server* s, uint64_t id, mpm_particle_arrays* out, mpm_step_info* info, double* latency_ms) {
    if (!s || id >= s->reqs.size()) return MPM_ERR_INVALID_ARGUMENT;
    const Request& r = *s->reqs[id];
    if (r.state == RQ_FAILED) return (int)(r.status & MPM_ERR_EXIT_MASK) ? (int)(r.status & MPM_ERR_EXIT_MASK) : MPM_ERR_HIP;
    if (r.state != RQ_DONE) return MPM_ERR_NOT_READY;
    if (info) *info = r.info;
    if (latency_ms) *latency_ms = 1e3 * (r.t_done - r.arrival);
    if (out) {
        double* dst[10] = {out->x, out->v, out->F, out->Bmat, out->eps_pl_dev, out->eps_pl_vol, out->eps_pl_vol_pradhana,
                           out->delta_gamma, out->viscosity, out->muI};
        for (int k = 0; k < 10; ++k) {
            if (!dst[k]) continue;
            if (r.out[k].empty()) return MPM_ERR_NOT_READY;  // results were not kept
            std::memcpy(dst[k], r.out[k].data(), sizeof(double) * r.out[k].size());
        }
    }
    return 0;
}

}  // extern "C"
