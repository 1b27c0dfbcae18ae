// fmm3d_server.hip -- the serving loop: a scheduler thread that keeps
// velocity models resident, admits requests by device-memory footprint and
// runs continuous batching on the column-walk engine (fmm3d_engine.h).
//
// Each (grid, model id, precision, order) gets one context opened as an
// engine with `cap` source slots, cap chosen so that the model and its
// slots fit the server's HBM budget.  A request's sources take free slots
// of their model's context as they appear; every loop iteration queues one
// chunk of sweeps on every context that has sources in flight (each on its
// own stream, so contexts of different grids overlap on the device), then
// collects the converged sources, reads their receiver times and frees
// their slots.  Idle contexts are evicted, least recently used first, when
// a new model needs the memory.  Point sources first need their refined
// neighbourhood solved (float64, on a refined sub-grid): prep workers make
// those fixed sets, each on its own refined-grid context and stream, while
// the scheduler keeps sweeping, and point sources are admitted in order as
// their sets are ready.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fmm3d.h"
#include "fmm3d_engine.h"

namespace {

using Clock = std::chrono::steady_clock;

constexpr int SWEEPS_PER_CHUNK = 4;   // default sweeps between collections (FMM3D_SERVER_CHUNK)
constexpr int PREP_WORKERS = 4;    // threads solving refined point-source neighbourhoods

struct Request {
    int64_t ticket = 0;
    fmm3d_grid_desc grid{};
    uint64_t model_id = 0;
    int kind = FMM3D_SRC_TELE, nsrc = 0, refine = 5, radius = 10, nrc = 0;
    int order = FMM3D_ORDER_MIXED, precision = FMM3D_PREC_FP32;
    std::vector<double> velv, faces, src, rcv;
    std::vector<float> times;
    bool want_rows = false;                // trace rays and keep Frechet rows (fmm3d_request.rows)
    std::vector<std::vector<int64_t>> rrp; // per source: nrc+1 row offsets
    std::vector<std::vector<int32_t>> rcols;
    std::vector<std::vector<float>> rvals;
    int admitted = 0, done = 0;
    int status = FMM3D_OK;
    std::atomic<bool> finished{false};     // set under the mutex, read by the scheduler without it
    // point sources: fixed sets made by the prep workers (under the server
    // mutex): next_prep is the next source to hand out, pready marks the
    // finished ones and prepared counts the ready prefix the scheduler may
    // admit; prep_busy counts workers holding the request
    std::vector<std::vector<int64_t>> pnode;
    std::vector<std::vector<double>> ptv;
    std::vector<char> pready;
    int next_prep = 0, prepared = 0, prep_busy = 0;
    int retired = 0;                       // sources that left their slots (under the mutex)
    bool released = false;
    Clock::time_point t_submit, t_done;
};

struct Model {
    std::string key;
    fmm3d_ctx* ctx = nullptr;
    int cap = 0;
    double bytes = 0.0;
    std::vector<std::pair<Request*, int>> owner;   // per slot: request and source index
    int in_flight = 0;
    bool launched = false;
    Clock::time_point last_used;
};

std::string model_key(const Request& r)
{
    std::string k(reinterpret_cast<const char*>(&r.grid), sizeof r.grid);
    k.append(reinterpret_cast<const char*>(&r.model_id), sizeof r.model_id);
    k.append(reinterpret_cast<const char*>(&r.precision), sizeof r.precision);
    k.append(reinterpret_cast<const char*>(&r.order), sizeof r.order);
    return k;
}

}  // namespace

struct fmm3d_server {
    int device = 0;
    double budget = 0.0;
    int max_slots = 64;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_prep;
    std::deque<Request*> queue;                          // not yet fully admitted
    std::deque<Request*> prep_q;                         // point requests with sources to hand out
    std::vector<std::thread> prep_th;                    // prep workers
    std::map<int64_t, std::unique_ptr<Request>> reqs;    // live tickets
    std::vector<std::unique_ptr<Model>> models;          // scheduler thread only
    int64_t next_ticket = 1;
    int chunk = SWEEPS_PER_CHUNK;                        // sweeps queued per context per loop iteration
    bool stop = false;
    std::thread th;
    // statistics (under mu)
    std::vector<double> lat_ms;
    int64_t req_done = 0, src_done = 0;
    bool any = false;
    Clock::time_point t_first, t_last;
    double hbm_used = 0.0;
    int in_flight = 0;
    int nmodels = 0;

    void run();
    void prep();
    void reap();
    Model* model_for(Request* r);
    void finish(Request* r, int status);
};

// Why a request failed, to stderr when FMM3D_SERVER_TRACE is set.
static void trace_failure(const Request* r, int status, const char* what)
{
    static const bool on = getenv("FMM3D_SERVER_TRACE") != nullptr;
    if (on) fprintf(stderr, "fmm3d server: request %lld: status %d: %s\n", (long long)r->ticket, status, what);
}

void fmm3d_server::finish(Request* r, int status)
{
    std::lock_guard<std::mutex> lk(mu);
    if (r->finished) return;
    if (status != FMM3D_OK && r->status == FMM3D_OK) r->status = status;
    r->finished = true;
    r->t_done = Clock::now();
    lat_ms.push_back(std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count());
    req_done++;
    t_last = r->t_done;
    cv_done.notify_all();
}

// The context of a request's model: found, or created if the budget allows
// (evicting idle contexts, least recently used first).  Null: not now.
Model* fmm3d_server::model_for(Request* r)
{
    const std::string key = model_key(*r);
    for (auto& m : models)
        if (m->key == key) return m.get();
    const double mb = (double)fmm3d_engine_model_bytes(&r->grid);
    const double sb = (double)fmm3d_engine_slot_bytes(&r->grid, r->precision);
    if (mb + sb > budget) {
        // not even alone with one slot: waiting would never end
        finish(r, FMM3D_E_NOMEM);
        return nullptr;
    }
    for (;;) {
        double used = 0.0;
        for (auto& m : models) used += m->bytes;
        const double room = budget - used - mb;
        if (room >= sb) break;
        // evict the least recently used idle context
        Model* victim = nullptr;
        for (auto& m : models)
            if (m->in_flight == 0 && (!victim || m->last_used < victim->last_used)) victim = m.get();
        if (!victim) return nullptr;
        fmm3d_destroy(victim->ctx);
        models.erase(std::remove_if(models.begin(), models.end(), [&](const std::unique_ptr<Model>& m) { return m.get() == victim; }),
                     models.end());
    }
    double used = 0.0;
    for (auto& m : models) used += m->bytes;
    int cap = (int)std::min<double>(max_slots, floor((budget - used - mb) / sb));
    if (cap < 1) return nullptr;
    auto m = std::make_unique<Model>();
    m->key = key;
    int rc = fmm3d_create(&r->grid, device, &m->ctx);
    if (rc == FMM3D_OK) rc = fmm3d_set_model_vtx(m->ctx, r->velv.data());
    if (rc == FMM3D_OK) rc = fmm3d_engine_open(m->ctx, cap, r->precision, r->order);
    if (rc != FMM3D_OK) {
        if (m->ctx) fmm3d_destroy(m->ctx);
        finish(r, rc);
        return nullptr;
    }
    m->cap = cap;
    m->bytes = mb + cap * sb;
    m->owner.assign(cap, { nullptr, -1 });
    m->last_used = Clock::now();
    models.push_back(std::move(m));
    return models.back().get();
}

// Drop released requests (their result was taken) that no thread still
// holds: not queued for or in preparation, not in the admission queue, no
// source left in a slot.  Called with the mutex held.
void fmm3d_server::reap()
{
    for (auto it = reqs.begin(); it != reqs.end();) {
        Request* r = it->second.get();
        if (r->released && r->prep_busy == 0 && r->retired == r->admitted &&
            std::find(queue.begin(), queue.end(), r) == queue.end() &&
            std::find(prep_q.begin(), prep_q.end(), r) == prep_q.end())
            it = reqs.erase(it);
        else
            ++it;
    }
}

// A prep worker: fixed sets of point sources, beside the scheduler's
// sweeps.  It takes a run of sources from the oldest request with sources
// left -- an even share of what is left among the workers, at most 32 --
// and solves their refined blocks together.  Each worker has its own
// refined-grid context (and stream).
void fmm3d_server::prep()
{
    // the scheduler may be capturing a sweep graph meanwhile
    hipStreamCaptureMode mode = hipStreamCaptureModeRelaxed;
    (void)hipThreadExchangeStreamCaptureMode(&mode);
    (void)hipSetDevice(device);
    fmm3d_ctx* fine = nullptr;
    fmm3d_ctx* shadow = nullptr;
    std::string shadow_key;
    std::vector<std::vector<int64_t>> node;
    std::vector<std::vector<double>> tv;
    std::string err;
    for (;;) {
        Request* r = nullptr;
        int q = -1, b = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_prep.wait(lk, [&] { return stop || !prep_q.empty(); });
            if (stop) break;
            r = prep_q.front();
            if (r->finished || r->next_prep >= r->nsrc) {
                prep_q.pop_front();
                reap();   // a failed request may have been released meanwhile
                continue;
            }
            q = r->next_prep;
            b = std::min(32, std::max(1, (r->nsrc - q + PREP_WORKERS - 1) / PREP_WORKERS));
            r->next_prep += b;
            if (r->next_prep >= r->nsrc) prep_q.pop_front();
            r->prep_busy++;
        }
        // the grid and model view, kept while requests share them
        const std::string key = model_key(*r);
        if (!shadow || key != shadow_key) {
            if (shadow) fmm3d_engine_shadow_free(shadow);
            shadow = fmm3d_engine_shadow(&r->grid, r->velv.data(), device);
            shadow_key = key;
        }
        const int rc = fmm3d_engine_point_sets(shadow, &fine, err, b, r->src.data() + 3 * q, r->refine, r->radius,
                                               r->order, node, tv);
        if (rc != FMM3D_OK) {
            trace_failure(r, rc, err.c_str());
            finish(r, rc);
        }
        std::lock_guard<std::mutex> lk(mu);
        if (rc == FMM3D_OK) {
            for (int i = 0; i < b; ++i) {
                r->pnode[q + i].swap(node[i]);
                r->ptv[q + i].swap(tv[i]);
                r->pready[q + i] = 1;
            }
            while (r->prepared < r->nsrc && r->pready[r->prepared]) r->prepared++;
            cv_work.notify_all();
        }
        r->prep_busy--;
        reap();
    }
    if (shadow) fmm3d_engine_shadow_free(shadow);
    if (fine) fmm3d_destroy(fine);
}

void fmm3d_server::run()
{
    hipStreamCaptureMode mode = hipStreamCaptureModeRelaxed;
    (void)hipThreadExchangeStreamCaptureMode(&mode);
    (void)hipSetDevice(device);
    std::vector<int64_t> node;
    std::vector<double> tv;
    std::vector<int> done, failed;
    std::vector<int> rslots, rnrc;
    std::vector<const double*> rrcv;
    std::vector<float*> rout;
    std::vector<int> fslots, fnrc;        // slots whose rows are traced together
    std::vector<const double*> frcv;
    std::vector<int64_t> frp;
    std::vector<int32_t> fcols;
    std::vector<float> fvals;
    for (;;) {
        std::vector<Request*> waiting;
        std::vector<int> ready;
        {
            std::unique_lock<std::mutex> lk(mu);
            if (stop) break;
            // with nothing in flight, sleep until a source can be admitted
            auto admissible = [&] {
                for (Request* r : queue)
                    if (!r->finished && (r->kind == FMM3D_SRC_TELE || r->prepared > r->admitted)) return true;
                return false;
            };
            if (in_flight == 0) cv_work.wait(lk, [&] { return stop || admissible(); });
            if (stop) break;
            waiting.assign(queue.begin(), queue.end());
            for (Request* r : waiting) ready.push_back(r->kind == FMM3D_SRC_TELE ? r->nsrc : r->prepared);
        }
        // admission: free slots of each request's model take its next sources
        for (size_t wi = 0; wi < waiting.size(); ++wi) {
            Request* r = waiting[wi];
            if (r->finished || r->admitted >= ready[wi]) continue;
            Model* m = model_for(r);
            if (!m) continue;
            for (int slot = 0; slot < m->cap && r->admitted < ready[wi] && !r->finished; ++slot) {
                if (m->owner[slot].first) continue;
                const int q = r->admitted;
                node.clear();
                tv.clear();
                int rc = FMM3D_OK;
                const double* src3 = nullptr;
                if (r->kind == FMM3D_SRC_TELE) {
                    // straight from the face, queued behind the running sweeps
                    const size_t NL = (size_t)r->grid.nnx * r->grid.nnz;
                    rc = fmm3d_engine_admit_face(m->ctx, slot, r->faces.data() + (size_t)q * NL);
                } else {
                    src3 = r->src.data() + 3 * q;
                    {
                        std::lock_guard<std::mutex> lk(mu);
                        node.swap(r->pnode[q]);
                        tv.swap(r->ptv[q]);
                    }
                    rc = fmm3d_engine_admit(m->ctx, slot, node.data(), tv.data(), (int64_t)node.size(), 1, src3);
                }
                if (rc != FMM3D_OK) {
                    trace_failure(r, rc, fmm3d_last_error(m->ctx));
                    finish(r, rc);
                    break;
                }
                m->owner[slot] = { r, q };
                m->in_flight++;
                std::lock_guard<std::mutex> lk(mu);
                r->admitted++;
                in_flight++;
            }
            m->last_used = Clock::now();
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            queue.erase(std::remove_if(queue.begin(), queue.end(),
                                       [](Request* r) { return r->finished || r->admitted >= r->nsrc; }),
                        queue.end());
            reap();
            double used = 0.0;
            for (auto& m : models) used += m->bytes;
            hbm_used = used;
            nmodels = (int)models.size();
        }
        // one chunk of sweeps on every context with sources in flight ...
        for (auto& m : models) {
            m->launched = false;
            if (m->in_flight == 0) continue;
            if (fmm3d_engine_launch(m->ctx, chunk) == FMM3D_OK) m->launched = true;
            else if (getenv("FMM3D_SERVER_TRACE")) fprintf(stderr, "fmm3d server: launch: %s\n", fmm3d_last_error(m->ctx));
        }
        // ... then the converged sources leave their slots
        for (auto& m : models) {
            if (!m->launched) continue;
            int rc = fmm3d_engine_collect(m->ctx, done, failed);
            if (rc != FMM3D_OK) { failed.clear(); done.clear(); }
            // the converged slots' receiver times in one launch and one wait
            rslots.clear(); rnrc.clear(); rrcv.clear(); rout.clear();
            for (int slot : done) {
                Request* r = m->owner[slot].first;
                if (!r) continue;
                rslots.push_back(slot);
                rnrc.push_back(r->nrc);
                rrcv.push_back(r->rcv.data());
                rout.push_back(r->times.data() + (size_t)m->owner[slot].second * r->nrc);
            }
            int rst = fmm3d_engine_receivers_many(m->ctx, (int)rslots.size(), rslots.data(), rnrc.data(),
                                                  rrcv.data(), rout.data());
            // rays and Frechet rows of the converged slots whose request asked
            // for them, all in one pass, then split per slot
            fslots.clear(); fnrc.clear(); frcv.clear();
            for (int slot : done) {
                Request* r = m->owner[slot].first;
                if (!r || !r->want_rows || rst != FMM3D_OK) continue;
                fslots.push_back(slot);
                fnrc.push_back(r->nrc);
                frcv.push_back(r->rcv.data());
            }
            if (!fslots.empty()) {
                const int frc = fmm3d_engine_frechet_many(m->ctx, (int)fslots.size(), fslots.data(), fnrc.data(),
                                                          frcv.data(), frp, fcols, fvals);
                int64_t ray = 0;
                for (size_t k = 0; k < fslots.size(); ++k) {
                    Request* r = m->owner[fslots[k]].first;
                    const int q = m->owner[fslots[k]].second;
                    if (frc == FMM3D_OK) {
                        const int64_t b0 = frp[ray], b1 = frp[ray + fnrc[k]];
                        r->rrp[q].resize((size_t)fnrc[k] + 1);
                        for (int i = 0; i <= fnrc[k]; ++i) r->rrp[q][i] = frp[ray + i] - b0;
                        r->rcols[q].assign(fcols.begin() + b0, fcols.begin() + b1);
                        r->rvals[q].assign(fvals.begin() + b0, fvals.begin() + b1);
                    } else {
                        trace_failure(r, frc, fmm3d_last_error(m->ctx));
                        std::lock_guard<std::mutex> lk(mu);
                        if (r->status == FMM3D_OK) r->status = frc;
                    }
                    ray += fnrc[k];
                }
            }
            for (int pass = 0; pass < 2; ++pass)
                for (int slot : (pass == 0 ? done : failed)) {
                    Request* r = m->owner[slot].first;
                    if (!r) continue;
                    int st = pass == 0 ? rst : FMM3D_E_NOCONV;
                    if (st != FMM3D_OK) trace_failure(r, st, pass == 0 ? fmm3d_last_error(m->ctx) : "no convergence");
                    fmm3d_engine_retire(m->ctx, slot);
                    m->owner[slot] = { nullptr, -1 };
                    m->in_flight--;
                    {
                        std::lock_guard<std::mutex> lk(mu);
                        in_flight--;
                        src_done++;
                        r->retired++;
                        if (st != FMM3D_OK && r->status == FMM3D_OK) r->status = st;
                    }
                    if (++r->done == r->nsrc) finish(r, FMM3D_OK);
                }
            m->last_used = Clock::now();
        }
    }
    for (auto& m : models) fmm3d_destroy(m->ctx);
    models.clear();
}

extern "C" int fmm3d_server_create(int device, double hbm_budget, int max_slots, fmm3d_server** out)
{
    if (!out || !(hbm_budget > 0.0) || max_slots < 1) return FMM3D_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return FMM3D_E_DEVICE;
    fmm3d_server* s = new fmm3d_server();
    s->device = device;
    s->budget = hbm_budget;
    s->max_slots = max_slots;
    if (const char* ch = getenv("FMM3D_SERVER_CHUNK")) s->chunk = std::min(8, std::max(1, atoi(ch)));
    s->th = std::thread([s] { s->run(); });
    for (int w = 0; w < PREP_WORKERS; ++w) s->prep_th.emplace_back([s] { s->prep(); });
    *out = s;
    return FMM3D_OK;
}

extern "C" int fmm3d_server_submit(fmm3d_server* s, const fmm3d_request* q, int64_t* ticket)
{
    if (!s || !q || !ticket || !q->velv || q->nsrc <= 0 || q->nrc < 0 || (q->nrc > 0 && !q->rcv)) return FMM3D_E_ARG;
    if (q->kind != FMM3D_SRC_TELE && q->kind != FMM3D_SRC_POINT) return FMM3D_E_ARG;
    if ((q->kind == FMM3D_SRC_TELE && !q->faces) || (q->kind == FMM3D_SRC_POINT && !q->src)) return FMM3D_E_ARG;
    if (q->precision != FMM3D_PREC_FP32 && q->precision != FMM3D_PREC_FP64) return FMM3D_E_ARG;
    if (q->order != FMM3D_ORDER_FIRST && q->order != FMM3D_ORDER_MIXED) return FMM3D_E_ARG;
    const fmm3d_grid_desc& g = q->grid;
    if (g.nnr < 3 || g.nnx < 3 || g.nnz < 3 || g.nvr < 2 || g.nvx < 2 || g.nvz < 2) return FMM3D_E_ARG;
    auto r = std::make_unique<Request>();
    r->grid = g;
    r->model_id = q->model_id;
    r->kind = q->kind;
    r->nsrc = q->nsrc;
    r->refine = q->refine;
    r->radius = q->radius;
    r->nrc = q->nrc;
    r->order = q->order;
    r->precision = q->precision;
    const size_t nv = (size_t)(g.nvr + 2) * (g.nvx + 2) * (g.nvz + 2);
    r->velv.assign(q->velv, q->velv + nv);
    if (q->kind == FMM3D_SRC_TELE) {
        r->faces.assign(q->faces, q->faces + (size_t)q->nsrc * g.nnx * g.nnz);
        for (double v : r->faces)
            if (v != v) return FMM3D_E_ARG;   // NaN boundary time
    }
    else r->src.assign(q->src, q->src + 3 * (size_t)q->nsrc);
    r->rcv.assign(q->rcv, q->rcv + 3 * (size_t)q->nrc);
    r->times.assign((size_t)q->nsrc * q->nrc, NAN);
    r->want_rows = q->rows != 0;
    if (r->want_rows) {
        r->rrp.assign(q->nsrc, std::vector<int64_t>((size_t)q->nrc + 1, 0));
        r->rcols.resize(q->nsrc);
        r->rvals.resize(q->nsrc);
    }
    if (q->kind == FMM3D_SRC_POINT) {
        r->pnode.resize(q->nsrc);
        r->ptv.resize(q->nsrc);
        r->pready.assign(q->nsrc, 0);
    }
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->any) { s->any = true; s->t_first = r->t_submit; }
    r->ticket = s->next_ticket++;
    *ticket = r->ticket;
    s->queue.push_back(r.get());
    if (r->kind == FMM3D_SRC_POINT) {
        s->prep_q.push_back(r.get());
        s->cv_prep.notify_all();
    }
    s->reqs[r->ticket] = std::move(r);
    s->cv_work.notify_all();
    return FMM3D_OK;
}

extern "C" int fmm3d_server_wait(fmm3d_server* s, int64_t ticket, double timeout_s)
{
    if (!s) return FMM3D_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return FMM3D_E_ARG;
    Request* r = it->second.get();
    if (r->finished) return FMM3D_OK;
    if (timeout_s <= 0.0) return FMM3D_E_TIMEOUT;
    const bool ok = s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return r->finished.load(); });
    return ok ? FMM3D_OK : FMM3D_E_TIMEOUT;
}

extern "C" int fmm3d_server_rows(fmm3d_server* s, int64_t ticket, int64_t* row_ptr, int32_t* cols, float* vals,
                                 int64_t cap)
{
    if (!s || !row_ptr || cap < 0) return FMM3D_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return FMM3D_E_ARG;
    Request* r = it->second.get();
    if (!r->finished) return FMM3D_E_TIMEOUT;
    if (r->released || !r->want_rows) return FMM3D_E_STATE;
    if (r->status != FMM3D_OK) return r->status;
    // source-major, receiver order: the layout of fmm3d_frechet
    row_ptr[0] = 0;
    int64_t k = 0;
    for (int q = 0; q < r->nsrc; ++q)
        for (int i = 0; i < r->nrc; ++i, ++k) row_ptr[k + 1] = row_ptr[k] + (r->rrp[q][i + 1] - r->rrp[q][i]);
    const int64_t need = row_ptr[k];
    if (need > cap) return FMM3D_E_CAPACITY;
    if (need > 0 && (!cols || !vals)) return FMM3D_E_ARG;
    int64_t at = 0;
    for (int q = 0; q < r->nsrc; ++q) {
        const size_t n = r->rcols[q].size();
        if (n) {
            memcpy(cols + at, r->rcols[q].data(), n * sizeof(int32_t));
            memcpy(vals + at, r->rvals[q].data(), n * sizeof(float));
        }
        at += (int64_t)n;
    }
    return FMM3D_OK;
}

extern "C" int fmm3d_server_result(fmm3d_server* s, int64_t ticket, float* times, double* latency_ms)
{
    if (!s) return FMM3D_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return FMM3D_E_ARG;
    Request* r = it->second.get();
    if (!r->finished) return FMM3D_E_TIMEOUT;
    if (r->released) return FMM3D_E_ARG;
    if (times) memcpy(times, r->times.data(), r->times.size() * sizeof(float));
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    const int st = r->status;
    // a failed request can still have sources in slots or in preparation:
    // it goes once no thread holds it
    r->released = true;
    s->reap();
    return st;
}

extern "C" int fmm3d_server_stats(fmm3d_server* s, fmm3d_server_info* o)
{
    if (!s || !o) return FMM3D_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *o = fmm3d_server_info{};
    o->requests_done = s->req_done;
    o->sources_done = s->src_done;
    std::vector<double> l = s->lat_ms;
    if (!l.empty()) {
        std::sort(l.begin(), l.end());
        auto pct = [&](double p) { return l[std::min(l.size() - 1, (size_t)floor(p * (l.size() - 1) + 0.5))]; };
        o->p50_ms = pct(0.50);
        o->p99_ms = pct(0.99);
    }
    if (s->any && s->req_done > 0) {
        const double span = std::chrono::duration<double>(s->t_last - s->t_first).count();
        o->source_solves_per_s = span > 0.0 ? s->src_done / span : 0.0;
    }
    o->hbm_bytes = s->hbm_used;
    o->models_resident = s->nmodels;
    o->sources_in_flight = s->in_flight;
    return FMM3D_OK;
}

extern "C" int fmm3d_server_destroy(fmm3d_server* s)
{
    if (!s) return FMM3D_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
        s->cv_prep.notify_all();
    }
    if (s->th.joinable()) s->th.join();
    for (auto& t : s->prep_th) t.join();
    delete s;
    return FMM3D_OK;
}
