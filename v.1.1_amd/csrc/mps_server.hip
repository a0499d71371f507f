// mps_server.hip -- serving form: a request loop with continuous batching.
//
// Requests (independent fixed-step solves) are validated and prepared (CSR, Dirichlet records,
// material constants) on the submitting thread.  One worker thread owns a batch (mps_batch.hip)
// and loops:
//   1. enqueue one graph launch (steps_per_launch steps of every active solve);
//   2. while it runs, admit queued requests in arrival order as long as a slot and an arena
//      block of the request's footprint are free (uploads go on a separate stream; the set-up
//      kernels queue behind the running launch, so an admitted solve joins the next launch);
//   3. wait for the launch, retire finished solves (copy u out, free slot and block).
// The element block size and resident-block multiple come from an online tuner (when the configuration
// leaves them 0): per mix size (active elements, in half-decade buckets) it times one real launch with each
// candidate, normalised by the element-steps the launch performed, and keeps the fastest for that bucket.
// Results do not depend on the choice, so the trials are production launches and nothing is replayed.
// State stays on the device for the whole solve; a failed solve (non-finite u) is retired with
// its error code and does not affect the others.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/mps_gpu.h"
#include "mps_api_internal.h"

using namespace mpsh;
using Clock = std::chrono::steady_clock;

namespace {

struct Request {
    int64_t id = 0;
    HostPrep hp;
    size_t footprint = 0;
    int32_t n_nodes = 0;
    int32_t state = MPS_REQ_QUEUED;
    int32_t slot = -1;
    int32_t status = MPS_OK;
    Clock::time_point t_submit, t_admit, t_done;
    std::vector<double> u;
};

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

// Online tuner of (element block size, grid multiple) per mix size.
struct Tuner {
    bool on = false;
    std::vector<std::pair<int, int>> cand;
    std::map<int, std::pair<int, int>> best;  // bucket -> choice
    int explore = -1;                         // bucket under exploration (-1: none)
    size_t next = 0;                          // candidate of the next timed launch
    std::vector<double> cost;                 // ms per element-step of each candidate
    int tunings = 0;
    Tuner() {
        for (int bs : {64, 128, 256})
            for (int gm : {1, 2, 4, 8, 16})
                if (bs * gm <= 4096) cand.emplace_back(bs, gm);
    }
    static int bucket(int64_t elements) {
        return elements > 0 ? (int)std::floor(2.0 * std::log10((double)elements)) : 0;
    }
};

}  // namespace

struct mps_server_s {
    mps_server_cfg cfg{};
    mps_batch batch = nullptr;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work;  // new submissions / shutdown
    std::condition_variable cv_done;  // completions
    bool stop = false;
    int64_t next_id = 1;
    std::deque<std::shared_ptr<Request>> queue;                 // submitted, not admitted
    std::unordered_map<int64_t, std::shared_ptr<Request>> all;  // requests not yet collected
    std::vector<std::shared_ptr<Request>> running;              // admitted (worker only)
    std::vector<double> lat_ms, queue_ms;                       // completed requests
    int64_t submitted = 0, completed = 0, failed = 0, launches = 0;
    Clock::time_point t_first, t_last;
    bool have_first = false;
    double busy_ms = 0.0;
    Tuner tuner;
    int cur_bs = 0, cur_gm = 0;  // tuning of the latest launch (under mu)
    std::string err;
};

// Before a launch: the tuning for the current mix -- the bucket's choice, or the next candidate while the
// bucket is explored.  Returns the bucket the launch belongs to.
static int tune_before_launch(mps_server_s* S) {
    Tuner& T = S->tuner;
    const int b = Tuner::bucket(mps_batch_active_elements(S->batch));
    if (!T.on) return b;
    auto it = T.best.find(b);
    if (it != T.best.end()) {
        T.explore = -1;
        (void)mps_batch_set_tuning(S->batch, it->second.first, it->second.second);
        return b;
    }
    if (T.explore != b) {  // a new mix size: time every candidate on the next launches
        T.explore = b;
        T.next = 0;
        T.cost.assign(T.cand.size(), 0.0);
    }
    (void)mps_batch_set_tuning(S->batch, T.cand[T.next].first, T.cand[T.next].second);
    return b;
}

// After a launch of `work` element-steps that took `ms` on the device, in bucket b.
static void tune_after_launch(mps_server_s* S, int b, double work, double ms) {
    Tuner& T = S->tuner;
    if (!T.on || T.explore != b || work <= 0.0 || ms <= 0.0) return;
    T.cost[T.next] = ms / work;
    if (++T.next < T.cand.size()) return;
    size_t k = 0;
    for (size_t i = 1; i < T.cost.size(); i++)
        if (T.cost[i] < T.cost[k]) k = i;
    T.best[b] = T.cand[k];
    T.explore = -1;
    std::lock_guard<std::mutex> lk(S->mu);
    T.tunings++;
}

static void retire(mps_server_s* S) {
    std::vector<std::shared_ptr<Request>> keep;
    for (auto& r : S->running) {
        mps_report rep;
        if (mps_batch_query(S->batch, r->slot, &rep) != MPS_OK || rep.status == MPS_RUNNING) {
            keep.push_back(r);
            continue;
        }
        const int32_t st = rep.status < 0 ? rep.status : MPS_OK;
        if (st == MPS_OK) {
            r->u.resize(3 * (size_t)r->n_nodes);
            (void)mps_batch_get_state(S->batch, r->slot, r->u.data(), nullptr, nullptr);
        }
        (void)mps_batch_remove(S->batch, r->slot);
        std::lock_guard<std::mutex> lk(S->mu);
        r->status = st;
        r->t_done = Clock::now();
        r->state = st == MPS_OK ? MPS_REQ_DONE : MPS_REQ_FAILED;
        if (st == MPS_OK) {
            S->completed++;
            S->lat_ms.push_back(ms_between(r->t_submit, r->t_done));
            S->queue_ms.push_back(ms_between(r->t_submit, r->t_admit));
        } else {
            S->failed++;
        }
        S->t_last = r->t_done;
        S->cv_done.notify_all();
    }
    S->running.swap(keep);
}

static void admit(mps_server_s* S) {
    for (;;) {
        std::shared_ptr<Request> r;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (S->queue.empty()) return;
            int free_slots = 0;
            size_t largest = 0;
            mps_batch_room(S->batch, &free_slots, &largest);
            r = S->queue.front();
            if (free_slots == 0 || r->footprint > largest) return;  // head of line waits for room
            S->queue.pop_front();
        }
        int32_t slot = -1;
        const int rc = mps_batch_add_prepared(S->batch, r->hp, &slot);
        std::lock_guard<std::mutex> lk(S->mu);
        r->hp = HostPrep();  // inputs are on the device (or rejected)
        if (rc != MPS_OK) {
            r->state = MPS_REQ_FAILED;
            r->status = rc;
            r->t_done = Clock::now();
            S->failed++;
            S->cv_done.notify_all();
            continue;
        }
        r->slot = slot;
        r->state = MPS_REQ_RUNNING;
        r->t_admit = Clock::now();
        S->running.push_back(r);
    }
}

static void worker_loop(mps_server_s* S) {
    (void)hipSetDevice(S->cfg.batch.device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->running.empty()) S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty(); });
            if (S->stop) {  // shutdown abandons queued and running solves
                for (auto& r : S->queue) {
                    r->state = MPS_REQ_FAILED;
                    r->status = MPS_ERR_STATE;
                }
                for (auto& r : S->running) {
                    r->state = MPS_REQ_FAILED;
                    r->status = MPS_ERR_STATE;
                }
                S->queue.clear();
                S->running.clear();
                S->cv_done.notify_all();
                return;
            }
        }
        if (S->running.empty()) {  // nothing on the device yet: admit, then launch next round
            admit(S);
            continue;
        }
        const auto t0 = Clock::now();
        const int bucket = tune_before_launch(S);
        const double work = mps_batch_launch_work(S->batch, 1);
        {
            int bs = 0, gm = 0;
            mps_batch_tuning(S->batch, &bs, &gm);
            std::lock_guard<std::mutex> lk(S->mu);
            S->cur_bs = bs;
            S->cur_gm = gm;
        }
        if (mps_batch_launch(S->batch, 1) != MPS_OK) {
            // the running solves fail: their slots and arena blocks are released once nothing queued
            // can still read them, and each counts as failed so that drain's predicate can hold
            const std::string why = mps_batch_last_error(S->batch);
            (void)mps_batch_sync(S->batch);
            for (auto& r : S->running) (void)mps_batch_remove(S->batch, r->slot);
            std::lock_guard<std::mutex> lk(S->mu);
            S->err = why;
            const auto now = Clock::now();
            for (auto& r : S->running) {
                r->state = MPS_REQ_FAILED;
                r->status = MPS_ERR_HIP;
                r->t_done = now;
                S->failed++;
            }
            if (!S->running.empty()) S->t_last = now;
            S->running.clear();
            S->cv_done.notify_all();
            continue;
        }
        S->launches++;
        admit(S);  // overlaps with the launch in flight
        (void)mps_batch_sync(S->batch);
        S->busy_ms += ms_between(t0, Clock::now());
        tune_after_launch(S, bucket, work, mps_batch_last_launch_ms(S->batch));
        retire(S);
    }
}

extern "C" int mps_server_create(const mps_server_cfg* cfg, mps_server* out) {
    if (!cfg || !out) return MPS_ERR_ARG;
    *out = nullptr;
    auto* S = new mps_server_s();
    S->cfg = *cfg;
    S->tuner.on = cfg->batch.elem_block == 0 && cfg->batch.grid_mult == 0;
    int rc = mps_batch_create(&cfg->batch, &S->batch);
    if (rc != MPS_OK) {
        delete S;
        return rc;
    }
    S->worker = std::thread(worker_loop, S);
    *out = S;
    return MPS_OK;
}

extern "C" int mps_server_submit(mps_server S, const mps_mesh* mesh, const mps_material* mat, const mps_bcs* bcs,
                                 const mps_solver_cfg* cfg, int64_t* id) {
    if (!S || !id) return MPS_ERR_ARG;
    auto r = std::make_shared<Request>();
    r->t_submit = Clock::now();
    std::string why;
    int rc = prepare(mesh, mat, bcs, cfg, &r->hp, &why);
    if (rc == MPS_OK) rc = mps_batch_check(S->batch, r->hp, &why);
    if (rc != MPS_OK) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->err = why;
        return rc;
    }
    r->footprint = mps_batch_footprint(S->batch, r->hp);
    r->n_nodes = r->hp.N;
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return MPS_ERR_STATE;
    if ((int64_t)r->footprint > S->cfg.batch.arena_bytes) {
        S->err = fmt("request footprint %zu exceeds the arena", r->footprint);
        return MPS_ERR_CAPACITY;
    }
    r->id = S->next_id++;
    *id = r->id;
    if (!S->have_first) {
        S->have_first = true;
        S->t_first = r->t_submit;
    }
    S->submitted++;
    S->all[r->id] = r;
    S->queue.push_back(r);
    S->cv_work.notify_one();
    return MPS_OK;
}

extern "C" int mps_server_poll(mps_server S, int64_t id, int32_t* state, double* u, double* lat_ms) {
    if (!S) return MPS_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return MPS_ERR_ARG;
    auto& r = it->second;
    if (state) *state = r->state;
    if (r->state == MPS_REQ_DONE || r->state == MPS_REQ_FAILED) {
        if (u && r->state == MPS_REQ_DONE) std::copy(r->u.begin(), r->u.end(), u);
        if (lat_ms) {
            lat_ms[0] = ms_between(r->t_submit, r->t_done);
            lat_ms[1] = r->state == MPS_REQ_DONE ? ms_between(r->t_submit, r->t_admit) : 0.0;
        }
        const int32_t st = r->status;
        if (u) S->all.erase(it);  // collected
        return st;
    }
    return MPS_OK;
}

extern "C" int mps_server_wait(mps_server S, int64_t id, double timeout_s, int32_t* state) {
    if (!S) return MPS_ERR_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return MPS_ERR_ARG;
    auto r = it->second;
    S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s),
                        [&] { return r->state == MPS_REQ_DONE || r->state == MPS_REQ_FAILED; });
    if (state) *state = r->state;
    return MPS_OK;
}

extern "C" int mps_server_drain(mps_server S, double timeout_s) {
    if (!S) return MPS_ERR_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    const bool ok = S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s),
                                        [&] { return S->completed + S->failed >= S->submitted; });
    return ok ? MPS_OK : MPS_ERR_STATE;
}

extern "C" int mps_server_get_stats(mps_server S, mps_server_stats* out) {
    if (!S || !out) return MPS_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    out->submitted = S->submitted;
    out->completed = S->completed;
    out->failed = S->failed;
    out->launches = S->launches;
    out->p50_ms = percentile(S->lat_ms, 0.50);
    out->p99_ms = percentile(S->lat_ms, 0.99);
    out->max_ms = S->lat_ms.empty() ? 0.0 : *std::max_element(S->lat_ms.begin(), S->lat_ms.end());
    out->queue_p50_ms = percentile(S->queue_ms, 0.50);
    const double span = S->have_first && S->completed ? ms_between(S->t_first, S->t_last) : 0.0;
    out->span_ms = span;
    out->solves_per_s = span > 0 ? 1e3 * (double)S->completed / span : 0.0;
    out->busy_ms = S->busy_ms;
    out->elem_block = S->cur_bs;
    out->grid_mult = S->cur_gm;
    out->tunings = S->tuner.tunings;
    out->_pad = 0;
    return MPS_OK;
}

extern "C" int mps_server_reset_stats(mps_server S) {
    if (!S) return MPS_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    S->lat_ms.clear();
    S->queue_ms.clear();
    S->submitted -= S->completed + S->failed;
    S->completed = S->failed = S->launches = 0;
    S->have_first = false;
    S->busy_ms = 0.0;
    return MPS_OK;
}

extern "C" int mps_server_destroy(mps_server S) {
    if (!S) return MPS_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv_work.notify_all();
    }
    if (S->worker.joinable()) S->worker.join();
    (void)mps_batch_destroy(S->batch);
    delete S;
    return MPS_OK;
}

extern "C" const char* mps_server_last_error(mps_server S) { return S ? S->err.c_str() : ""; }
