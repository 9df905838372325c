// The served loop: flock-simulation requests (birds, steps, seed) arrive at
// any time; a worker thread keeps every admitted flock resident in one batch
// (fh_batch), admits waiting requests whenever slots and HBM bytes free up,
// advances all active flocks together a few steps per iteration, and retires
// each flock when its step count is reached. Generation of a request's start
// state and the checksum of its final state run on a small host pool, so they
// overlap the device work.
//
// No counterpart in the reference (one flock per process, cuda/flock.cu:374).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch.h"
#include "host_util.h"
#include "model.h"

using namespace fh;
using Clock = std::chrono::steady_clock;

namespace {

// Records of one flock. Elements are default-initialised (left as they are):
// every word is written before it is read, and a zero fill of a fresh
// multi-megabyte buffer would fault in its pages on the serving thread.
template <class T>
struct NoFill : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoFill<U>;
    };
    NoFill() = default;
    template <class U>
    NoFill(const NoFill<U> &) noexcept {}
    template <class U>
    void construct(U *p) noexcept {
        ::new (static_cast<void *>(p)) U;
    }
    template <class U, class... A>
    void construct(U *p, A &&...a) {
        ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
    }
};
using Words = std::vector<float, NoFill<float>>;

// Page-locked host buffers for the asynchronous uploads and downloads of a
// resident batch, recycled by size class (powers of two from 1 MiB): the
// pinning cost is paid once per buffer, not per request.
class PinPool {
   public:
    ~PinPool() {
        for (auto &kv : free_)
            for (void *p : kv.second) (void)hipHostFree(p);
    }
    // A buffer of at least `bytes` (nullptr if pinning fails); *cls is its class.
    void *get(size_t bytes, size_t *cls) {
        size_t c = size_t(1) << 20;
        while (c < bytes) c <<= 1;
        *cls = c;
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto &v = free_[c];
            if (!v.empty()) {
                void *p = v.back();
                v.pop_back();
                return p;
            }
        }
        void *p = nullptr;
        if (hipHostMalloc(&p, c, hipHostMallocDefault) != hipSuccess) return nullptr;
        return p;
    }
    void put(void *p, size_t cls) {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu_);
        free_[cls].push_back(p);
    }

   private:
    std::mutex mu_;
    std::map<size_t, std::vector<void *>> free_;
};

// Recycled events marking the end of a flock's download.
class EventPool {
   public:
    ~EventPool() {
        for (hipEvent_t e : free_) (void)hipEventDestroy(e);
    }
    hipEvent_t get() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (!free_.empty()) {
                hipEvent_t e = free_.back();
                free_.pop_back();
                return e;
            }
        }
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
        return e;
    }
    void put(hipEvent_t e) {
        if (!e) return;
        std::lock_guard<std::mutex> lk(mu_);
        free_.push_back(e);
    }

   private:
    std::mutex mu_;
    std::vector<hipEvent_t> free_;
};

struct Job {
    fh_result res{};
    Words words;  // start state, then final state when kept (canonical batches)
    void *stage = nullptr;     // resident batches: the start state staged in pinned memory
    size_t stage_cls = 0;
    bool ready = false;        // start state generated
    uint32_t flock = 0;        // batch slot while active
    uint32_t left = 0;         // steps still to run
};

// FNV-1a 64 over the 32-bit words of the records (one xor and one multiply
// per word).
uint64_t fnv1a_words(const float *w, size_t count) {
    uint64_t h = 1469598103934665603ull;
    const uint32_t *p = reinterpret_cast<const uint32_t *>(w);
    for (size_t i = 0; i < count; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

// A retired flock whose checksum is still to be computed. In a resident
// batch its records are still on their way to `out` (pinned, staged form)
// until `done` completes.
struct Finish {
    fh_result res;
    Words words;
    void *out = nullptr;
    size_t out_cls = 0;
    hipEvent_t done = nullptr;
};

}  // namespace

struct fh_server {
    fh_server_config cfg{};
    fh_batch *batch = nullptr;
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_prep, cv_done;
    std::deque<uint64_t> queue;                  // submitted, in arrival order
    std::deque<uint64_t> prep_queue;             // waiting for generation
    std::deque<Finish> finish_queue;             // retired, checksum pending
    uint32_t finishing = 0;                      // finish tasks not yet in `done`
    std::unordered_map<uint64_t, Job> jobs;
    std::vector<uint64_t> active;                // in the batch
    std::deque<fh_result> done;
    struct Kept {
        Words words;
        float time;
    };
    std::unordered_map<uint64_t, Kept> kept;  // final states (keep_states)
    uint64_t next_id = 1;
    double bytes_active = 0.0;
    bool stop = false;
    fh_status fatal = FH_OK;
    std::string fatal_msg;
    fh_server_stats st{};
    std::thread worker;
    std::vector<std::thread> prep;
    PinPool pins;
    EventPool events;
    bool staged_mode = false;   // resident batch: staged starts, asynchronous joins and leaves
    uint32_t staged_waiting = 0;  // start states staged and not yet admitted
    uint32_t joins_in_flight = 0;  // admitted, joining the batch at the next apply

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

static double flock_bytes_d(const fh_server *s, uint32_t n) { return (double)fh_flock_bytes(n, &s->cfg.params); }

// Start states staged ahead of admission (resident batches): enough to fill
// every slot, so pinned memory stays bounded while a long queue waits.
static uint32_t stage_ahead(const fh_server *s) { return s->cfg.max_active; }

static void prep_loop(fh_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_prep.wait(lk, [&] {
            return s->stop || !s->finish_queue.empty() ||
                   (!s->prep_queue.empty() && (!s->staged_mode || s->staged_waiting < stage_ahead(s)));
        });
        if (s->stop) return;
        if (!s->finish_queue.empty()) {  // results first: they end a request's latency
            Finish f = std::move(s->finish_queue.front());
            s->finish_queue.pop_front();
            lk.unlock();
            if (f.out) {  // wait for the download, then the records in .flk form
                const hipError_t e = hipEventSynchronize(f.done);
                if (e == hipSuccess) {
                    f.words.resize((size_t)f.res.birds * FH_WORDS);
                    unstage_words(f.out, f.res.birds, f.words.data());
                } else {
                    f.res.status = FH_ERR_HIP;
                }
                s->pins.put(f.out, f.out_cls);
                s->events.put(f.done);
            }
            f.res.checksum = fnv1a_words(f.words.data(), f.words.size());
            lk.lock();
            f.res.done_s = s->now();
            if (s->cfg.keep_states) s->kept[f.res.id] = fh_server::Kept{std::move(f.words), f.res.final_time};
            s->done.push_back(f.res);
            s->finishing -= 1;
            s->cv_done.notify_all();
            continue;
        }
        const uint64_t id = s->prep_queue.front();
        s->prep_queue.pop_front();
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) continue;
        const uint32_t n = it->second.res.birds;
        const uint64_t seed = it->second.res.seed;
        const bool staged = s->staged_mode;
        if (staged) s->staged_waiting += 1;
        lk.unlock();
        Words w((size_t)n * FH_WORDS);
        fh_gen_request(n, seed, s->cfg.birds_per_cell, s->cfg.cell_for_density, w.data());
        void *stage = nullptr;
        size_t cls = 0;
        if (staged && (stage = s->pins.get(staged_size(n), &cls))) {
            stage_words(w.data(), n, stage);
            Words().swap(w);
        }
        lk.lock();
        it = s->jobs.find(id);
        if (it != s->jobs.end()) {
            if (staged && !stage) {
                it->second.res.status = FH_ERR_NOMEM;  // the worker fails the request
                s->staged_waiting -= 1;
            }
            it->second.words = std::move(w);
            it->second.stage = stage;
            it->second.stage_cls = cls;
            it->second.ready = true;
            s->cv_work.notify_all();
        } else {
            s->pins.put(stage, cls);
            if (staged) s->staged_waiting -= 1;
        }
    }
}

static void fail_server(fh_server *s, fh_status st) {
    s->fatal = st;
    s->fatal_msg = fh_last_error();
}

// Resident batches take per-flock step budgets (fh_batch_limit).
static bool budgets(const fh_server *s) { return (s->cfg.params.flags & FH_FLAG_RESIDENT) != 0; }

// Canonical batches: admissions and retirements between step batches, each a
// synchronous add, download and remove.
static void worker_loop_sync(fh_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    const double busy0 = s->now();
    (void)busy0;
    while (!s->stop && s->fatal == FH_OK) {
        // Admit in arrival order while a slot, the HBM budget and pool room allow.
        while (!s->queue.empty() && s->active.size() < s->cfg.max_active) {
            Job &j = s->jobs[s->queue.front()];
            if (!j.ready) break;  // head of line still being generated
            const double bytes = flock_bytes_d(s, j.res.birds);
            if (s->bytes_active + bytes > s->cfg.hbm_budget_bytes) break;
            uint32_t fid = 0;
            const double t_add = s->now();
            fh_status st = fh_batch_add(s->batch, j.words.data(), j.res.birds, 0.0f, &fid);
            s->st.admit_s += s->now() - t_add;
            if (st == FH_ERR_BUDGET && !s->active.empty()) break;  // pools fragmented: wait for room
            if (st) {
                j.res.status = st;
                j.res.done_s = s->now();
                s->done.push_back(j.res);
                s->jobs.erase(s->queue.front());
                s->queue.pop_front();
                s->cv_done.notify_all();
                continue;
            }
            j.flock = fid;
            j.left = j.res.steps;
            if (budgets(s) && (st = fh_batch_limit(s->batch, fid, j.left))) {
                fail_server(s, st);
                break;
            }
            j.res.admit_s = s->now();
            if (!s->cfg.keep_states) Words().swap(j.words);
            s->bytes_active += bytes;
            s->active.push_back(s->queue.front());
            s->queue.pop_front();
            s->st.peak_active = std::max<uint32_t>(s->st.peak_active, (uint32_t)s->active.size());
            s->st.peak_bytes = std::max(s->st.peak_bytes, s->bytes_active);
        }
        if (s->active.empty()) {
            s->cv_work.wait_for(lk, std::chrono::milliseconds(2));
            continue;
        }
        // Step every active flock together. With step budgets (resident
        // batch) each flock freezes on its own last step and the iteration
        // runs its full length; otherwise it stops at the first completion.
        uint32_t k = s->cfg.steps_per_iter ? s->cfg.steps_per_iter : 8;
        if (!budgets(s))
            for (uint64_t id : s->active) k = std::min(k, s->jobs[id].left);
        double bird_steps = 0.0;
        for (uint64_t id : s->active)
            bird_steps += (double)s->jobs[id].res.birds * std::min(k, s->jobs[id].left);
        lk.unlock();
        const double t_a = s->now();
        fh_status st = batch_step(s->batch, k, 0, nullptr);
        const double t_b = s->now();
        lk.lock();
        if (st) {
            fail_server(s, st);
            break;
        }
        s->st.busy_s += t_b - t_a;
        s->st.iterations += 1;
        s->st.bird_steps += bird_steps;
        s->st.steps += k;
        // Retire finished flocks.
        for (size_t i = 0; i < s->active.size();) {
            Job &j = s->jobs[s->active[i]];
            j.left -= std::min(k, j.left);
            if (j.left) {
                ++i;
                continue;
            }
            const double t_ret = s->now();
            Words out((size_t)j.res.birds * FH_WORDS);
            float t = 0.0f;
            st = fh_batch_download(s->batch, j.flock, out.data(), &t);
            if (!st) st = fh_batch_remove(s->batch, j.flock);
            s->st.retire_s += s->now() - t_ret;
            if (st) {
                fail_server(s, st);
                break;
            }
            j.res.final_time = t;
            j.res.status = FH_OK;
            s->bytes_active -= flock_bytes_d(s, j.res.birds);
            s->finish_queue.push_back(Finish{j.res, std::move(out)});
            s->finishing += 1;
            s->cv_prep.notify_one();
            s->jobs.erase(s->active[i]);
            s->active.erase(s->active.begin() + (long)i);
        }
    }
    s->cv_done.notify_all();
}

// A request that cannot be served: its result goes out with `st`.
static void fail_job(fh_server *s, uint64_t id, fh_status st) {
    Job &j = s->jobs[id];
    j.res.status = st;
    j.res.done_s = s->now();
    s->done.push_back(j.res);
    if (j.stage) s->pins.put(j.stage, j.stage_cls);
    s->jobs.erase(id);
    s->cv_done.notify_all();
}

// Resident batches: a pipeline over two streams. Each iteration applies the
// joins and leaves queued during the previous one (one splice and one scan,
// the only GPU-idle part), launches the steps, and while they run admits the
// next flocks: their staged start states (prepared by the prep threads in
// pinned memory) upload on the copy stream. When the steps end, the flocks
// that finished are flushed and their records download on the copy stream
// during the next iteration; the prep threads turn them into results.
static void worker_loop_resident(fh_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    std::vector<uint64_t> joining;  // uploads queued; in the layout after the next apply
    std::vector<std::pair<void *, size_t>> applied;  // staging of the joins of the last apply
    const uint32_t k_iter = s->cfg.steps_per_iter ? s->cfg.steps_per_iter : 8;
    double last_end = -1.0;  // when the last step batch ended (-1: the GPU had nothing to run)
    // Admits queued requests in arrival order while a slot, the HBM budget and
    // pool room allow; `leaving` counts the slots and bytes of flocks that
    // finish with the steps in flight.
    // (It runs while steps are in flight: its time is not GPU-idle time.)
    auto admit = [&](uint32_t leaving_slots, double leaving_bytes) {
        while (!s->queue.empty() && s->fatal == FH_OK) {
            const uint64_t id = s->queue.front();
            Job &j = s->jobs[id];
            if (!j.ready) break;  // head of line still being generated
            if (j.res.status == FH_ERR_NOMEM) {
                s->queue.pop_front();
                fail_job(s, id, FH_ERR_NOMEM);
                continue;
            }
            if (s->active.size() - leaving_slots + joining.size() >= s->cfg.max_active) break;
            const double bytes = flock_bytes_d(s, j.res.birds);
            if (s->bytes_active - leaving_bytes + bytes > s->cfg.hbm_budget_bytes) break;
            uint32_t fid = 0;
            lk.unlock();
            fh_status st = batch_join_async(s->batch, j.stage, j.res.birds, 0.0f, j.res.steps, &fid);
            lk.lock();
            if (st == FH_ERR_BUDGET && !(s->active.empty() && joining.empty())) break;  // wait for room
            s->queue.pop_front();
            s->staged_waiting -= 1;
            s->cv_prep.notify_all();
            if (st) {
                fail_job(s, id, st);
                continue;
            }
            j.flock = fid;
            j.left = j.res.steps;
            j.res.admit_s = s->now();
            s->bytes_active += bytes;
            joining.push_back(id);
        }
        s->joins_in_flight = (uint32_t)joining.size();
    };
    while (!s->stop && s->fatal == FH_OK) {
        // GPU idle from here to the launch of the steps.
        const double g0 = s->now();
        const bool had_joins = !joining.empty();
        lk.unlock();
        fh_status st = batch_apply(s->batch);
        lk.lock();
        if (st) {
            fail_server(s, st);
            break;
        }
        const double g1 = s->now();
        if (had_joins) s->st.admit_s += g1 - g0;
        else s->st.retire_s += g1 - g0;
        for (uint64_t id : joining) {
            Job &j = s->jobs[id];
            applied.emplace_back(j.stage, j.stage_cls);
            j.stage = nullptr;
            s->active.push_back(id);
        }
        joining.clear();
        s->joins_in_flight = 0;
        s->st.peak_active = std::max<uint32_t>(s->st.peak_active, (uint32_t)s->active.size());
        s->st.peak_bytes = std::max(s->st.peak_bytes, s->bytes_active);
        if (s->active.empty()) {
            last_end = -1.0;
            admit(0, 0.0);
            if (joining.empty()) s->cv_work.wait_for(lk, std::chrono::milliseconds(2));
            continue;
        }
        uint32_t leaving = 0;
        double leaving_bytes = 0.0, bird_steps = 0.0;
        for (uint64_t id : s->active) {
            const Job &j = s->jobs[id];
            bird_steps += (double)j.res.birds * std::min(k_iter, j.left);
            if (j.left <= k_iter) {
                leaving += 1;
                leaving_bytes += flock_bytes_d(s, j.res.birds);
            }
        }
        lk.unlock();
        const double t_a = s->now();
        st = batch_step_async(s->batch, k_iter);
        lk.lock();
        if (last_end >= 0.0) s->st.gap_s += t_a - last_end;
        if (st) {
            fail_server(s, st);
            break;
        }
        admit(leaving, leaving_bytes);  // overlaps the steps
        lk.unlock();
        st = batch_wait(s->batch);
        const double t_b = s->now();
        lk.lock();
        if (st) {
            fail_server(s, st);
            break;
        }
        s->st.busy_s += t_b - t_a;
        s->st.iterations += 1;
        s->st.bird_steps += bird_steps;
        s->st.steps += k_iter;
        last_end = t_b;
        // the apply before these steps waited for its joins' uploads
        for (auto &pc : applied) s->pins.put(pc.first, pc.second);
        applied.clear();
        // Retire the flocks whose budget is spent: flush and download queued.
        for (size_t i = 0; i < s->active.size();) {
            Job &j = s->jobs[s->active[i]];
            j.left -= std::min(k_iter, j.left);
            if (j.left) {
                ++i;
                continue;
            }
            Finish f;
            f.res = j.res;
            f.out = s->pins.get(staged_size(j.res.birds), &f.out_cls);
            f.done = s->events.get();
            float t = 0.0f;
            lk.unlock();
            st = (f.out && f.done) ? batch_leave_async(s->batch, j.flock, f.out, &t, f.done)
                                   : set_error(FH_ERR_NOMEM, "served loop: no pinned buffer or event for a result");
            lk.lock();
            if (st) {
                s->pins.put(f.out, f.out_cls);
                s->events.put(f.done);
                fail_server(s, st);
                break;
            }
            f.res.final_time = t;
            f.res.status = FH_OK;
            s->bytes_active -= flock_bytes_d(s, j.res.birds);
            s->finish_queue.push_back(std::move(f));
            s->finishing += 1;
            s->jobs.erase(s->active[i]);
            s->active.erase(s->active.begin() + (long)i);
        }
        s->st.retire_s += s->now() - t_b;
        s->cv_prep.notify_all();
    }
    // Whatever was staged or in flight: wait for the copies before the buffers go.
    lk.unlock();
    (void)batch_wait(s->batch);
    lk.lock();
    for (auto &pc : applied) s->pins.put(pc.first, pc.second);
    s->cv_done.notify_all();
}

static void worker_loop(fh_server *s) {
    if (s->staged_mode)
        worker_loop_resident(s);
    else
        worker_loop_sync(s);
}

extern "C" {

fh_status fh_server_config_default(fh_server_config *c) {
    clear_error();
    if (!c) return set_error(FH_ERR_INVALID, "fh_server_config_default: NULL");
    std::memset(c, 0, sizeof *c);
    c->hbm_budget_bytes = 200e9;
    c->max_active = 64;
    c->bird_capacity = 40u << 20;
    c->bucket_capacity = 64u << 20;
    c->steps_per_iter = 8;
    c->prep_threads = 4;
    c->birds_per_cell = 0.64f;
    c->cell_for_density = 4.0f;
    fh_params_default(&c->params);
    c->params.flags = FH_FLAG_CANONICAL;
    return FH_OK;
}

fh_status fh_server_create(const fh_server_config *cfg, fh_server **out) {
    clear_error();
    if (!cfg || !out) return set_error(FH_ERR_INVALID, "fh_server_create: NULL argument");
    *out = nullptr;
    if (cfg->max_active == 0 || !(cfg->hbm_budget_bytes > 0) || !(cfg->birds_per_cell > 0) ||
        !(cfg->cell_for_density > 0))
        return set_error(FH_ERR_INVALID, "fh_server_create: bad config");
    fh_server *s = new fh_server();
    s->cfg = *cfg;
    fh_params prm = cfg->params;
    prm.flags &= ~FH_FLAG_GRAPH;  // membership changes every few steps: launch eagerly
    fh_status st = fh_batch_create(cfg->bird_capacity, cfg->bucket_capacity, cfg->max_active, &prm, &s->batch);
    if (st) {
        delete s;
        return st;
    }
    s->t0 = Clock::now();
    s->staged_mode = budgets(s);
    s->worker = std::thread(worker_loop, s);
    const uint32_t np = cfg->prep_threads ? cfg->prep_threads : 1;
    for (uint32_t i = 0; i < np; ++i) s->prep.emplace_back(prep_loop, s);
    *out = s;
    return FH_OK;
}

fh_status fh_server_destroy(fh_server *s) {
    clear_error();
    if (!s) return FH_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_prep.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (auto &t : s->prep)
        if (t.joinable()) t.join();
    // staged starts never admitted, and results still on their way
    for (auto &kv : s->jobs) s->pins.put(kv.second.stage, kv.second.stage_cls);
    for (Finish &f : s->finish_queue) {
        if (f.done) (void)hipEventSynchronize(f.done);
        s->pins.put(f.out, f.out_cls);
        s->events.put(f.done);
    }
    fh_batch_destroy(s->batch);
    delete s;
    return FH_OK;
}

fh_status fh_server_submit(fh_server *s, uint32_t birds, uint32_t steps, uint64_t seed, uint64_t *id) {
    clear_error();
    if (!s) return set_error(FH_ERR_INVALID, "fh_server_submit: NULL server");
    if (birds == 0 || steps == 0) return set_error(FH_ERR_INVALID, "fh_server_submit: birds and steps must be > 0");
    if (flock_bytes_d(s, birds) > s->cfg.hbm_budget_bytes || birds > s->cfg.bird_capacity)
        return set_error(FH_ERR_BUDGET, "fh_server_submit: %u birds can never fit the HBM budget", birds);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->fatal) return set_error(s->fatal, "fh_server_submit: server stopped: %s", s->fatal_msg.c_str());
    const uint64_t rid = s->next_id++;
    Job j;
    j.res.id = rid;
    j.res.birds = birds;
    j.res.steps = steps;
    j.res.seed = seed;
    j.res.arrival_s = s->now();
    j.res.status = -1;
    s->jobs.emplace(rid, std::move(j));
    s->queue.push_back(rid);
    s->prep_queue.push_back(rid);
    s->st.submitted += 1;
    s->cv_prep.notify_one();
    if (id) *id = rid;
    return FH_OK;
}

fh_status fh_server_poll(fh_server *s, fh_result *out, uint32_t cap, uint32_t *count) {
    clear_error();
    if (!s || !count || (cap && !out)) return set_error(FH_ERR_INVALID, "fh_server_poll: bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    uint32_t k = 0;
    while (k < cap && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.pop_front();
    }
    *count = k;
    if (s->fatal) return set_error(s->fatal, "server stopped: %s", s->fatal_msg.c_str());
    return FH_OK;
}

fh_status fh_server_drain(fh_server *s, double timeout_s) {
    clear_error();
    if (!s) return set_error(FH_ERR_INVALID, "fh_server_drain: NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    const auto until = Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(timeout_s));
    while (s->fatal == FH_OK && (!s->queue.empty() || !s->active.empty() || s->joins_in_flight || s->finishing)) {
        if (s->cv_done.wait_until(lk, until) == std::cv_status::timeout)
            return set_error(FH_ERR_STATE, "fh_server_drain: timed out with %zu queued, %zu active",
                             s->queue.size(), s->active.size());
    }
    if (s->fatal) return set_error(s->fatal, "server stopped: %s", s->fatal_msg.c_str());
    return FH_OK;
}

fh_status fh_server_stats_get(fh_server *s, fh_server_stats *out) {
    clear_error();
    if (!s || !out) return set_error(FH_ERR_INVALID, "fh_server_stats_get: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->wall_s = s->now();
    out->queued = (uint32_t)s->queue.size();
    out->active = (uint32_t)s->active.size();
    return FH_OK;
}

fh_status fh_server_take_state(fh_server *s, uint64_t id, float *words, uint32_t cap, float *time) {
    clear_error();
    if (!s || !words) return set_error(FH_ERR_INVALID, "fh_server_take_state: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->kept.find(id);
    if (it == s->kept.end()) return set_error(FH_ERR_INVALID, "fh_server_take_state: no kept state for %llu",
                                              (unsigned long long)id);
    const uint32_t n = (uint32_t)(it->second.words.size() / FH_WORDS);
    if (n > cap) return set_error(FH_ERR_INVALID, "fh_server_take_state: %u birds exceed %u", n, cap);
    std::memcpy(words, it->second.words.data(), it->second.words.size() * 4);
    if (time) *time = it->second.time;
    s->kept.erase(it);
    return FH_OK;
}

// Open-loop replay: request i is submitted `arrival_s[i]` seconds after the
// start, whatever the server is doing; returns when every request is done.
fh_status fh_server_replay(const fh_server_config *cfg, const uint32_t *birds, const uint32_t *steps,
                           const uint64_t *seeds, const double *arrival_s, uint32_t count, fh_result *results,
                           fh_server_stats *stats) {
    clear_error();
    if (!cfg || (count && (!birds || !steps || !seeds || !arrival_s || !results)))
        return set_error(FH_ERR_INVALID, "fh_server_replay: NULL argument");
    fh_server *s = nullptr;
    fh_status st = fh_server_create(cfg, &s);
    if (st) return st;
    std::vector<uint64_t> ids(count, 0);
    std::unordered_map<uint64_t, uint32_t> index;
    for (uint32_t i = 0; i < count && st == FH_OK; ++i) {
        const auto at = s->t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival_s[i]));
        std::this_thread::sleep_until(at);
        st = fh_server_submit(s, birds[i], steps[i], seeds[i], &ids[i]);
        if (st == FH_OK) index[ids[i]] = i;
        else {
            results[i] = fh_result{};
            results[i].status = st;
            st = FH_OK;
        }
    }
    if (st == FH_OK) st = fh_server_drain(s, 3600.0);
    std::vector<fh_result> buf(256);
    uint32_t got = 0;
    do {
        fh_status pst = fh_server_poll(s, buf.data(), (uint32_t)buf.size(), &got);
        if (pst && st == FH_OK) st = pst;
        for (uint32_t k = 0; k < got; ++k) {
            auto it = index.find(buf[k].id);
            if (it != index.end()) results[it->second] = buf[k];
        }
    } while (got);
    if (stats) fh_server_stats_get(s, stats);
    std::string msg = fh_last_error();
    fh_server_destroy(s);
    if (st) return set_error(st, "%s", msg.c_str());
    return FH_OK;
}

}  // extern "C"
