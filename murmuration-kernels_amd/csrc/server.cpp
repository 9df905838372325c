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

struct Job {
    fh_result res{};
    Words words;  // start state, then final state when kept
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

// A retired flock whose checksum is still to be computed.
struct Finish {
    fh_result res;
    Words words;
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

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

static double flock_bytes_d(const fh_server *s, uint32_t n) { return (double)fh_flock_bytes(n, &s->cfg.params); }

static void prep_loop(fh_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_prep.wait(lk, [&] { return s->stop || !s->prep_queue.empty() || !s->finish_queue.empty(); });
        if (s->stop) return;
        if (!s->finish_queue.empty()) {  // results first: they end a request's latency
            Finish f = std::move(s->finish_queue.front());
            s->finish_queue.pop_front();
            lk.unlock();
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
        lk.unlock();
        Words w((size_t)n * FH_WORDS);
        fh_gen_request(n, seed, s->cfg.birds_per_cell, s->cfg.cell_for_density, w.data());
        lk.lock();
        it = s->jobs.find(id);
        if (it != s->jobs.end()) {
            it->second.words = std::move(w);
            it->second.ready = true;
            s->cv_work.notify_all();
        }
    }
}

static void fail_server(fh_server *s, fh_status st) {
    s->fatal = st;
    s->fatal_msg = fh_last_error();
}

// Resident batches take per-flock step budgets (fh_batch_limit).
static bool budgets(const fh_server *s) { return (s->cfg.params.flags & FH_FLAG_RESIDENT) != 0; }

static void worker_loop(fh_server *s) {
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
    while (s->fatal == FH_OK && (!s->queue.empty() || !s->active.empty() || s->finishing)) {
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
