// The batched multi-flock engine shared by fh_batch_* and the served loop.
#pragma once

#include <condition_variable>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/flock_hip.h"
#include "layout.h"

namespace fh {

// First-fit allocator over [0, capacity) in whole units (tiles of birds or
// tiles of buckets), with coalescing of neighbouring free runs.
class RangeAlloc {
   public:
    static constexpr uint32_t NONE = 0xffffffffu;
    void reset(uint32_t capacity) {
        free_.clear();
        used_.clear();
        if (capacity) free_[0] = capacity;
    }
    uint32_t alloc(uint32_t len);
    void release(uint32_t at);
    uint32_t high_water() const;

   private:
    std::map<uint32_t, uint32_t> free_, used_;  // start -> length
};

// Splits body(lo, hi) over [0, n) between the calling thread and `helpers`
// threads that sleep on a condition variable between calls. One caller at a
// time (a batch is used by one thread at a time).
class ParFor {
   public:
    explicit ParFor(unsigned helpers);
    ~ParFor();
    void run(size_t n, const std::function<void(size_t, size_t)> &body);

   private:
    void loop(unsigned part);
    const unsigned parts_;
    std::mutex mu_;
    std::condition_variable go_, done_;
    const std::function<void(size_t, size_t)> *body_ = nullptr;
    size_t n_ = 0;
    uint64_t gen_ = 0;
    unsigned pending_ = 0;
    bool stop_ = false;
    std::vector<std::thread> th_;
};

size_t batch_bytes_per_bird();
size_t batch_bytes_per_bucket();
// Extra device bytes of the resident layout per bird and per bucket.
size_t resident_bytes_per_bird(bool flight);
size_t resident_bytes_per_bucket();

}  // namespace fh

struct fh_batch {
    fh_params prm{};
    uint32_t bird_cap = 0, bucket_cap_units = 0, flock_cap = 0;
    hipStream_t stream = nullptr;
    void *arena = nullptr;
    size_t arena_bytes = 0;
    fh::Pools P{};
    fh::Seg *d_segs = nullptr;
    uint32_t *d_tile_seg = nullptr;
    std::vector<fh::Seg> segs;        // host mirror of the segment table
    std::vector<float> times;         // host mirror of each flock clock
    std::vector<uint32_t> tile_seg;   // host mirror of tile -> segment
    fh::RangeAlloc tiles, bucket_units;
    fh::Geometry geo{};
    bool hashed = false;        // counts hold the histogram of every live bird
    bool counts_dirty = false;  // histogram filled and not yet consumed by a scan
    hipGraphExec_t graph = nullptr;
    bool graph_valid = false;
    // Resident layout (FH_FLAG_RESIDENT): every flock's state kept in bucket
    // order between steps, as for one flock (records R[parity], flown
    // records F for the fused flight, work arrays W, offsets per parity).
    void *res_arena = nullptr;
    fh::Records R[2]{}, F{};
    fh::ResWork W{};
    uint32_t *off[2]{};
    uint32_t *res_due = nullptr;  // due counter per parity, then the persistent decide's chunk counters
    int32_t *d_shift = nullptr;   // per flock: slot shift of a splice (flocks joining or leaving)
    bool resident = false;
    uint32_t parity = 0;
    hipGraphExec_t rgraph[2]{};
    bool rgraph_valid = false;
    // Lazy rebuilds (FH_FLAG_LAZY), as for one flock: the layout is rebuilt
    // every LAZY_PERIOD steps; in between each step flies the records in
    // place (alternating between R[parity]'s positions, velocities and roll
    // rates and F's) and lists the birds out of their slot's bucket. The
    // segment table keeps the clocks and budgets of the last rebuild; a lazy
    // step derives each flock's from them and its phase, and the rebuild
    // commits the period's steps. Membership, budget and download calls
    // first rebuild the layout of the current positions (settle).
    bool lazy = false;
    uint32_t phase = 0;  // steps since the last rebuild
    uint32_t fq = 0;     // 1: the current positions, velocities and roll rates are F's
    uint32_t cpar = 0;   // due-counter parity of the next lazy step
    fh::MoverLists ml;
    hipGraphExec_t lgraph[2][2]{};  // one period of lazy steps from phase 0: layout parity, counter parity
    // Pinned host staging for fh_batch_add / fh_batch_download (grown on
    // demand, reused): copies from pageable memory are staged by the runtime
    // and run at a fraction of the link's speed.
    void *pin = nullptr;
    size_t pin_bytes = 0;
    // Helper threads for the record conversions of large flocks (made on
    // first use).
    std::unique_ptr<fh::ParFor> par;
    // Joins and leaves queued for the next batch_apply (resident batches):
    // per segment 0, PENDING_JOIN or PENDING_LEAVE. Uploads and downloads of
    // the queued flocks run on `copy`, a second stream, while steps run on
    // `stream`; `up_done` follows the last upload, `flushed` the last flush.
    std::vector<uint8_t> pending;
    std::vector<uint32_t> joins, leaves;
    hipStream_t copy = nullptr;
    hipEvent_t up_done = nullptr, flushed = nullptr;
    bool uploads = false;  // an upload was queued since the last apply
    // Pinned staging of the tables an apply uploads (two generations of the
    // segment table and of tile -> segment, and the splice shifts), so its
    // copies are asynchronous.
    void *tbl = nullptr;
    size_t tbl_bytes = 0;
};

namespace fh {
fh_status batch_step(fh_batch *b, uint32_t steps, uint32_t warmup, fh_pass_ms *out);

// The host form of a flock's records as the pools hold them (structure of
// arrays: pos, vel, steer as float4, then spin, centrality, neighbours), the
// unit of the asynchronous joins and leaves below.
size_t staged_size(uint32_t n);
void stage_words(const float *words, uint32_t n, void *dst);
void unstage_words(const void *src, uint32_t n, float *words);

// Asynchronous membership changes of a resident batch, for the served loop.
// batch_join_async reserves the flock's ranges and queues the copy of its
// staged records (`src`, pinned host memory the caller keeps until the next
// batch_wait returns) on the copy stream; it may run while steps are in
// flight. batch_leave_async (no steps in flight) flushes the flock's slots to
// bird order and queues their copy into `dst` (pinned) on the copy stream,
// then records `done`; *time gets the flock's clock. batch_apply applies
// every queued join and leave with one splice and one scan (after the
// uploads), without waiting on the host. The flock takes at most `steps`
// steps (its budget).
fh_status batch_join_async(fh_batch *b, const void *src, uint32_t n, float time, uint32_t steps, uint32_t *fid);
fh_status batch_leave_async(fh_batch *b, uint32_t fid, void *dst, float *time, hipEvent_t done);
fh_status batch_apply(fh_batch *b);
// `steps` steps of every live flock, launched and not waited for; batch_wait
// waits for them and reports kernel errors.
fh_status batch_step_async(fh_batch *b, uint32_t steps);
fh_status batch_wait(fh_batch *b);
}
