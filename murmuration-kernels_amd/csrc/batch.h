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
    // Pinned host staging for fh_batch_add / fh_batch_download (grown on
    // demand, reused): copies from pageable memory are staged by the runtime
    // and run at a fraction of the link's speed.
    void *pin = nullptr;
    size_t pin_bytes = 0;
    // Helper threads for the record conversions of large flocks (made on
    // first use).
    std::unique_ptr<fh::ParFor> par;
};

namespace fh {
fh_status batch_step(fh_batch *b, uint32_t steps, uint32_t warmup, fh_pass_ms *out);
}
