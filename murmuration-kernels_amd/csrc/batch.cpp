// Many independent flocks stepped together: one launch per stage for all of
// them. Flocks share one set of device pools; flock f owns a run of 256-bird
// tiles and a run of buckets (its own buckets_for(n_f)), so the counting sort
// is segmented by (flock, bucket) simply by giving each flock its own bucket
// range in one global bucket space. Each flock keeps its own clock.
//
// The reference steps one flock per process (cuda/flock.cu:374-523); this is
// the batched growth on top of the same passes.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "batch.h"
#include "host_util.h"
#include "model.h"

namespace fh {
hipError_t launch_fill(uint32_t *dst, uint32_t count, uint32_t value, hipStream_t s);

// First-fit allocator over [0, capacity) in whole units.
uint32_t RangeAlloc::alloc(uint32_t len) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second >= len) {
            const uint32_t at = it->first, left = it->second - len;
            free_.erase(it);
            if (left) free_[at + len] = left;
            used_[at] = len;
            return at;
        }
    }
    return NONE;
}

void RangeAlloc::release(uint32_t at) {
    auto u = used_.find(at);
    if (u == used_.end()) return;
    uint32_t len = u->second;
    used_.erase(u);
    auto nx = free_.lower_bound(at);
    if (nx != free_.end() && at + len == nx->first) {
        len += nx->second;
        nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {
        auto pv = std::prev(nx);
        if (pv->first + pv->second == at) {
            pv->second += len;
            return;
        }
    }
    free_[at] = len;
}

uint32_t RangeAlloc::high_water() const { return used_.empty() ? 0u : used_.rbegin()->first + used_.rbegin()->second; }

}  // namespace fh

using namespace fh;

static fh_status bcheck(const fh_batch *b, const char *fn) {
    return b ? FH_OK : set_error(FH_ERR_INVALID, "%s: NULL batch", fn);
}

// Queued membership changes of a segment (fh_batch::pending).
constexpr uint8_t PENDING_JOIN = 1, PENDING_LEAVE = 2;

// A live flock that takes part in the steps: not one still waiting to join.
static bool in_layout(const fh_batch *b, uint32_t f) {
    return b->segs[f].live && (f >= b->pending.size() || b->pending[f] != PENDING_JOIN);
}

// Geometry of the live flocks: bird range up to the highest live tile, bucket
// range up to the highest live bucket.
static void refresh_geometry(fh_batch *b) {
    b->rgraph_valid = false;
    b->geo.segs = b->d_segs;
    b->geo.tile_seg = b->d_tile_seg;
    b->geo.n_birds = b->tiles.high_water() * TILE;
    b->geo.n_buckets = b->bucket_units.high_water() * SCAN_TILE;
    uint32_t slots = 0;
    for (uint32_t f = 0; f < b->segs.size(); ++f)
        if (in_layout(b, f)) slots += b->segs[f].n;
    b->geo.n_slots = slots;
    b->graph_valid = false;
}

static void set_slot_bases(fh_batch *b);

// Uploads the segment table and the tile -> segment entries [t_lo, t_hi).
static fh_status push_tables(fh_batch *b, uint32_t t_lo, uint32_t t_hi) {
    // The device advances each clock by DT per step; the host mirror in
    // times[] does the same f32 additions, so it is the current value.
    for (size_t f = 0; f < b->segs.size(); ++f) b->segs[f].time = b->times[f];
    set_slot_bases(b);
    if (!b->segs.empty())
        FH_HIP(hipMemcpyAsync(b->d_segs, b->segs.data(), b->segs.size() * sizeof(Seg), hipMemcpyHostToDevice,
                              b->stream));
    t_hi = std::min<uint32_t>(t_hi, (uint32_t)b->tile_seg.size());
    if (t_lo < t_hi)
        FH_HIP(hipMemcpyAsync(b->d_tile_seg + t_lo, b->tile_seg.data() + t_lo, (size_t)(t_hi - t_lo) * 4,
                              hipMemcpyHostToDevice, b->stream));
    FH_HIP(hipStreamSynchronize(b->stream));
    return FH_OK;
}

// Pinned staging of at least `bytes` (hipHostMalloc). The stream is
// synchronised first, so no copy of an earlier call still reads or writes the
// buffer when the caller fills it or when it is replaced by a larger one.
static fh_status pinned(fh_batch *b, size_t bytes) {
    FH_HIP(hipStreamSynchronize(b->stream));
    if (b->pin_bytes >= bytes) return FH_OK;
    if (b->pin) {
        (void)hipHostFree(b->pin);
        b->pin = nullptr;
        b->pin_bytes = 0;
    }
    const size_t want = std::max<size_t>(bytes, (size_t)1 << 20);
    hipError_t e = hipHostMalloc(&b->pin, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        b->pin = nullptr;
        return hip_fail(e, "hipHostMalloc(batch staging)");
    }
    b->pin_bytes = want;
    return FH_OK;
}

// The six per-bird arrays of n birds laid out one after another in the
// staging buffer (pos, vel, steer: float4; spin, cen: float; nb: uint32).
struct Staged {
    float4 *pos, *vel, *steer;
    float *spin, *cen;
    uint32_t *nb;
};
ParFor::ParFor(unsigned helpers) : parts_(helpers + 1) {
    for (unsigned k = 1; k <= helpers; ++k) th_.emplace_back([this, k] { loop(k); });
}

ParFor::~ParFor() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    go_.notify_all();
    for (auto &t : th_) t.join();
}

void ParFor::loop(unsigned part) {
    uint64_t seen = 0;
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        go_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        const size_t n = n_;
        const auto *body = body_;
        lk.unlock();
        (*body)(n * part / parts_, n * (part + 1) / parts_);
        lk.lock();
        if (--pending_ == 0) done_.notify_one();
    }
}

void ParFor::run(size_t n, const std::function<void(size_t, size_t)> &body) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        body_ = &body;
        n_ = n;
        pending_ = parts_ - 1;
        ++gen_;
    }
    go_.notify_all();
    body(0, n / parts_);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return pending_ == 0; });
}

// The .flk word <-> pool conversions of one flock: on the calling thread for
// small flocks, split four ways for large ones (they run while the GPU idles
// in the served loop's admissions and retirements).
static void convert(fh_batch *b, uint32_t n, const std::function<void(size_t, size_t)> &body) {
    if (n < (1u << 15)) {
        body(0, n);
        return;
    }
    if (!b->par) b->par = std::make_unique<ParFor>(3u);
    b->par->run(n, body);
}

static Staged staged(void *pin, uint32_t n) {
    char *q = static_cast<char *>(pin);
    Staged st;
    st.pos = reinterpret_cast<float4 *>(q);
    st.vel = st.pos + n;
    st.steer = st.vel + n;
    st.spin = reinterpret_cast<float *>(st.steer + n);
    st.cen = st.spin + n;
    st.nb = reinterpret_cast<uint32_t *>(st.cen + n);
    return st;
}
static size_t staged_bytes(uint32_t n) { return (size_t)n * (3 * sizeof(float4) + 3 * 4); }

static fh_status batch_hash(fh_batch *b) {
    if (b->counts_dirty) FH_HIP(launch_fill(b->P.counts, b->bucket_cap_units * SCAN_TILE, 0u, b->stream));
    FH_HIP(launch_hash_count(b->geo, b->P, (b->prm.flags & FH_FLAG_WAVE_AGG) != 0, b->stream));
    b->counts_dirty = true;
    b->hashed = true;
    return FH_OK;
}

// One batched step: every live flock's clock += DT on the device, then the
// passes over all flocks (relative due test, fused histogram).
static fh_status batch_step_launches(fh_batch *b, hipEvent_t *ev) {
    if (ev) FH_HIP(hipEventRecord(ev[0], b->stream));
    if (!b->hashed) {
        fh_status st = batch_hash(b);
        if (st) return st;
    }
    if (ev) FH_HIP(hipEventRecord(ev[1], b->stream));
    // the scan also advances every live flock's device clock by DT
    FH_HIP(launch_scan(b->geo, b->P, b->d_segs, (uint32_t)b->segs.size(), DT, b->stream));
    b->counts_dirty = false;
    if (ev) FH_HIP(hipEventRecord(ev[2], b->stream));
    FH_HIP(launch_scatter(b->geo, b->P, (b->prm.flags & FH_FLAG_CANONICAL) != 0, true, b->stream));
    if (ev) FH_HIP(hipEventRecord(ev[3], b->stream));
    // the decide launch records ev[6] and ev[7] at its kernel's start and end
    FH_HIP(launch_decide(b->geo, b->P, 0.0f, b->prm.decide_block, b->prm.birds_per_thread, b->prm.decide_group,
                         walk_flags(b->prm.flags), b->stream, ev ? DecideTiming{ev[6], ev[7]} : DecideTiming{}));
    if (ev) FH_HIP(hipEventRecord(ev[4], b->stream));
    FH_HIP(launch_integrate(b->geo, b->P, true, (b->prm.flags & FH_FLAG_WAVE_AGG) != 0, b->stream));
    b->counts_dirty = true;
    b->hashed = true;
    if (ev) FH_HIP(hipEventRecord(ev[5], b->stream));
    return FH_OK;
}

static fh_status batch_capture(fh_batch *b) {
    if (b->graph_valid) return FH_OK;
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    hipGraph_t g = nullptr;
    FH_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const bool h0 = b->hashed, d0 = b->counts_dirty;
    fh_status st = batch_step_launches(b, nullptr);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    b->hashed = h0;
    b->counts_dirty = d0;
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->graph = nullptr;
        return hip_fail(e, "hipGraphInstantiate");
    }
    b->graph_valid = true;
    return FH_OK;
}

// ---- resident layout ------------------------------------------------------

static fh_status batch_alloc_resident(fh_batch *b) {
    if (b->res_arena) return FH_OK;
    const size_t nn = b->bird_cap ? b->bird_cap : 1;
    const size_t nb = (size_t)b->bucket_cap_units * SCAN_TILE;
    const bool flight = (b->prm.flags & FH_FLAG_FLIGHT) != 0;
    Arena a;
    size_t o_p4[2], o_v4[2], o_s4[2], o_spin[2], o_id[2], o_key[2];
    for (int q = 0; q < 2; ++q) {
        o_p4[q] = a.take<float4>(nn);
        o_v4[q] = a.take<float4>(nn);
        o_s4[q] = a.take<float4>(nn);
        o_spin[q] = a.take<float>(nn);
        o_id[q] = a.take<uint32_t>(nn);
        o_key[q] = a.take<uint32_t>(nn);
    }
    const size_t o_arank = a.take<uint32_t>(nn);
    const size_t o_mbits = a.take<uint32_t>((nn + TILE - 1) / TILE * (TILE / 32));
    const size_t o_arr = a.take<uint32_t>(nb), o_abase = a.take<uint32_t>(nb);
    const size_t o_off1 = a.take<uint32_t>(nb + 1), o_due = a.take<uint32_t>(4);
    const size_t o_tsum = a.take<uint32_t>(nb / RES_TILE);
    const size_t o_tsum2 = a.take<uint32_t>(nb / RES_TILE / RES_SUPER + 1);
    const size_t o_fp4 = flight ? a.take<float4>(nn) : 0, o_fv4 = flight ? a.take<float4>(nn) : 0;
    const size_t o_fspin = flight ? a.take<float>(nn) : 0;
    const size_t o_shift = a.take<int32_t>(b->flock_cap ? b->flock_cap : 1);
    // mover lists (FH_FLAG_LAZY): about one list head per 8 birds, per phase
    const bool lazy = b->lazy;
    const uint32_t mb = lazy ? mover_heads(nn) : 4096u;
    const size_t o_mhead = lazy ? a.take<float4>((size_t)LAZY_PERIOD * mb) : 0;
    const size_t o_movf = lazy ? a.take<uint32_t>((size_t)LAZY_PERIOD * mb) : 0;  // right after: one fill clears both
    size_t o_mp4[2]{}, o_mv4[2]{}, o_mnext[2]{};
    for (int k = 0; lazy && k < 2; ++k) {
        o_mp4[k] = a.take<float4>(nn);
        o_mv4[k] = a.take<float4>(nn);
        o_mnext[k] = a.take<uint32_t>(nn);
    }
    hipError_t e = hipMalloc(&b->res_arena, a.used);
    if (e != hipSuccess) {
        b->res_arena = nullptr;
        return hip_fail(e, "hipMalloc(batch resident arena)");
    }
    char *base = static_cast<char *>(b->res_arena);
    for (int q = 0; q < 2; ++q) {
        b->R[q].p4 = reinterpret_cast<float4 *>(base + o_p4[q]);
        b->R[q].v4 = reinterpret_cast<float4 *>(base + o_v4[q]);
        b->R[q].s4 = reinterpret_cast<float4 *>(base + o_s4[q]);
        b->R[q].spin = reinterpret_cast<float *>(base + o_spin[q]);
        b->R[q].id = reinterpret_cast<uint32_t *>(base + o_id[q]);
        b->R[q].key = reinterpret_cast<uint32_t *>(base + o_key[q]);
    }
    b->W.arank = reinterpret_cast<uint32_t *>(base + o_arank);
    b->W.mbits = reinterpret_cast<uint32_t *>(base + o_mbits);
    b->W.arr = reinterpret_cast<uint32_t *>(base + o_arr);
    b->W.counts = b->P.counts;
    b->W.abase = reinterpret_cast<uint32_t *>(base + o_abase);
    b->W.tsum = reinterpret_cast<uint32_t *>(base + o_tsum);
    b->W.tsum2 = reinterpret_cast<uint32_t *>(base + o_tsum2);
    if (flight) {
        b->F.p4 = reinterpret_cast<float4 *>(base + o_fp4);
        b->F.v4 = reinterpret_cast<float4 *>(base + o_fv4);
        b->F.spin = reinterpret_cast<float *>(base + o_fspin);
    }
    b->off[0] = b->P.offsets;
    b->off[1] = reinterpret_cast<uint32_t *>(base + o_off1);
    b->res_due = reinterpret_cast<uint32_t *>(base + o_due);
    b->d_shift = reinterpret_cast<int32_t *>(base + o_shift);
    if (lazy) {
        b->ml.heads = mb;
        b->ml.head4 = reinterpret_cast<float4 *>(base + o_mhead);
        b->ml.ovf = reinterpret_cast<uint32_t *>(base + o_movf);
        for (int k = 0; k < 2; ++k) {
            b->ml.p4[k] = reinterpret_cast<float4 *>(base + o_mp4[k]);
            b->ml.v4[k] = reinterpret_cast<float4 *>(base + o_mv4[k]);
            b->ml.next[k] = reinterpret_cast<uint32_t *>(base + o_mnext[k]);
        }
    }
    // arrivals start at zero and every rebuild zeroes them again
    FH_HIP(hipMemsetAsync(b->res_arena, 0, a.used, b->stream));
    FH_HIP(hipStreamSynchronize(b->stream));
    return FH_OK;
}

// Every list of a lazy batch empty (at the start of each period after a
// settle or an entry; the steps of a period clear the lists they reuse).
static fh_status clear_mover_lists(fh_batch *b) {
    if (b->lazy && b->ml.head4)
        FH_HIP(launch_fill(reinterpret_cast<uint32_t *>(b->ml.head4), 5u * LAZY_PERIOD * b->ml.heads, MOVER_NIL,
                           b->stream));
    return FH_OK;
}

// Bird-order state of every live flock -> resident records: the histogram,
// the scan with the counts kept, every bird to its slot, the tiles' totals.
static fh_status batch_enter_resident(fh_batch *b) {
    fh_status st = batch_alloc_resident(b);
    if (st) return st;
    if (!b->hashed && (st = batch_hash(b))) return st;
    Pools p = b->P;
    p.offsets = b->off[0];
    FH_HIP(launch_scan(b->geo, p, nullptr, 0u, 0.0f, b->stream, true));
    FH_HIP(launch_resident_enter(b->geo, p, b->R[0], b->stream));
    FH_HIP(launch_resident_tile_sums(b->geo, p, b->W, b->stream));
    FH_HIP(launch_fill(b->res_due, 4, 0u, b->stream));  // parity 0's counters may hold an earlier step
    b->parity = 0;
    b->resident = true;
    b->hashed = false;  // the counts now follow the layout; the ranks are stale
    b->counts_dirty = false;
    b->phase = b->fq = b->cpar = 0;
    if (b->lazy) {
        if ((st = clear_mover_lists(b))) return st;
        if ((st = push_tables(b, 0, 0))) return st;  // the flocks' slot ranges (the lists need them)
    }
    return FH_OK;
}

// Flocks join and leave a resident batch in place. Each flock's records fill
// as many consecutive slots as it has birds, flocks in the order of their
// bucket ranges; a splice moves the records of the flocks after the one that
// joins or leaves, keeping every record's order, and the offsets follow from
// a scan of the counts (kept). This replaces a flush of every flock to bird
// order and a full re-entry (hash, scan and placement of every bird).

// The slot where each live flock's records begin.
static std::vector<uint32_t> resident_slot_starts(const fh_batch *b) {
    std::vector<uint32_t> order;
    for (uint32_t i = 0; i < b->segs.size(); ++i)
        if (in_layout(b, i)) order.push_back(i);
    std::sort(order.begin(), order.end(),
              [&](uint32_t x, uint32_t y) { return b->segs[x].bucket_base < b->segs[y].bucket_base; });
    std::vector<uint32_t> start(b->segs.size(), 0u);
    uint32_t at = 0;
    for (uint32_t i : order) {
        start[i] = at;
        at += b->segs[i].n;
    }
    return start;
}

// Each segment's first slot in the resident layout (slot_base; pushed with
// the segment table).
static void set_slot_bases(fh_batch *b) {
    const std::vector<uint32_t> start = resident_slot_starts(b);
    for (size_t f = 0; f < b->segs.size(); ++f) b->segs[f].slot_base = start[f];
}

// Splice shifts: the live flocks whose buckets lie after flock f's move by
// `by` slots, f itself by `own`, the rest stay.
static fh_status push_shifts(fh_batch *b, uint32_t f, int32_t by, int32_t own) {
    std::vector<int32_t> sh(b->segs.size(), 0);
    for (uint32_t i = 0; i < b->segs.size(); ++i)
        if (i != f && b->segs[i].live && b->segs[i].bucket_base > b->segs[f].bucket_base) sh[i] = by;
    sh[f] = own;
    FH_HIP(hipMemcpyAsync(b->d_shift, sh.data(), sh.size() * sizeof(int32_t), hipMemcpyHostToDevice, b->stream));
    FH_HIP(hipStreamSynchronize(b->stream));  // sh is freed on return
    return FH_OK;
}

// Buckets of a flock's range (whole scan tiles).
static uint32_t seg_bucket_span(const Seg &s) { return (s.mask + SCAN_TILE) / SCAN_TILE * SCAN_TILE; }

// The offsets of the current parity from the counts, and fresh chunk
// counters for the next step's decide.
static fh_status resident_rescan(fh_batch *b) {
    Pools p = b->P;
    p.offsets = b->off[b->parity];
    FH_HIP(launch_scan(b->geo, p, nullptr, 0u, 0.0f, b->stream, true));
    FH_HIP(launch_fill(b->res_due, 4, 0u, b->stream));
    return FH_OK;
}

// Flock f leaves (before its tables entries are cleared): the flocks after it
// move down by its birds; its counts and tile totals become zero.
static fh_status resident_leave_one(fh_batch *b, uint32_t f) {
    const Seg s = b->segs[f];
    const uint32_t q = b->parity;
    fh_status st = push_shifts(b, f, -(int32_t)s.n, SPLICE_DROP);
    if (st) return st;
    FH_HIP(launch_resident_splice(b->geo, b->P, b->R[q], b->R[1u - q], b->d_shift, b->geo.n_slots,
                                  b->geo.n_slots - s.n, b->stream));
    const uint32_t span = seg_bucket_span(s);
    FH_HIP(launch_fill(b->P.counts + s.bucket_base, span, 0u, b->stream));
    FH_HIP(launch_fill(b->W.tsum + s.bucket_base / RES_TILE, span / RES_TILE, 0u, b->stream));
    b->parity = 1u - q;
    return FH_OK;
}

// Flock f joins (its birds uploaded, its table entries pushed, the geometry
// refreshed): its histogram, the flocks after it move up by its birds, the
// offsets and tile totals, and its birds to their slots.
static fh_status resident_join_one(fh_batch *b, uint32_t f) {
    const Seg s = b->segs[f];
    const uint32_t q = b->parity;
    const uint32_t span = seg_bucket_span(s);
    FH_HIP(launch_fill(b->P.counts + s.bucket_base, span, 0u, b->stream));
    Geometry g = b->geo;
    g.tile_begin = s.bird_base / TILE;
    g.tile_count = (s.n + TILE - 1) / TILE;
    FH_HIP(launch_hash_count(g, b->P, (b->prm.flags & FH_FLAG_WAVE_AGG) != 0, b->stream));
    fh_status st = push_shifts(b, f, (int32_t)s.n, 0);
    if (st) return st;
    FH_HIP(launch_resident_splice(b->geo, b->P, b->R[q], b->R[1u - q], b->d_shift, b->geo.n_slots - s.n,
                                  b->geo.n_slots, b->stream));
    b->parity = 1u - q;
    if ((st = resident_rescan(b))) return st;
    Pools p = b->P;
    p.offsets = b->off[b->parity];
    FH_HIP(launch_resident_tile_sums(b->geo, p, b->W, b->stream));
    FH_HIP(launch_resident_enter_range(b->geo, p, b->R[b->parity], s.bird_base, s.n, b->stream));
    return FH_OK;
}

// One resident step over every live flock (the events as for one flock).
static fh_status batch_resident_launches(fh_batch *b, hipEvent_t *ev) {
    const uint32_t q = b->parity;
    const uint32_t nseg = (uint32_t)b->segs.size();
    Pools p = b->P;
    p.offsets = b->off[q];
    p.due_count = b->res_due + q;
    p.work = b->res_due + 2 + q;
    const bool fused = (b->prm.flags & FH_FLAG_FLIGHT) != 0;
    if (ev) FH_HIP(hipEventRecord(ev[0], b->stream));
    FH_HIP(launch_decide_resident(b->geo, p, b->R[q], b->prm.decide_block, b->prm.birds_per_thread,
                                  b->prm.decide_group, walk_flags(b->prm.flags), b->res_due + (1u - q),
                                  b->res_due + 2 + (1u - q), b->W, b->F, b->stream,
                                  ev ? DecideTiming{ev[6], ev[7]} : DecideTiming{}));
    if (ev) FH_HIP(hipEventRecord(ev[1], b->stream));
    if (!fused) FH_HIP(launch_resident_integrate(b->geo, p, b->R[q], b->W, b->stream));
    if (ev) FH_HIP(hipEventRecord(ev[2], b->stream));
    Records src = b->R[q];
    if (fused) {
        src.p4 = b->F.p4;
        src.v4 = b->F.v4;
        src.spin = b->F.spin;
    }
    FH_HIP(launch_resident_rebuild(b->geo, p, src, b->R[1u - q], b->W, b->off[q], b->off[1u - q], b->d_segs, nseg,
                                   b->stream));
    if (ev)
        for (int k = 3; k < 6; ++k) FH_HIP(hipEventRecord(ev[k], b->stream));
    return FH_OK;
}

// Drops the resident step graphs when the geometry changed since they were
// captured.
static void drop_stale_resident_graphs(fh_batch *b) {
    if (b->rgraph_valid) return;
    for (auto &gx : b->rgraph)
        if (gx) {
            (void)hipGraphExecDestroy(gx);
            gx = nullptr;
        }
    for (auto &row : b->lgraph)
        for (auto &gx : row)
            if (gx) {
                (void)hipGraphExecDestroy(gx);
                gx = nullptr;
            }
    b->rgraph_valid = true;
}

static fh_status batch_capture_resident(fh_batch *b) {
    drop_stale_resident_graphs(b);
    const uint32_t q = b->parity;
    if (b->rgraph[q]) return FH_OK;
    hipGraph_t g = nullptr;
    FH_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    fh_status st = batch_resident_launches(b, nullptr);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&b->rgraph[q], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->rgraph[q] = nullptr;
        return hip_fail(e, "hipGraphInstantiate");
    }
    return FH_OK;
}


// ---- lazy rebuilds (FH_FLAG_LAZY) ------------------------------------------

// The records of the current lazy step: R[parity]'s, with the positions,
// velocities and roll rates F's when `f`.
static Records batch_lazy_view(const fh_batch *b, uint32_t f) {
    Records r = b->R[b->parity];
    if (f) {
        r.p4 = b->F.p4;
        r.v4 = b->F.v4;
        r.spin = b->F.spin;
    }
    return r;
}

// One lazy step of every live flock: decide (its walks visit the mover lists
// too) with the fused flight from the current records into the other set.
// Each flock's clock and budget follow from the segment table and the phase.
// Between rebuilds the flight lists its movers for the next step; every
// LAZY_PERIOD-th step it counts them against the layout instead, and the
// rebuild follows, committing the period's steps. Events as
// batch_resident_launches.
static fh_status batch_lazy_launches(fh_batch *b, hipEvent_t *ev) {
    const uint32_t q = b->parity, c = b->cpar;
    const bool rebuild = b->phase + 1u == LAZY_PERIOD;
    Pools p = b->P;
    p.offsets = b->off[q];
    p.due_count = b->res_due + c;
    p.work = b->res_due + 2 + c;
    const Records cur = batch_lazy_view(b, b->fq), flown = batch_lazy_view(b, b->fq ^ 1u);
    const int wf = walk_flags(b->prm.flags) | WALK_MOVERS | (rebuild ? 0 : WALK_LAZY);
    ResWork w = with_mover_lists(b->W, b->ml, b->phase);
    w.mclock_now = nullptr;  // clocks from the segment table and the phase
    w.mclock_next = nullptr;
    w.mtick = nullptr;
    w.mphase = b->phase;
    if (ev) FH_HIP(hipEventRecord(ev[0], b->stream));
    FH_HIP(launch_decide_resident(b->geo, p, cur, b->prm.decide_block, b->prm.birds_per_thread, b->prm.decide_group,
                                  wf, b->res_due + (1u - c), b->res_due + 2 + (1u - c), w, flown, b->stream,
                                  ev ? DecideTiming{ev[6], ev[7]} : DecideTiming{}));
    if (ev) {
        FH_HIP(hipEventRecord(ev[1], b->stream));
        FH_HIP(hipEventRecord(ev[2], b->stream));
    }
    if (rebuild) {
        FH_HIP(launch_resident_rebuild(b->geo, p, flown, b->R[1u - q], b->W, b->off[q], b->off[1u - q], b->d_segs,
                                       (uint32_t)b->segs.size(), b->stream, LAZY_PERIOD));
        b->parity = 1u - q;
        b->fq = 0;
        b->phase = 0;
    } else {
        b->fq ^= 1u;
        b->phase += 1u;
    }
    b->cpar = 1u - c;
    if (ev)
        for (int k = 3; k < 6; ++k) FH_HIP(hipEventRecord(ev[k], b->stream));
    return FH_OK;
}

// A lazy batch between rebuilds -> the layout of the current positions: the
// movers counted against their slots' buckets without a flight, then the
// rebuild, which commits the steps of the period so far; every list empty.
// Nothing to do at a rebuild boundary.
static fh_status batch_settle(fh_batch *b) {
    if (!b->lazy || !b->resident || b->phase == 0) return FH_OK;
    const uint32_t q = b->parity;
    Pools p = b->P;
    p.offsets = b->off[q];
    const Records cur = batch_lazy_view(b, b->fq);
    FH_HIP(launch_resident_account(b->geo, p, cur, b->W, b->stream));
    FH_HIP(launch_resident_rebuild(b->geo, p, cur, b->R[1u - q], b->W, b->off[q], b->off[1u - q], b->d_segs,
                                   (uint32_t)b->segs.size(), b->stream, b->phase));
    b->parity = 1u - q;
    b->fq = 0;
    b->phase = 0;
    return clear_mover_lists(b);
}

// LAZY_PERIOD lazy steps from phase 0 as one graph, captured once per layout
// parity and counter parity (the host state is restored after the capture;
// the caller replays it and applies its effect).
static fh_status batch_capture_lazy(fh_batch *b) {
    drop_stale_resident_graphs(b);
    const uint32_t q = b->parity, c = b->cpar;
    hipGraphExec_t *out = &b->lgraph[q][c];
    if (*out) return FH_OK;
    hipGraph_t g = nullptr;
    FH_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    fh_status st = FH_OK;
    for (uint32_t k = 0; k < LAZY_PERIOD && !st; ++k) st = batch_lazy_launches(b, nullptr);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    b->parity = q;
    b->cpar = c;
    b->phase = 0;
    b->fq = 0;
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        *out = nullptr;
        return hip_fail(e, "hipGraphInstantiate");
    }
    return FH_OK;
}

namespace fh {

size_t resident_bytes_per_bird(bool flight) {
    // two records (p4, v4, s4: 48; spin, id, key: 12), arank, mover bits
    return 2 * 60 + 4 + 1 + (flight ? 36 : 0);
}
size_t resident_bytes_per_bucket() { return 4 /*arr*/ + 4 /*abase*/ + 4 /*offsets*/ + 1 /*tile totals*/; }

// One step of every live flock (launched, not waited for), with per-pass
// events when `ev` is given.
// The host mirror of one step: each stepping flock's clock advances by DT
// (one f32 addition, as on the device) and its budget by one step.
static void mirror_step(fh_batch *b) {
    for (size_t f = 0; f < b->segs.size(); ++f)
        if (in_layout(b, (uint32_t)f) && (!b->resident || b->segs[f].left)) {
            b->times[f] += DT;
            if (b->resident && b->segs[f].left != SEG_UNLIMITED) b->segs[f].left -= 1u;
        }
}

static fh_status one_step(fh_batch *b, bool graph, hipEvent_t *ev) {
    mirror_step(b);
    fh_status st = FH_OK;
    if (b->resident && b->lazy) {
        st = batch_lazy_launches(b, ev);  // whole periods replay as graphs (lazy_period)
    } else if (b->resident) {
        if (graph) {
            if ((st = batch_capture_resident(b))) return st;
            hipError_t e = hipGraphLaunch(b->rgraph[b->parity], b->stream);
            if (e != hipSuccess) st = hip_fail(e, "hipGraphLaunch");
        } else {
            st = batch_resident_launches(b, ev);
        }
        b->parity ^= 1u;
    } else if (graph) {
        if (!b->hashed && (st = batch_hash(b))) return st;
        if ((st = batch_capture(b))) return st;
        hipError_t e = hipGraphLaunch(b->graph, b->stream);
        if (e != hipSuccess) st = hip_fail(e, "hipGraphLaunch");
        b->hashed = true;
        b->counts_dirty = true;
    } else {
        st = batch_step_launches(b, ev);
    }
    return st;
}

// LAZY_PERIOD steps of a lazy batch at phase 0 as one graph replay.
static fh_status lazy_period(fh_batch *b) {
    fh_status st = batch_capture_lazy(b);
    if (st) return st;
    hipError_t e = hipGraphLaunch(b->lgraph[b->parity][b->cpar], b->stream);
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    for (uint32_t k = 0; k < LAZY_PERIOD; ++k) mirror_step(b);
    b->parity ^= 1u;  // one rebuild per period; the counter parity comes back
    return FH_OK;
}

// Whether the next LAZY_PERIOD of `steps` steps (from step s, the first
// `warmup` untimed) replay as one graph: a lazy batch at phase 0, graphs on,
// no per-pass timing, and the period on one side of the warm-up boundary.
static bool period_fits(const fh_batch *b, bool graph, bool per_pass, uint32_t s, uint32_t steps, uint32_t warmup) {
    return graph && !per_pass && b->resident && b->lazy && b->phase == 0 && s + LAZY_PERIOD <= steps &&
           (s >= warmup || s + LAZY_PERIOD <= warmup);
}

fh_status batch_step(fh_batch *b, uint32_t steps, uint32_t warmup, fh_pass_ms *out) {
    if (warmup > steps) warmup = steps;
    const bool graph = (b->prm.flags & FH_FLAG_GRAPH) != 0;
    const bool res = (b->prm.flags & FH_FLAG_RESIDENT) != 0;
    const uint32_t timed = steps - warmup;
    if (res && steps && !b->resident && b->geo.n_slots) {
        fh_status st0 = batch_enter_resident(b);
        if (st0) return st0;
    }
    const bool per_pass = out && !graph;
    std::vector<hipEvent_t> ev;
    if (out) {
        ev.resize(per_pass ? (size_t)timed * 8 : 2);
        for (auto &x : ev) FH_HIP(hipEventCreate(&x));
    }
    auto destroy = [&] {
        for (auto &x : ev) (void)hipEventDestroy(x);
    };
    fh_status st = FH_OK;
    std::chrono::steady_clock::time_point t0{}, t1{};
    for (uint32_t s = 0; s < steps && st == FH_OK;) {
        if (s == warmup) {
            (void)hipStreamSynchronize(b->stream);
            t0 = std::chrono::steady_clock::now();
            if (out && graph) (void)hipEventRecord(ev[0], b->stream);
        }
        if (period_fits(b, graph, per_pass, s, steps, warmup)) {
            st = lazy_period(b);
            s += LAZY_PERIOD;
        } else {
            st = one_step(b, graph, (per_pass && s >= warmup) ? &ev[(size_t)(s - warmup) * 8] : nullptr);
            s += 1;
        }
    }
    if (st) {
        (void)hipStreamSynchronize(b->stream);
        destroy();
        return st;
    }
    if (out && graph && timed) FH_HIP(hipEventRecord(ev[1], b->stream));
    FH_HIP(hipStreamSynchronize(b->stream));
    t1 = std::chrono::steady_clock::now();
    if (out) {
        std::memset(out, 0, sizeof *out);
        out->timed_steps = timed;
        if (timed) {
            out->wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
            if (per_pass) {
                // events per step: six between the passes, then decide's kernel start and end
                double acc[5] = {0, 0, 0, 0, 0}, dec = 0.0, total = 0.0;
                for (uint32_t s = 0; s < timed; ++s) {
                    hipEvent_t *e = &ev[(size_t)s * 8];
                    float ms = 0.0f;
                    for (int k = 0; k < 5; ++k) {
                        FH_HIP(hipEventElapsedTime(&ms, e[k], e[k + 1]));
                        acc[k] += ms;
                    }
                    FH_HIP(hipEventElapsedTime(&ms, e[6], e[7]));
                    dec += ms;
                    FH_HIP(hipEventElapsedTime(&ms, e[0], e[5]));
                    total += ms;
                }
                out->decide = dec / timed;
                out->step_total = total / timed;
                if (b->resident) {  // decide, integrate, then the rebuild (scan and move, reported as scatter)
                    out->grid_scatter = acc[2] / timed;
                    out->integrate = acc[1] / timed;
                } else {
                    out->grid_histogram = acc[0] / timed;
                    out->grid_scan = acc[1] / timed;
                    out->grid_scatter = acc[2] / timed;
                    out->integrate = acc[4] / timed;
                }
            } else {
                float ms = 0.0f;
                FH_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
                out->step_total = ms / timed;
            }
        }
        uint32_t due = 0;
        const uint32_t *last = !b->resident ? b->P.due_count
                               : b->lazy    ? b->res_due + (b->cpar ^ 1u)
                                            : b->res_due + (b->parity ^ 1u);
        FH_HIP(hipMemcpy(&due, last, 4, hipMemcpyDeviceToHost));
        out->due_last = due;
    }
    destroy();
    uint32_t err = 0;
    FH_HIP(hipMemcpy(&err, b->P.err, 4, hipMemcpyDeviceToHost));
    if (err & ERR_SLOT) return set_error(FH_ERR_HIP, "kernel reported a slot outside the sorted range (0x%x)", err);
    return FH_OK;
}

size_t staged_size(uint32_t n) { return staged_bytes(n); }

void stage_words(const float *words, uint32_t n, void *dst) {
    const Staged h = staged(dst, n);
    for (size_t i = 0; i < n; ++i) {
        const float *w = words + i * FH_WORDS;
        h.pos[i] = make_float4(w[PX], w[PY], w[PZ], w[NEXT]);
        h.vel[i] = make_float4(w[VX], w[VY], w[VZ], w[ROLL]);
        h.steer[i] = make_float4(w[SX], w[SY], w[SZ], w[RADIUS]);
        h.spin[i] = w[SPIN];
        std::memcpy(&h.nb[i], &w[NB], 4);
        h.cen[i] = w[CEN];
    }
}

void unstage_words(const void *src, uint32_t n, float *words) {
    const Staged h = staged(const_cast<void *>(src), n);
    for (size_t i = 0; i < n; ++i) {
        float *w = words + i * FH_WORDS;
        w[PX] = h.pos[i].x; w[PY] = h.pos[i].y; w[PZ] = h.pos[i].z;
        w[VX] = h.vel[i].x; w[VY] = h.vel[i].y; w[VZ] = h.vel[i].z;
        w[ROLL] = h.vel[i].w;
        w[SPIN] = h.spin[i];
        w[RADIUS] = h.steer[i].w;
        w[SX] = h.steer[i].x; w[SY] = h.steer[i].y; w[SZ] = h.steer[i].z;
        w[NEXT] = h.pos[i].w;
        std::memcpy(&w[NB], &h.nb[i], 4);
        w[CEN] = h.cen[i];
    }
}

// The six arrays of the staged records of segment s and their pool addresses.
struct Part {
    void *pool;
    void *host;
    size_t bytes;
};
static void staged_parts(fh_batch *b, const Seg &s, void *host, Part out[6]) {
    const Staged h = staged(host, s.n);
    const size_t o = s.bird_base, n = s.n;
    out[0] = {b->P.pos4 + o, h.pos, n * sizeof(float4)};
    out[1] = {b->P.vel4 + o, h.vel, n * sizeof(float4)};
    out[2] = {b->P.steer4 + o, h.steer, n * sizeof(float4)};
    out[3] = {b->P.spin + o, h.spin, n * 4};
    out[4] = {b->P.cen + o, h.cen, n * 4};
    out[5] = {b->P.nb + o, h.nb, n * 4};
}

fh_status batch_join_async(fh_batch *b, const void *src, uint32_t n, float time, uint32_t steps, uint32_t *fid) {
    if (!(b->prm.flags & FH_FLAG_RESIDENT)) return set_error(FH_ERR_INVALID, "batch_join_async: needs FH_FLAG_RESIDENT");
    if (!src || !n || !fid) return set_error(FH_ERR_INVALID, "batch_join_async: need records, n > 0 and fid");
    const uint32_t buckets = flock_buckets(n, b->prm);
    const uint32_t ntile = (n + TILE - 1) / TILE;
    const uint32_t nunit = (buckets + SCAN_TILE - 1) / SCAN_TILE;
    uint32_t f = NO_SEG;
    for (uint32_t i = 0; i < b->segs.size(); ++i)
        if (!b->segs[i].live && !b->pending[i]) {
            f = i;
            break;
        }
    if (f == NO_SEG) {
        if (b->segs.size() >= b->flock_cap) return set_error(FH_ERR_BUDGET, "batch_join_async: flock slots full");
        f = (uint32_t)b->segs.size();
        b->segs.push_back(Seg{});
        b->times.push_back(0.0f);
        b->pending.push_back(0);
    }
    const uint32_t t0 = b->tiles.alloc(ntile);
    if (t0 == RangeAlloc::NONE) return set_error(FH_ERR_BUDGET, "batch_join_async: no room for %u birds", n);
    const uint32_t u0 = b->bucket_units.alloc(nunit);
    if (u0 == RangeAlloc::NONE) {
        b->tiles.release(t0);
        return set_error(FH_ERR_BUDGET, "batch_join_async: no room for %u buckets", buckets);
    }
    Seg sg{};
    sg.bird_base = t0 * TILE;
    sg.n = n;
    sg.bucket_base = u0 * SCAN_TILE;
    sg.mask = buckets - 1;
    sg.cell = b->prm.cell;
    sg.inv_cell = 1.0f / b->prm.cell;
    sg.time = time;
    sg.live = 1;
    sg.mode = (b->prm.flags & FH_FLAG_SPATIAL) ? SEG_SPATIAL : 0u;
    sg.left = steps;
    Part parts[6];
    staged_parts(b, sg, const_cast<void *>(src), parts);
    hipError_t e = hipSuccess;
    for (const Part &pt : parts)
        if (e == hipSuccess) e = hipMemcpyAsync(pt.pool, pt.host, pt.bytes, hipMemcpyHostToDevice, b->copy);
    if (e == hipSuccess) e = hipEventRecord(b->up_done, b->copy);
    if (e != hipSuccess) {
        // the ranges stay reserved: an upload may be in flight into them
        return hip_fail(e, "batch_join_async: upload");
    }
    b->uploads = true;
    b->segs[f] = sg;
    b->times[f] = time;
    b->pending[f] = PENDING_JOIN;
    for (uint32_t t = 0; t < ntile; ++t) b->tile_seg[t0 + t] = f;
    b->joins.push_back(f);
    *fid = f;
    return FH_OK;
}

fh_status batch_leave_async(fh_batch *b, uint32_t fid, void *dst, float *time, hipEvent_t done) {
    if (fid >= b->segs.size() || !b->segs[fid].live || b->pending[fid])
        return set_error(FH_ERR_INVALID, "batch_leave_async: no flock %u in the layout", fid);
    if (!b->resident) return set_error(FH_ERR_STATE, "batch_leave_async: the batch has not stepped");
    if (!dst) return set_error(FH_ERR_INVALID, "batch_leave_async: NULL destination");
    fh_status st = batch_settle(b);
    if (st) return st;
    const Seg sg = b->segs[fid];
    FH_HIP(launch_resident_flush_range(b->geo, b->P, b->R[b->parity], resident_slot_starts(b)[fid], sg.n,
                                       b->stream));
    FH_HIP(hipEventRecord(b->flushed, b->stream));
    FH_HIP(hipStreamWaitEvent(b->copy, b->flushed, 0));
    Part parts[6];
    staged_parts(b, sg, dst, parts);
    for (const Part &pt : parts) FH_HIP(hipMemcpyAsync(pt.host, pt.pool, pt.bytes, hipMemcpyDeviceToHost, b->copy));
    if (done) FH_HIP(hipEventRecord(done, b->copy));
    if (time) *time = b->times[fid];
    b->pending[fid] = PENDING_LEAVE;
    b->leaves.push_back(fid);
    return FH_OK;
}

// Pinned staging of the tables one apply uploads: segment tables A and B,
// tile -> segment mirrors A and B, and the splice shifts.
static fh_status table_staging(fh_batch *b, Seg **sa, Seg **sb, uint32_t **ta, uint32_t **tb, int32_t **sh) {
    const size_t ns = b->flock_cap, nt = b->tile_seg.size();
    const size_t segs_b = (ns * sizeof(Seg) + 255) & ~size_t(255), tiles_b = (nt * 4 + 255) & ~size_t(255);
    const size_t need = 2 * segs_b + 2 * tiles_b + ns * 4;
    if (b->tbl_bytes < need) {
        if (b->tbl) (void)hipHostFree(b->tbl);
        b->tbl = nullptr;
        b->tbl_bytes = 0;
        FH_HIP(hipHostMalloc(&b->tbl, need, hipHostMallocDefault));
        b->tbl_bytes = need;
    }
    char *p = static_cast<char *>(b->tbl);
    *sa = reinterpret_cast<Seg *>(p);
    *sb = reinterpret_cast<Seg *>(p + segs_b);
    *ta = reinterpret_cast<uint32_t *>(p + 2 * segs_b);
    *tb = reinterpret_cast<uint32_t *>(p + 2 * segs_b + tiles_b);
    *sh = reinterpret_cast<int32_t *>(p + 2 * segs_b + 2 * tiles_b);
    return FH_OK;
}

// Uploads the segment table and the tile -> segment entries of the given
// segments' ranges from pinned staging `seg_st` / `tile_st` (no host wait).
static fh_status push_tables_async(fh_batch *b, Seg *seg_st, uint32_t *tile_st, const std::vector<uint32_t> &fs) {
    set_slot_bases(b);
    for (size_t f = 0; f < b->segs.size(); ++f) {
        seg_st[f] = b->segs[f];
        seg_st[f].time = b->times[f];
    }
    if (!b->segs.empty())
        FH_HIP(hipMemcpyAsync(b->d_segs, seg_st, b->segs.size() * sizeof(Seg), hipMemcpyHostToDevice, b->stream));
    for (uint32_t f : fs) {
        const uint32_t t0 = b->segs[f].bird_base / TILE, nt = (b->segs[f].n + TILE - 1) / TILE;
        std::memcpy(tile_st + t0, b->tile_seg.data() + t0, (size_t)nt * 4);
        FH_HIP(hipMemcpyAsync(b->d_tile_seg + t0, tile_st + t0, (size_t)nt * 4, hipMemcpyHostToDevice, b->stream));
    }
    return FH_OK;
}

fh_status batch_apply(fh_batch *b) {
    if (b->joins.empty() && b->leaves.empty()) return FH_OK;
    fh_status st0 = batch_settle(b);
    if (st0) return st0;
    if (b->uploads) {
        FH_HIP(hipStreamWaitEvent(b->stream, b->up_done, 0));
        b->uploads = false;
    }
    Seg *sa, *sb;
    uint32_t *ta, *tb;
    int32_t *sh;
    fh_status st = table_staging(b, &sa, &sb, &ta, &tb, &sh);
    if (st) return st;
    if (!b->resident) {
        // Not stepped yet: the flocks join the bird-order state, and the next
        // step hashes every flock and enters the resident layout.
        if (!b->leaves.empty()) return set_error(FH_ERR_STATE, "batch_apply: leave before the first step");
        for (uint32_t f : b->joins) b->pending[f] = 0;
        if ((st = push_tables_async(b, sa, ta, b->joins))) return st;
        refresh_geometry(b);
        b->hashed = false;
        b->joins.clear();
        return FH_OK;
    }
    const uint32_t q = b->parity;
    // Slot ranges before (every flock in the layout, the leaving ones too) and
    // after (without the leaving flocks, with the joining ones), in bucket order.
    std::vector<uint32_t> order;
    for (uint32_t i = 0; i < b->segs.size(); ++i)
        if (b->segs[i].live) order.push_back(i);
    std::sort(order.begin(), order.end(),
              [&](uint32_t x, uint32_t y) { return b->segs[x].bucket_base < b->segs[y].bucket_base; });
    std::vector<uint32_t> old_start(b->segs.size(), 0u), new_start(b->segs.size(), 0u);
    uint32_t n_cur = 0, n_new = 0;
    for (uint32_t i : order) {
        if (b->pending[i] != PENDING_JOIN) {
            old_start[i] = n_cur;
            n_cur += b->segs[i].n;
        }
        if (b->pending[i] != PENDING_LEAVE) {
            new_start[i] = n_new;
            n_new += b->segs[i].n;
        }
    }
    for (uint32_t i = 0; i < b->segs.size(); ++i)
        sh[i] = b->pending[i] == PENDING_LEAVE ? SPLICE_DROP
                : (b->segs[i].live && !b->pending[i]) ? (int32_t)new_start[i] - (int32_t)old_start[i]
                                                      : 0;
    // Tables A: the joining flocks in, the leaving ones still there (the
    // splice finds a record's flock through its bird id).
    if ((st = push_tables_async(b, sa, ta, b->joins))) return st;
    Geometry gu = b->geo;  // the ranges of the leaving and the joining flocks together
    gu.n_birds = b->tiles.high_water() * TILE;
    gu.n_buckets = b->bucket_units.high_water() * SCAN_TILE;
    gu.n_slots = n_new;
    for (uint32_t f : b->leaves) FH_HIP(launch_fill(b->P.counts + b->segs[f].bucket_base, seg_bucket_span(b->segs[f]), 0u, b->stream));
    for (uint32_t f : b->joins) {
        const Seg &sg = b->segs[f];
        FH_HIP(launch_fill(b->P.counts + sg.bucket_base, seg_bucket_span(sg), 0u, b->stream));
        Geometry g = gu;
        g.tile_begin = sg.bird_base / TILE;
        g.tile_count = (sg.n + TILE - 1) / TILE;
        FH_HIP(launch_hash_count(g, b->P, (b->prm.flags & FH_FLAG_WAVE_AGG) != 0, b->stream));
    }
    FH_HIP(hipMemcpyAsync(b->d_shift, sh, b->segs.size() * sizeof(int32_t), hipMemcpyHostToDevice, b->stream));
    FH_HIP(launch_resident_splice(gu, b->P, b->R[q], b->R[1u - q], b->d_shift, n_cur, n_new, b->stream));
    b->parity = 1u - q;
    // Tables B: the leaving flocks out; their ranges free for later joins
    // (their downloads run before any later upload on the copy stream).
    std::vector<uint32_t> gone = b->leaves;
    for (uint32_t f : gone) {
        Seg &sg = b->segs[f];
        const uint32_t t0 = sg.bird_base / TILE, nt = (sg.n + TILE - 1) / TILE;
        for (uint32_t t = 0; t < nt; ++t) b->tile_seg[t0 + t] = NO_SEG;
        b->tiles.release(t0);
        b->bucket_units.release(sg.bucket_base / SCAN_TILE);
        sg.live = 0;
        b->pending[f] = 0;
    }
    for (uint32_t f : b->joins) b->pending[f] = 0;
    if ((st = push_tables_async(b, sb, tb, gone))) return st;
    refresh_geometry(b);
    if ((st = resident_rescan(b))) return st;
    Pools p = b->P;
    p.offsets = b->off[b->parity];
    FH_HIP(launch_resident_tile_sums(b->geo, p, b->W, b->stream));
    for (uint32_t f : b->joins)
        FH_HIP(launch_resident_enter_range(b->geo, p, b->R[b->parity], b->segs[f].bird_base, b->segs[f].n, b->stream));
    b->joins.clear();
    b->leaves.clear();
    return FH_OK;
}

fh_status batch_step_async(fh_batch *b, uint32_t steps) {
    if (!b->joins.empty() || !b->leaves.empty())
        return set_error(FH_ERR_STATE, "batch_step_async: membership changes queued (batch_apply first)");
    const bool graph = (b->prm.flags & FH_FLAG_GRAPH) != 0;
    if ((b->prm.flags & FH_FLAG_RESIDENT) && steps && !b->resident && b->geo.n_slots) {
        fh_status st = batch_enter_resident(b);
        if (st) return st;
    }
    for (uint32_t k = 0; k < steps;) {
        fh_status st;
        if (period_fits(b, graph, false, k, steps, 0)) {
            st = lazy_period(b);
            k += LAZY_PERIOD;
        } else {
            st = one_step(b, graph, nullptr);
            k += 1;
        }
        if (st) return st;
    }
    return FH_OK;
}

fh_status batch_wait(fh_batch *b) {
    FH_HIP(hipStreamSynchronize(b->stream));
    uint32_t err = 0;
    FH_HIP(hipMemcpy(&err, b->P.err, 4, hipMemcpyDeviceToHost));
    if (err & ERR_SLOT) return set_error(FH_ERR_HIP, "kernel reported a slot outside the sorted range (0x%x)", err);
    return FH_OK;
}

size_t batch_bytes_per_bird() {
    // pos4, vel4, steer4 (16 B each), sbird (32 B) + spin, cen, nb, key, rank,
    // snext, sorted_tmp (4 B each)
    return 5 * 16 + 7 * 4;
}

size_t batch_bytes_per_bucket() { return 4 /*counts*/ + 4 /*offsets*/; }

}  // namespace fh

extern "C" {

fh_status fh_batch_create(uint32_t bird_capacity, uint32_t bucket_capacity, uint32_t flock_capacity,
                          const fh_params *params, fh_batch **out) {
    clear_error();
    if (!out) return set_error(FH_ERR_INVALID, "fh_batch_create: NULL out");
    *out = nullptr;
    if (flock_capacity == 0 || flock_capacity > (1u << 20) || bird_capacity > (1u << 30) ||
        bucket_capacity > (1u << 30))
        return set_error(FH_ERR_INVALID, "fh_batch_create: bad capacity");
    fh_params prm;
    fh_status st = resolve_params(params, &prm, "fh_batch_create");
    if (st) return st;
    st = require_device();
    if (st) return st;

    fh_batch *b = new fh_batch();
    b->prm = prm;
    b->lazy = (prm.flags & FH_FLAG_LAZY) != 0;  // with FH_FLAG_RESIDENT and FH_FLAG_FLIGHT (resolve_params)
    const uint32_t tile_cap = (bird_capacity + TILE - 1) / TILE;
    b->bird_cap = tile_cap * TILE;
    b->bucket_cap_units = (bucket_capacity + SCAN_TILE - 1) / SCAN_TILE;
    b->flock_cap = flock_capacity;
    b->tiles.reset(tile_cap);
    b->bucket_units.reset(b->bucket_cap_units);
    const size_t nb = (size_t)b->bucket_cap_units * SCAN_TILE;
    const size_t nn = b->bird_cap ? b->bird_cap : 1;

    Arena a;
    size_t o_pos = a.take<float4>(nn), o_vel = a.take<float4>(nn), o_steer = a.take<float4>(nn);
    size_t o_sbird = a.take<float4>(2 * nn);
    size_t o_spin = a.take<float>(nn), o_cen = a.take<float>(nn), o_nb = a.take<uint32_t>(nn);
    size_t o_key = a.take<uint32_t>(nn), o_rank = a.take<uint32_t>(nn), o_snext = a.take<float>(nn);
    size_t o_tmp = a.take<uint32_t>(nn);
    size_t o_counts = a.take<uint32_t>(nb), o_offsets = a.take<uint32_t>(nb + 1);
    size_t o_tsum = a.take<uint32_t>(b->bucket_cap_units + 1), o_tpre = a.take<uint32_t>(b->bucket_cap_units + 1);
    size_t o_err = a.take<uint32_t>(4), o_due = a.take<uint32_t>(4);
    size_t o_bigc = a.take<uint32_t>(4), o_big = a.take<uint32_t>(big_list_cap((uint32_t)nn));
    size_t o_occ = a.take<uint32_t>(nb / 32);
    size_t o_segs = a.take<Seg>(flock_capacity), o_tseg = a.take<uint32_t>(tile_cap ? tile_cap : 1);
    hipError_t e = hipMalloc(&b->arena, a.used);
    if (e != hipSuccess) {
        delete b;
        return hip_fail(e, "hipMalloc(batch arena)");
    }
    b->arena_bytes = a.used;
    char *base = static_cast<char *>(b->arena);
    b->P.pos4 = reinterpret_cast<float4 *>(base + o_pos);
    b->P.vel4 = reinterpret_cast<float4 *>(base + o_vel);
    b->P.steer4 = reinterpret_cast<float4 *>(base + o_steer);
    b->P.sbird = reinterpret_cast<float4 *>(base + o_sbird);
    b->P.spin = reinterpret_cast<float *>(base + o_spin);
    b->P.cen = reinterpret_cast<float *>(base + o_cen);
    b->P.nb = reinterpret_cast<uint32_t *>(base + o_nb);
    b->P.key = reinterpret_cast<uint32_t *>(base + o_key);
    b->P.rank = reinterpret_cast<uint32_t *>(base + o_rank);
    b->P.snext = reinterpret_cast<float *>(base + o_snext);
    b->P.sorted_tmp = reinterpret_cast<uint32_t *>(base + o_tmp);
    b->P.counts = reinterpret_cast<uint32_t *>(base + o_counts);
    b->P.offsets = reinterpret_cast<uint32_t *>(base + o_offsets);
    b->P.tile_sums = reinterpret_cast<uint32_t *>(base + o_tsum);
    b->P.tile_prefix = reinterpret_cast<uint32_t *>(base + o_tpre);
    b->P.err = reinterpret_cast<uint32_t *>(base + o_err);
    b->P.due_count = reinterpret_cast<uint32_t *>(base + o_due);
    b->P.work = b->P.due_count + 1;
    b->P.big_count = reinterpret_cast<uint32_t *>(base + o_bigc);
    b->P.big = reinterpret_cast<uint32_t *>(base + o_big);
    b->P.big_cap = big_list_cap((uint32_t)nn);
    b->P.occ = reinterpret_cast<uint32_t *>(base + o_occ);
    b->d_segs = reinterpret_cast<Seg *>(base + o_segs);
    b->d_tile_seg = reinterpret_cast<uint32_t *>(base + o_tseg);
    b->tile_seg.assign(tile_cap, NO_SEG);

    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&b->copy, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->up_done, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->flushed, hipEventDisableTiming)) != hipSuccess) {
        if (b->flushed) (void)hipEventDestroy(b->flushed);
        if (b->up_done) (void)hipEventDestroy(b->up_done);
        if (b->copy) (void)hipStreamDestroy(b->copy);
        if (b->stream) (void)hipStreamDestroy(b->stream);
        (void)hipFree(b->arena);
        delete b;
        return hip_fail(e, "fh_batch_create: streams and events");
    }
    if ((e = hipMemsetAsync(b->arena, 0, a.used, b->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(b->stream)) != hipSuccess) {
        fh_batch_destroy(b);
        return hip_fail(e, "hipMemsetAsync(batch arena)");
    }
    if ((st = push_tables(b, 0, tile_cap))) {
        const std::string msg = fh_last_error();
        fh_batch_destroy(b);
        return set_error(st, "%s", msg.c_str());
    }
    refresh_geometry(b);
    *out = b;
    return FH_OK;
}

fh_status fh_batch_destroy(fh_batch *b) {
    clear_error();
    if (!b) return FH_OK;
    (void)hipStreamSynchronize(b->stream);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    for (auto &gx : b->rgraph)
        if (gx) (void)hipGraphExecDestroy(gx);
    for (auto &row : b->lgraph)
        for (auto &gx : row)
            if (gx) (void)hipGraphExecDestroy(gx);
    if (b->copy) (void)hipStreamSynchronize(b->copy);
    if (b->res_arena) (void)hipFree(b->res_arena);
    if (b->pin) (void)hipHostFree(b->pin);
    if (b->tbl) (void)hipHostFree(b->tbl);
    (void)hipFree(b->arena);
    if (b->up_done) (void)hipEventDestroy(b->up_done);
    if (b->flushed) (void)hipEventDestroy(b->flushed);
    if (b->copy) (void)hipStreamDestroy(b->copy);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return FH_OK;
}

fh_status fh_batch_add(fh_batch *b, const float *words, uint32_t n, float time, uint32_t *flock_id) {
    clear_error();
    fh_status st = bcheck(b, "fh_batch_add");
    if (st) return st;
    if (!flock_id || (n && !words) || n == 0)
        return set_error(FH_ERR_INVALID, "fh_batch_add: need n > 0, words and flock_id");
    if ((st = batch_settle(b))) return st;  // a lazy batch: the layout and clocks of the current step
    const uint32_t buckets = flock_buckets(n, b->prm);
    const uint32_t ntile = (n + TILE - 1) / TILE;
    const uint32_t nunit = (buckets + SCAN_TILE - 1) / SCAN_TILE;
    // find a free segment slot
    uint32_t f = NO_SEG;
    for (uint32_t i = 0; i < b->segs.size(); ++i)
        if (!b->segs[i].live && !b->pending[i]) {
            f = i;
            break;
        }
    if (f == NO_SEG) {
        if (b->segs.size() >= b->flock_cap) return set_error(FH_ERR_BUDGET, "fh_batch_add: flock slots full");
        f = (uint32_t)b->segs.size();
        b->segs.push_back(Seg{});
        b->times.push_back(0.0f);
        b->pending.push_back(0);
    }
    const uint32_t t0 = b->tiles.alloc(ntile);
    if (t0 == RangeAlloc::NONE) return set_error(FH_ERR_BUDGET, "fh_batch_add: no room for %u birds", n);
    const uint32_t u0 = b->bucket_units.alloc(nunit);
    if (u0 == RangeAlloc::NONE) {
        b->tiles.release(t0);
        return set_error(FH_ERR_BUDGET, "fh_batch_add: no room for %u buckets", buckets);
    }
    Seg s{};
    s.bird_base = t0 * TILE;
    s.n = n;
    s.bucket_base = u0 * SCAN_TILE;
    s.mask = buckets - 1;
    s.cell = b->prm.cell;
    s.inv_cell = 1.0f / b->prm.cell;
    s.time = time;
    s.live = 1;
    s.mode = (b->prm.flags & FH_FLAG_SPATIAL) ? SEG_SPATIAL : 0u;
    // Until the tables are pushed, no launch sees the flock: its tiles map to
    // no segment on the device, so a failed upload leaves the batch as it was.
    auto unwind = [&](fh_status e) {
        b->tiles.release(t0);
        b->bucket_units.release(u0);
        return e;
    };
    if ((st = pinned(b, staged_bytes(n)))) return unwind(st);
    const Staged h = staged(b->pin, n);
    convert(b, n, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            const float *w = words + i * FH_WORDS;
            h.pos[i] = make_float4(w[PX], w[PY], w[PZ], w[NEXT]);
            h.vel[i] = make_float4(w[VX], w[VY], w[VZ], w[ROLL]);
            h.steer[i] = make_float4(w[SX], w[SY], w[SZ], w[RADIUS]);
            h.spin[i] = w[SPIN];
            std::memcpy(&h.nb[i], &w[NB], 4);
            h.cen[i] = w[CEN];
        }
    });
    const size_t o = s.bird_base;
    hipError_t e = hipSuccess;
    const struct {
        void *dst;
        const void *src;
        size_t bytes;
    } up[6] = {{b->P.pos4 + o, h.pos, n * sizeof(float4)}, {b->P.vel4 + o, h.vel, n * sizeof(float4)},
               {b->P.steer4 + o, h.steer, n * sizeof(float4)}, {b->P.spin + o, h.spin, n * 4u},
               {b->P.cen + o, h.cen, n * 4u}, {b->P.nb + o, h.nb, n * 4u}};
    for (const auto &u : up)
        if (e == hipSuccess) e = hipMemcpyAsync(u.dst, u.src, u.bytes, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);  // the staging is free again
    if (e != hipSuccess) return unwind(hip_fail(e, "fh_batch_add: upload"));
    b->segs[f] = s;
    b->times[f] = time;
    for (uint32_t t = 0; t < ntile; ++t) b->tile_seg[t0 + t] = f;
    if ((st = push_tables(b, t0, t0 + ntile))) {
        b->segs[f].live = 0;
        for (uint32_t t = 0; t < ntile; ++t) b->tile_seg[t0 + t] = NO_SEG;
        (void)push_tables(b, t0, t0 + ntile);
        return unwind(st);
    }
    // Count the new flock's birds into the histogram if the others are already
    // counted; otherwise the next step hashes everything.
    refresh_geometry(b);
    if (b->resident) {  // joins the resident layout in place
        if ((st = resident_join_one(b, f))) return st;
    } else if (b->hashed) {
        Geometry g = b->geo;
        g.tile_begin = t0;
        g.tile_count = ntile;
        FH_HIP(launch_hash_count(g, b->P, (b->prm.flags & FH_FLAG_WAVE_AGG) != 0, b->stream));
    }
    *flock_id = f;
    return FH_OK;
}

fh_status fh_batch_remove(fh_batch *b, uint32_t flock_id) {
    clear_error();
    fh_status st = bcheck(b, "fh_batch_remove");
    if (st) return st;
    if (flock_id >= b->segs.size() || !b->segs[flock_id].live)
        return set_error(FH_ERR_INVALID, "fh_batch_remove: no live flock %u", flock_id);
    if ((st = batch_settle(b))) return st;
    const bool splice = b->resident;  // leaves the resident layout in place
    if (splice && (st = resident_leave_one(b, flock_id))) return st;
    Seg &s = b->segs[flock_id];
    if (b->counts_dirty)  // drop this flock's birds from the pending histogram
        FH_HIP(launch_fill(b->P.counts + s.bucket_base, s.mask + 1, 0u, b->stream));
    const uint32_t t0 = s.bird_base / TILE, ntile = (s.n + TILE - 1) / TILE;
    for (uint32_t t = 0; t < ntile; ++t) b->tile_seg[t0 + t] = NO_SEG;
    b->tiles.release(t0);
    b->bucket_units.release(s.bucket_base / SCAN_TILE);
    s.live = 0;
    if ((st = push_tables(b, t0, t0 + ntile))) return st;
    refresh_geometry(b);
    if (splice && (st = resident_rescan(b))) return st;
    return FH_OK;
}

fh_status fh_batch_download(fh_batch *b, uint32_t flock_id, float *words, float *time) {
    clear_error();
    fh_status st = bcheck(b, "fh_batch_download");
    if (st) return st;
    if (flock_id >= b->segs.size() || !b->segs[flock_id].live)
        return set_error(FH_ERR_INVALID, "fh_batch_download: no live flock %u", flock_id);
    const Seg &s = b->segs[flock_id];
    const uint32_t n = s.n;
    if (!words) return set_error(FH_ERR_INVALID, "fh_batch_download: NULL words");
    if ((st = batch_settle(b))) return st;
    if (b->resident)  // this flock's slots back to bird order; the layout stays
        FH_HIP(launch_resident_flush_range(b->geo, b->P, b->R[b->parity], resident_slot_starts(b)[flock_id], n,
                                           b->stream));
    if ((st = pinned(b, staged_bytes(n)))) return st;
    const Staged h = staged(b->pin, n);
    const size_t o = s.bird_base;
    FH_HIP(hipMemcpyAsync(h.pos, b->P.pos4 + o, n * sizeof(float4), hipMemcpyDeviceToHost, b->stream));
    FH_HIP(hipMemcpyAsync(h.vel, b->P.vel4 + o, n * sizeof(float4), hipMemcpyDeviceToHost, b->stream));
    FH_HIP(hipMemcpyAsync(h.steer, b->P.steer4 + o, n * sizeof(float4), hipMemcpyDeviceToHost, b->stream));
    FH_HIP(hipMemcpyAsync(h.spin, b->P.spin + o, n * 4, hipMemcpyDeviceToHost, b->stream));
    FH_HIP(hipMemcpyAsync(h.cen, b->P.cen + o, n * 4, hipMemcpyDeviceToHost, b->stream));
    FH_HIP(hipMemcpyAsync(h.nb, b->P.nb + o, n * 4, hipMemcpyDeviceToHost, b->stream));
    FH_HIP(hipStreamSynchronize(b->stream));
    convert(b, n, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            float *w = words + i * FH_WORDS;
            w[PX] = h.pos[i].x; w[PY] = h.pos[i].y; w[PZ] = h.pos[i].z;
            w[VX] = h.vel[i].x; w[VY] = h.vel[i].y; w[VZ] = h.vel[i].z;
            w[ROLL] = h.vel[i].w;
            w[SPIN] = h.spin[i];
            w[RADIUS] = h.steer[i].w;
            w[SX] = h.steer[i].x; w[SY] = h.steer[i].y; w[SZ] = h.steer[i].z;
            w[NEXT] = h.pos[i].w;
            std::memcpy(&w[NB], &h.nb[i], 4);
            w[CEN] = h.cen[i];
        }
    });
    if (time) *time = b->times[flock_id];
    return FH_OK;
}

fh_status fh_batch_limit(fh_batch *b, uint32_t flock_id, uint32_t steps) {
    clear_error();
    fh_status st = bcheck(b, "fh_batch_limit");
    if (st) return st;
    if (!(b->prm.flags & FH_FLAG_RESIDENT))
        return set_error(FH_ERR_INVALID, "fh_batch_limit: step budgets need FH_FLAG_RESIDENT");
    if (flock_id >= b->segs.size() || !b->segs[flock_id].live)
        return set_error(FH_ERR_INVALID, "fh_batch_limit: no live flock %u", flock_id);
    if ((st = batch_settle(b))) return st;  // budgets change only at a rebuild boundary
    b->segs[flock_id].left = steps;
    return push_tables(b, 0, 0);
}

fh_status fh_batch_step(fh_batch *b, uint32_t steps, uint32_t warmup, fh_pass_ms *out) {
    clear_error();
    fh_status st = bcheck(b, "fh_batch_step");
    return st ? st : batch_step(b, steps, warmup, out);
}

fh_status fh_batch_info(const fh_batch *b, uint32_t *live_flocks, uint32_t *live_birds, uint32_t *bucket_range,
                        size_t *device_bytes) {
    clear_error();
    fh_status st = bcheck(b, "fh_batch_info");
    if (st) return st;
    uint32_t lf = 0;
    for (const Seg &s : b->segs) lf += s.live ? 1u : 0u;
    if (live_flocks) *live_flocks = lf;
    if (live_birds) *live_birds = b->geo.n_slots;
    if (bucket_range) *bucket_range = b->geo.n_buckets;
    if (device_bytes) *device_bytes = b->arena_bytes;
    return FH_OK;
}

size_t fh_flock_bytes(uint32_t n, const fh_params *params) {
    fh_params prm;
    if (resolve_params(params, &prm, "fh_flock_bytes")) fh_params_default(&prm);
    const uint32_t buckets = flock_buckets(n, prm);
    const size_t birds = (size_t)((n + TILE - 1) / TILE) * TILE;
    const size_t bk = (size_t)((buckets + SCAN_TILE - 1) / SCAN_TILE) * SCAN_TILE;
    size_t bytes = birds * batch_bytes_per_bird() + bk * batch_bytes_per_bucket() + bk / 8 /*occupancy bits*/;
    if (prm.flags & FH_FLAG_RESIDENT)
        bytes += birds * resident_bytes_per_bird((prm.flags & FH_FLAG_FLIGHT) != 0) + bk * resident_bytes_per_bucket();
    if (prm.flags & FH_FLAG_LAZY) {  // two sets of mover list entries, and 4 arrays of list heads
        const size_t mb = mover_heads(birds);
        bytes += birds * 2 * 36 + mb * 4 * 20 + 64;
    }
    return bytes;
}

}  // extern "C"
