// Device data layout of the flock step and the launchers of its kernels.
//
// Bird state lives in bird-index order as three float4 vector fields plus
// three scalar fields (FP32 structure of arrays; the w lanes carry the scalar
// each pass reads next to that vector):
//   pos4   = (px, py, pz, next_decision)   integrate: read+write; decide: w += DU
//   vel4   = (vx, vy, vz, roll)            integrate: read+write
//   steer4 = (sx, sy, sz, radius)          decide: write (due birds); integrate: read
//   spin, neighbours, centrality           integrate / decide write only
// The grid is the reference's counting sort (cuda/flock.cu:105-111):
//   key[bird]        bucket of each bird (grid_histogram)
//   rank[bird]       arrival order of the bird inside its bucket
//   counts[bucket]   histogram, zeroed again by the scan
//   offsets[bucket]  exclusive scan, buckets + 1 entries
//   occ[bucket / 32] bit (bucket % 32) set when the bucket holds a bird (scan)
// and the scatter writes the sorted copies decide walks, in slot order, as one
// 32-byte record per slot (a candidate's heading shares its position's line):
//   sbird[2 slot]     = (px, py, pz, bird index bits)
//   sbird[2 slot + 1] = (hx, hy, hz, radius), h = hit_heading(velocity), the
//                       heading an aligned neighbour adds (model.rs:88-90),
//                       computed once per bird instead of once per hit
//   snext[slot] = next_decision (or next_decision - time, see Seg::time).
//
// Several flocks can share one set of pools ("segments"): flock f owns birds
// [bird_base, bird_base + n) and buckets [bucket_base, bucket_base + mask + 1).
// Bird ranges start on a multiple of TILE birds so each 256-lane workgroup of
// a bird-order pass belongs to exactly one flock (tile_seg[] maps tile ->
// flock). A single flock is segment 0 with tile_seg == nullptr.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "model.h"

namespace fh {

constexpr uint32_t TILE = 256;        // birds per workgroup of the bird-order passes
constexpr uint32_t SCAN_TILE = 4096;  // buckets per workgroup of the scan (256 lanes x 16)
constexpr uint32_t RES_TILE = 1024;   // buckets per workgroup of the resident rebuild scan (256 lanes x 4)
constexpr uint32_t CANON_CAP = 1024;  // largest bucket the per-slot canonical pass reorders
constexpr uint32_t CANON_BIG_CAP = 16384;  // largest bucket the LDS sort of big buckets reorders
constexpr uint32_t NO_SEG = 0xffffffffu;

// Entries of the big-bucket list for `slots` sorted birds: every listed
// bucket holds more than CANON_CAP of them.
constexpr uint32_t big_list_cap(uint32_t slots) { return slots / (CANON_CAP + 1u) + 1u; }

// Error bits a kernel may raise in the device error word.
constexpr uint32_t ERR_SLOT = 1u;     // a scatter slot fell outside [0, n_slots)
constexpr uint32_t ERR_CANON = 2u;    // a bucket exceeded CANON_BIG_CAP: left in arrival order
constexpr uint32_t ERR_NBCAP = 4u;    // neighbour list longer than the caller's buffer

struct Seg {
    uint32_t bird_base;    // first bird of the flock in the pools
    uint32_t n;            // birds
    uint32_t bucket_base;  // first bucket of the flock
    uint32_t mask;         // buckets - 1 (a power of two minus one)
    float cell;            // grid cell edge [m]
    float inv_cell;        // 1 / cell
    float time;            // clock of the step in flight
    uint32_t live;         // 0 for a freed segment
    uint32_t mode;         // SEG_* bits
    uint32_t left = 0xffffffffu;  // resident steps left (SEG_UNLIMITED: no budget); 0: frozen
    uint32_t slot_base = 0;       // first slot of the flock's records in a resident layout
};

constexpr uint32_t SEG_SPATIAL = 1u;  // buckets by spatial_hash instead of the reference hash
constexpr uint32_t SEG_UNLIMITED = 0xffffffffu;

// Bucket of cell (cx, cy, cz) in segment sg.
__device__ __forceinline__ uint32_t bucket_of_cell(const Seg &sg, int32_t cx, int32_t cy, int32_t cz) {
    const uint32_t h = (sg.mode & SEG_SPATIAL) ? spatial_hash(cx, cy, cz) : bucket_hash(cx, cy, cz);
    return sg.bucket_base + (h & sg.mask);
}

struct Pools {
    float4 *pos4, *vel4, *steer4;
    float *spin, *cen;
    uint32_t *nb;
    uint32_t *key, *rank;
    uint32_t *counts, *offsets, *tile_sums, *tile_prefix;
    uint32_t *occ;         // occupancy bitmap of the buckets (one bit each)
    float4 *sbird;         // sorted records, two float4 per slot (see above)
    float *snext;
    uint32_t *sorted_tmp;  // canonical mode: slot -> bird before reordering
    uint32_t *err;         // device error word (ERR_* bits)
    uint32_t *due_count;   // birds that decided in the current step (zeroed by the scan)
    uint32_t *work;        // persistent decide: next chunk to take (zeroed by the scan)
    uint32_t *big;         // canonical mode: buckets with more than CANON_CAP birds
    uint32_t *big_count;   // entries of `big` this step (zeroed by the scan)
    uint32_t big_cap;      // capacity of `big`: big_list_cap(bird capacity)
};

// forward_of(v) (model.rs:15-20) as a walk's hit sums it: v times the
// hardware reciprocal square root of v.v (1 ulp) instead of a correctly
// rounded root and division. The resident walks compute it per aligned hit
// from the neighbour's velocity; the scatter stores it in the sorted copies,
// so both paths add the same bits.
__device__ __forceinline__ f3 hit_heading(f3 v) {
    const float l2 = dot3(v, v);
    return l2 > 1e-12f ? v * __builtin_amdgcn_rsqf(l2) : mk3(1.f, 0.f, 0.f);
}

// [offsets[b], offsets[b + 1]) with one 8-byte load: the two words share a
// cache line (except when b is the last word of one), so one request fetches
// both instead of two requests for the same line. gfx950 loads dword-aligned
// 8-byte values natively.
__device__ __forceinline__ uint2 bucket_range(const uint32_t *offsets, uint32_t b) {
    uint2 r;
    __builtin_memcpy(&r, offsets + b, sizeof r);
    return r;
}

// Resident layout (one flock, fh_step with FH_FLAG_RESIDENT): the state
// itself is kept in slot order between steps, in two buffers the steps
// alternate between. Slot s holds bird id[s], in bucket key[s]:
//   p4 = (px, py, pz, next_decision), v4 = (vx, vy, vz, roll),
//   s4 = (sx, sy, sz, radius), spin.
struct Records {
    float4 *p4, *v4, *s4;
    float *spin;
    uint32_t *id, *key;
};

// Per-step work arrays of the resident rebuild. A bird whose new bucket
// differs from its old one ("mover") counts itself out of the old bucket
// (dep) and into the new one (arr, returning its arrival rank).
struct ResWork {
    uint32_t *arank;  // per old slot (movers): arrival rank in the new bucket
    uint32_t *mbits;  // per old slot: mover bit (32 slots per word)
    uint32_t *arr;     // per bucket: arrivals (zero between steps)
    uint32_t *counts;  // per bucket: birds in the layout (the flight moves each mover's count)
    uint32_t *abase;      // per bucket: first slot of its arrivals in the new layout
    uint32_t *tsum;       // per RES_TILE buckets: their count total, kept from step to step
    uint32_t *tsum2;      // per RES_SUPER tiles: their total, summed by the rebuild (large layouts only)
    // Mover lists (FH_FLAG_LAZY): between rebuilds the records stay in the
    // buckets of the last rebuild, flown in place, and every bird whose cell
    // now hashes to another bucket than its slot's (`key`) is also listed by
    // its current bucket & mmask. The first bird of a list sits in the list's
    // head record itself, head4[m] = (position, slot in w), so a walk reads it
    // with the head; w = MOVER_NIL: empty list; MOVER_OVF set in w: more birds
    // follow from ovf[m] through next[] (entries indexed by slot: position
    // with the slot in w, p4; velocity, v4, also for the head's bird). A step
    // reads the lists its previous step wrote (mr_*; none when mr_head4 is
    // null), writes the next step's (mw_*), and clears the heads that a later
    // step of the period writes (mc_*).
    const float4 *mr_head4;
    const uint32_t *mr_ovf, *mr_next;
    const float4 *mr_p4, *mr_v4;
    float4 *mw_head4;
    uint32_t *mw_ovf, *mw_next;
    float4 *mw_p4, *mw_v4;
    float4 *mc_head4[2];
    uint32_t *mc_ovf[2];
    uint32_t mmask;     // lists - 1 (a power of two)
    // The clock of a lazy flock: the step reads it from mclock_now (its due
    // test) and writes the next step's to mclock_next; a step without a
    // rebuild also commits it to the flock's segment, mtick (a rebuild's
    // scan does that otherwise).
    const float *mclock_now;
    float *mclock_next;
    Seg *mtick;
    // A batch's lazy steps (mclock_now null): the steps taken since the
    // segment table was last committed (the rebuild commits them all), from
    // which each flock's clock and budget follow (seg_step).
    uint32_t mphase;
};
constexpr uint32_t MOVER_NIL = 0xffffffffu;
constexpr uint32_t MOVER_OVF = 0x80000000u;  // in a head's slot word: the list continues at ovf

// Steps per rebuild with FH_FLAG_LAZY.
constexpr uint32_t LAZY_PERIOD = 4;

// Mover list heads per phase: the power of two at or above one per 8 birds,
// from 4096 up to MOVER_HEADS_MAX (2^21 measured 0.8% faster than 2^20 and
// 2^22 on a 9.3M-bird batch of 256 flocks; a 1M-bird flock's 2^17 is flat
// between 2^16 and 2^18).
constexpr uint32_t MOVER_HEADS_MAX = 1u << 21;
inline uint32_t mover_heads(size_t birds) {
    uint32_t mb = 4096;
    while (mb < birds / 8 && mb < MOVER_HEADS_MAX) mb <<= 1;
    return mb;
}

// The mover lists of one flock or one batch (FH_FLAG_LAZY): `heads` list
// heads per phase (a power of two), LAZY_PERIOD arrays of them (phase 0's stay
// empty), then LAZY_PERIOD arrays of overflow heads in the same region (one
// fill of 5 * LAZY_PERIOD * heads words clears both), and two sets of list
// entries by slot (position, velocity, next), steps alternating between them.
struct MoverLists {
    uint32_t heads = 0;
    float4 *head4 = nullptr;
    uint32_t *ovf = nullptr;
    float4 *p4[2]{}, *v4[2]{};
    uint32_t *next[2]{};
};

// The list fields of `w` for a step at `phase` of the period: it reads the
// lists the step before wrote (heads of its phase; none at phase 0), writes
// the next phase's (not at the rebuild, the period's last step), and clears
// the heads a later phase writes: the first step clears the last writer's,
// the rebuild the others. Entries alternate between the two sets (written at
// phases 0 and 2, and 1).
inline ResWork with_mover_lists(ResWork w, const MoverLists &m, uint32_t phase) {
    static_assert(LAZY_PERIOD == 4, "the rebuild clears two head arrays: phases 1 and 2");
    const uint32_t nh = m.heads;
    auto head = [&](uint32_t k) { return m.head4 + (size_t)k * nh; };
    auto ovf = [&](uint32_t k) { return m.ovf + (size_t)k * nh; };
    const uint32_t rd = (phase + 1u) % 2u, wr = phase % 2u;
    w.mr_head4 = phase ? head(phase) : nullptr;
    w.mr_ovf = ovf(phase);
    w.mr_next = m.next[rd];
    w.mr_p4 = m.p4[rd];
    w.mr_v4 = m.v4[rd];
    const bool writes = phase + 1u < LAZY_PERIOD;
    w.mw_head4 = writes ? head(phase + 1u) : nullptr;
    w.mw_ovf = writes ? ovf(phase + 1u) : nullptr;
    w.mw_next = m.next[wr];
    w.mw_p4 = m.p4[wr];
    w.mw_v4 = m.v4[wr];
    w.mc_head4[0] = w.mc_head4[1] = nullptr;
    w.mc_ovf[0] = w.mc_ovf[1] = nullptr;
    if (phase == 0) {
        w.mc_head4[0] = head(LAZY_PERIOD - 1u);
        w.mc_ovf[0] = ovf(LAZY_PERIOD - 1u);
    } else if (phase + 1u == LAZY_PERIOD) {
        for (uint32_t k = 1; k + 1u < LAZY_PERIOD; ++k) {  // phases 1 and 2
            w.mc_head4[k - 1u] = head(k);
            w.mc_ovf[k - 1u] = ovf(k);
        }
    }
    w.mmask = nh - 1u;
    return w;
}

// Up to RES_DIRECT_TILES tiles, each workgroup of the rebuild scan adds up the
// totals of the tiles before it itself (at most 8 loads per lane, one launch).
// Beyond, a first launch sums the totals of each RES_SUPER consecutive tiles,
// and a workgroup adds up the super totals before its super tile and the
// tile totals before it inside its super tile: a bounded number of loads per
// workgroup, so the scan stays linear in the buckets.
constexpr uint32_t RES_DIRECT_TILES = 2048;
constexpr uint32_t RES_SUPER = 256;

// Launch geometry of one build/step over the pools.
struct Geometry {
    const Seg *segs;           // device segment table
    const uint32_t *tile_seg;  // device tile -> segment map, nullptr for one flock
    uint32_t n_birds;          // bird-index range covered (sum of padded ranges)
    uint32_t n_slots;          // birds actually sorted (sum of n)
    uint32_t n_buckets;        // total buckets over all segments
    uint32_t tile_begin;       // first tile of a bird-order launch (0 = from the start)
    uint32_t tile_count;       // tiles of a bird-order launch (0 = up to n_birds)
};

// Segment (flock) of the bird in slot s of resident records: through its bird
// id and the tile -> segment map in a batch, segment 0 for one flock.
__device__ __forceinline__ uint32_t slot_segment(const Geometry &g, const Records &rec, uint32_t s) {
    return g.tile_seg ? g.tile_seg[rec.id[s] / TILE] : 0u;
}

// Advances the clock of every live segment with steps left by DT and spends
// one step of its budget. A segment with no steps left is frozen: its birds
// are neither due nor moved, and its clock stays. Segments are independent:
// thread `lane` of `lanes` (a grid-stride split over the launch) commits
// segments lane, lane + lanes, ..., so a batch's clocks take one round of
// loads instead of one dependent round per flock.
__device__ __forceinline__ void commit_clocks(Seg *segs, uint32_t nseg, uint32_t lane, uint32_t lanes,
                                              uint32_t steps = 1u) {
    for (uint32_t i = lane; i < nseg; i += lanes) {
        Seg &sg = segs[i];
        if (!sg.live) continue;
        float t = sg.time;
        uint32_t left = sg.left;
        for (uint32_t k = 0; k < steps && left; ++k) {  // one f32 addition per step, as the host mirror
            t = t + DT;
            if (left != SEG_UNLIMITED) --left;
        }
        sg.time = t;
        sg.left = left;
    }
}

// A flock's clock and whether it steps, `k` steps after its segment entry was
// last committed: the clock of k committed steps (one DT addition each, while
// its budget lasts) and whether steps are left after them. With k = 0: the
// entry's own clock, and whether the flock is not frozen.
__device__ __forceinline__ bool seg_step(const Seg &sg, uint32_t k, float &time) {
    float t = sg.time;
    uint32_t left = sg.left;
    for (uint32_t i = 0; i < k && left; ++i) {
        t = t + DT;
        if (left != SEG_UNLIMITED) --left;
    }
    time = t;
    return left != 0u;
}

// Launchers (stream-ordered, no host synchronisation). Each returns the
// hipError_t of its launch.
hipError_t launch_hash_count(const Geometry &g, const Pools &p, bool wave_agg, hipStream_t s);
// Exclusive scan of the counts into offsets (and the occupancy bitmap). With
// tick_nseg > 0 it also advances the clocks of the live segments among
// tick_segs[0, tick_nseg) by dt before any later launch reads them. The
// counts are zeroed for the next histogram unless keep_counts.
hipError_t launch_scan(const Geometry &g, const Pools &p, Seg *tick_segs, uint32_t tick_nseg, float dt,
                       hipStream_t s, bool keep_counts = false);
hipError_t launch_scatter(const Geometry &g, const Pools &p, bool canonical, bool relative_time,
                          hipStream_t s);
// Walk switches of decide (bit set = enabled).
constexpr int WALK_TILED = 1;  // dense neighbourhoods through LDS
constexpr int WALK_OCC = 2;    // skip empty buckets by the occupancy bitmap
constexpr int WALK_PERSIST = 4;  // a fixed grid of workgroups takes chunks from a counter
constexpr int WALK_XCD = 8;      // each XCD's workgroups take one contiguous run of chunks
constexpr int WALK_WINDOW = 16;  // stage the positions of the chunk and a halo in LDS
constexpr int WALK_FLIGHT = 32;  // resident: fly each chunk's birds right after deciding them
constexpr int WALK_LAZY = 64;    // fused flight of a step without a rebuild: movers go to the mover lists
constexpr int WALK_MOVERS = 128;  // the walks also visit the mover lists (FH_FLAG_LAZY; resident, fused flight)
// decide's LDS window: the chunk and a halo of half a chunk on each side, at
// most WINDOW_MAX slots and within DECIDE_LDS_MAX bytes of dynamic LDS together
// with the chunk's due-slot list (32 bytes per slot: position and velocity).
constexpr uint32_t WINDOW_MAX = 4096;
constexpr uint32_t DECIDE_LDS_MAX = 52u << 10;  // + at most 8.3 KB static: under 64 KB
FH_HD uint32_t window_cap(uint32_t chunk) {
    const uint32_t room = chunk * 4u < DECIDE_LDS_MAX ? (DECIDE_LDS_MAX - chunk * 4u) / 32u : 0u;
    return room < WINDOW_MAX ? room : WINDOW_MAX;
}
FH_HD uint32_t window_halo(uint32_t chunk) {
    const uint32_t cap = window_cap(chunk);
    return chunk >= cap ? 0u : (chunk / 2 < (cap - chunk) / 2 ? chunk / 2 : (cap - chunk) / 2);
}
FH_HD uint32_t window_slots(uint32_t chunk) {
    const uint32_t w = chunk + 2u * window_halo(chunk), cap = window_cap(chunk);
    return w < cap ? w : cap;
}
// Dynamic LDS of one decide workgroup: the due-slot list, then the window.
inline size_t decide_lds_bytes(uint32_t chunk, int walk_flags) {
    return chunk * sizeof(uint32_t) + ((walk_flags & WALK_WINDOW) ? 2u * window_slots(chunk) * sizeof(float4) : 0u);
}
// XCDs of an MI355X: workgroup i of a launch is dispatched to XCD i % 8, and
// each XCD has its own L2.
constexpr uint32_t XCDS = 8;
// Events the decide launch records itself (kernel start and end), or none.
struct DecideTiming {
    hipEvent_t start = nullptr, stop = nullptr;
};
hipError_t launch_decide(const Geometry &g, const Pools &p, float time, uint32_t block,
                         uint32_t birds_per_thread, uint32_t group, int walk_flags, hipStream_t s,
                         const DecideTiming &tm = DecideTiming{});
// decide over resident records (p.offsets: the current offsets). Adds the
// due birds to *p.due_count, takes chunks from *p.work (persistent), and
// zeroes *due_clear and *work_clear (the other step's counters). With
// WALK_FLIGHT, each chunk's birds then fly (resident_flight) from rec into
// `flown` (positions, velocities, roll rates) with the mover accounting in w;
// rec itself keeps the positions every walk reads.
hipError_t launch_decide_resident(const Geometry &g, const Pools &p, const Records &rec, uint32_t block,
                                  uint32_t birds_per_thread, uint32_t group, int walk_flags, uint32_t *due_clear,
                                  uint32_t *work_clear, const ResWork &w, const Records &flown, hipStream_t s,
                                  const DecideTiming &tm = DecideTiming{});
// Resident passes (one flock): records from the bird-order state and a grid
// built by hash + scan (counts kept); banked flight in slot order with the
// movers counted; the rebuild of the grid and the move of every record to its
// slot in `nxt`; records back to the bird-order state. The rebuild reads the
// records to move from `cur` and commits the clocks and step budgets of the
// ntick segments in `tick` (+= DT for each live one with steps left): the
// flight, fused into decide or not, reads the budgets, so the commit waits
// for the launch after it. In a batch (g.tile_seg set) every slot's flock is
// found through its bird id; the enter pass skips the padding birds of each
// flock's range.
hipError_t launch_resident_enter(const Geometry &g, const Pools &p, const Records &rec, hipStream_t s);
// The enter pass for the birds [bird_lo, bird_lo + count) only (a flock that
// joins a resident batch).
hipError_t launch_resident_enter_range(const Geometry &g, const Pools &p, const Records &rec, uint32_t bird_lo,
                                       uint32_t count, hipStream_t s);
// A batch's resident records when flocks join or leave: slot s of `cur`
// (s < n_cur) moves to s + shift[its flock] in `nxt`, and the records of a
// flock whose shift is SPLICE_DROP are left out; n_nxt bounds the new slots.
constexpr int32_t SPLICE_DROP = INT32_MIN;
hipError_t launch_resident_splice(const Geometry &g, const Pools &p, const Records &cur, const Records &nxt,
                                  const int32_t *shift, uint32_t n_cur, uint32_t n_nxt, hipStream_t s);
// Each resident scan tile's count total (after the entry scan, counts kept).
hipError_t launch_resident_tile_sums(const Geometry &g, const Pools &p, const ResWork &w, hipStream_t s);
hipError_t launch_resident_integrate(const Geometry &g, const Pools &p, const Records &rec, const ResWork &w,
                                     hipStream_t s);
// tick_steps: the steps each ticked clock and budget commits (the steps of a
// lazy batch's period, or of its part before a settle; 1 otherwise).
hipError_t launch_resident_rebuild(const Geometry &g, const Pools &p, const Records &cur, const Records &nxt,
                                   const ResWork &w, const uint32_t *off_cur, uint32_t *off_nxt, Seg *tick,
                                   uint32_t ntick, hipStream_t s, uint32_t tick_steps = 1u);
hipError_t launch_resident_flush(const Geometry &g, const Pools &p, const Records &rec, hipStream_t s);
// The movers of the current records `cur` against the buckets of their slots,
// counted as a fused flight counts them at a rebuild (arrival ranks, counts,
// tile totals, mover bits), without flying: with launch_resident_rebuild
// after it (no tick), the layout of the current positions.
hipError_t launch_resident_account(const Geometry &g, const Pools &p, const Records &cur, const ResWork &w,
                                   hipStream_t s);
// The flush of the slots [slot_lo, slot_lo + count) only (one flock of a batch).
hipError_t launch_resident_flush_range(const Geometry &g, const Pools &p, const Records &rec, uint32_t slot_lo,
                                       uint32_t count, hipStream_t s);
hipError_t launch_integrate(const Geometry &g, const Pools &p, bool hash, bool wave_agg,
                            hipStream_t s);
hipError_t launch_neighbours(const Geometry &g, const Pools &p, uint32_t bird, int walk,
                             uint32_t *ids, uint32_t cap, uint32_t *count, hipStream_t s);
hipError_t launch_set_time(Seg *segs, uint32_t nseg, float time, float dt, hipStream_t s);
// *clock = time (a lazy flock's clock, before a run of steps).
hipError_t launch_set_clock(float *clock, float time, hipStream_t s);

// Resource usage of the kernels, for reports (name: "decide", "integrate",
// "scatter", "hash_count", "scan").
hipError_t kernel_attributes(const char *name, hipFuncAttributes *out);

}  // namespace fh
