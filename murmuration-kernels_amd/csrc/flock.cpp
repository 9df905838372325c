// One flock behind the reference pass API: the handle owns one device arena,
// one HIP stream (default priority) and the host copy of the flock clock.
// Pass order and clock follow cuda/flock.cu:452-470.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "host_util.h"
#include "layout.h"
#include "model.h"

namespace fh {
hipError_t grid_kernel_attributes(const char *name, hipFuncAttributes *out);
hipError_t decide_kernel_attributes(const char *name, hipFuncAttributes *out);
hipError_t integrate_kernel_attributes(const char *name, hipFuncAttributes *out);
hipError_t launch_fill(uint32_t *dst, uint32_t count, uint32_t value, hipStream_t s);
}  // namespace fh

using namespace fh;

// Grid state of the current positions.
enum class Stage { None, Hashed, Scanned, Scattered, Decided };

struct fh_flock {
    uint32_t n = 0, buckets = 0, padded_buckets = 0;
    fh_params prm{};
    float time = 0.0f;
    hipStream_t stream = nullptr;
    void *arena = nullptr;
    Pools P{};
    Seg seg{};
    Seg *d_seg = nullptr;
    Geometry geo{};
    Stage stage = Stage::None;
    bool counts_dirty = false;  // histogram filled and not yet consumed by a scan
    bool absolute_due = true;   // snext holds next_decision (not next - time)
    uint32_t *d_nbr = nullptr;  // neighbour read-back buffer
    uint32_t nbr_cap = 0;
    hipGraphExec_t graph = nullptr;
    // Resident layout (FH_FLAG_RESIDENT): records R[q] and offsets off[q] of
    // the current layout, q = parity; a step moves them to the other buffers.
    void *res_arena = nullptr;
    Records R[2]{};
    Records F{};  // fused flight: the flown positions, velocities and roll rates
    ResWork W{};
    uint32_t *off[2]{};
    uint32_t *res_due = nullptr;  // due counter of each parity, then the persistent decide's chunk counters
    bool resident = false;
    uint32_t parity = 0;
    hipGraphExec_t rgraph[2]{};
    hipGraphExec_t graph_run[2]{};   // GRAPH_RUNS[i] fused steps from the Hashed stage
    hipGraphExec_t rgraph_run[2]{};  // GRAPH_RUNS[i] resident steps from parity 0
    // FH_FLAG_LAZY: the layout is rebuilt every LAZY_PERIOD steps. In between,
    // each step flies the records in place (alternating between R[parity]'s
    // positions, velocities and roll rates and F's), and appends its movers
    // to the mover lists, which the next step's walks visit.
    bool lazy = false;
    uint32_t phase = 0;  // steps since the last rebuild
    uint32_t fq = 0;     // 1: the current positions, velocities and roll rates are F's
    uint32_t cpar = 0;   // due-counter parity of the next step (the step count's parity)
    uint32_t mheads = 0;                   // list heads per phase (a power of two)
    float4 *mhead4 = nullptr;              // LAZY_PERIOD arrays of list heads (phase 0's stay empty),
    uint32_t *movf = nullptr;              // then LAZY_PERIOD arrays of overflow heads (one region)
    float *mclock = nullptr;               // the clock by step parity (cpar)
    float4 *mp4[2]{}, *mv4[2]{};           // two sets of list entries (by slot): steps alternate
    uint32_t *mnext[2]{};
    hipGraphExec_t lgraph_run[2][2][2]{};  // GRAPH_RUNS[i] lazy steps from phase 0, layout parity, counter parity
};

// LAZY_PERIOD (steps per rebuild with FH_FLAG_LAZY) divides every GRAPH_RUNS
// entry, so that a graph run starts and ends on a rebuild.

// Events per timed step with per-pass timing: six on the stream between the
// passes, two recorded by the decide launch at its kernel's start and end.
constexpr size_t EV_PER_STEP = 8;

static fh_status check_handle(const fh_flock *h, const char *fn) {
    if (!h) return set_error(FH_ERR_INVALID, "%s: NULL handle", fn);
    return FH_OK;
}

static fh_status read_error_word(fh_flock *h) {
    uint32_t err = 0;
    FH_HIP(hipMemcpyAsync(&err, h->P.err, 4, hipMemcpyDeviceToHost, h->stream));
    FH_HIP(hipStreamSynchronize(h->stream));
    if (err & ERR_SLOT) return set_error(FH_ERR_HIP, "kernel reported a slot outside the sorted range (err 0x%x)", err);
    if (err & ERR_NBCAP) return set_error(FH_ERR_HIP, "kernel reported error 0x%x", err);
    return FH_OK;
}

// ---- passes on the handle's stream -----------------------------------------

static fh_status do_hash(fh_flock *h) {
    if (h->counts_dirty) FH_HIP(launch_fill(h->P.counts, h->padded_buckets, 0u, h->stream));
    FH_HIP(launch_hash_count(h->geo, h->P, (h->prm.flags & FH_FLAG_WAVE_AGG) != 0, h->stream));
    h->counts_dirty = true;
    h->stage = Stage::Hashed;
    return FH_OK;
}

static fh_status do_scan(fh_flock *h, float tick = 0.0f) {
    FH_HIP(launch_scan(h->geo, h->P, h->d_seg, tick != 0.0f ? 1u : 0u, tick, h->stream));
    h->counts_dirty = false;
    h->stage = Stage::Scanned;
    return FH_OK;
}

static fh_status do_scatter(fh_flock *h, bool relative) {
    FH_HIP(launch_scatter(h->geo, h->P, (h->prm.flags & FH_FLAG_CANONICAL) != 0, relative, h->stream));
    h->absolute_due = !relative;
    h->stage = Stage::Scattered;
    return FH_OK;
}

static fh_status do_decide(fh_flock *h, float time_arg, const DecideTiming &tm = DecideTiming{}) {
    FH_HIP(launch_decide(h->geo, h->P, time_arg, h->prm.decide_block, h->prm.birds_per_thread,
                         h->prm.decide_group, walk_flags(h->prm.flags), h->stream, tm));
    h->stage = Stage::Decided;
    return FH_OK;
}

static fh_status do_integrate(fh_flock *h, bool hash) {
    if (hash && h->counts_dirty) FH_HIP(launch_fill(h->P.counts, h->padded_buckets, 0u, h->stream));
    FH_HIP(launch_integrate(h->geo, h->P, hash, (h->prm.flags & FH_FLAG_WAVE_AGG) != 0, h->stream));
    if (hash) {
        h->counts_dirty = true;
        h->stage = Stage::Hashed;
    } else {
        h->stage = Stage::None;
    }
    return FH_OK;
}

// One step after time += DT, in the relative-due form (scatter stores
// next - time, decide tests <= 0) so the same launches replay in a graph.
static fh_status step_launches(fh_flock *h, bool fused, hipEvent_t *ev) {
    if (ev) FH_HIP(hipEventRecord(ev[0], h->stream));
    if (!fused || h->stage != Stage::Hashed) {
        fh_status st = do_hash(h);
        if (st) return st;
    }
    if (ev) FH_HIP(hipEventRecord(ev[1], h->stream));
    fh_status st = do_scan(h, DT);  // the scan also advances the device clock by DT
    if (st) return st;
    if (ev) FH_HIP(hipEventRecord(ev[2], h->stream));
    if ((st = do_scatter(h, true))) return st;
    if (ev) FH_HIP(hipEventRecord(ev[3], h->stream));
    // the decide launch records ev[6] and ev[7] at its kernel's start and end
    if ((st = do_decide(h, 0.0f, ev ? DecideTiming{ev[6], ev[7]} : DecideTiming{}))) return st;
    if (ev) FH_HIP(hipEventRecord(ev[4], h->stream));
    if ((st = do_integrate(h, fused))) return st;
    if (ev) FH_HIP(hipEventRecord(ev[5], h->stream));
    return FH_OK;
}

static fh_status sync_clock(fh_flock *h) {
    // The device clock must equal the host clock before a relative-due step.
    FH_HIP(launch_set_time(h->d_seg, 1, h->time, 0.0f, h->stream));
    return FH_OK;
}

static fh_status capture_graph(fh_flock *h) {
    if (h->graph) return FH_OK;
    // The graph body is the fused step from the Hashed stage.
    hipGraph_t g = nullptr;
    FH_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const Stage stage0 = h->stage;
    const bool dirty0 = h->counts_dirty;
    fh_status st = step_launches(h, true, nullptr);
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    h->stage = stage0;
    h->counts_dirty = dirty0;
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        h->graph = nullptr;
        return hip_fail(e, "hipGraphInstantiate");
    }
    return FH_OK;
}

// ---- resident layout ------------------------------------------------------

static fh_status alloc_resident(fh_flock *h) {
    if (h->res_arena) return FH_OK;
    const size_t nn = h->n ? h->n : 1;
    const size_t nb = h->padded_buckets;
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
    const bool flight = (h->prm.flags & FH_FLAG_FLIGHT) != 0;
    const size_t o_fp4 = flight ? a.take<float4>(nn) : 0, o_fv4 = flight ? a.take<float4>(nn) : 0;
    const size_t o_fspin = flight ? a.take<float>(nn) : 0;
    const size_t o_off1 = a.take<uint32_t>(nb + 1), o_due = a.take<uint32_t>(4);
    const size_t o_tsum = a.take<uint32_t>(nb / RES_TILE);
    const size_t o_tsum2 = a.take<uint32_t>(nb / RES_TILE / RES_SUPER + 1);
    // mover lists: about one list head per 8 birds, per phase
    const bool lazy = (h->prm.flags & FH_FLAG_LAZY) != 0;
    const uint32_t mb = lazy ? mover_heads(nn) : 4096u;
    const size_t o_mhead = lazy ? a.take<float4>((size_t)LAZY_PERIOD * mb) : 0;
    const size_t o_movf = lazy ? a.take<uint32_t>((size_t)LAZY_PERIOD * mb) : 0;  // right after: one fill clears both
    const size_t o_mclock = lazy ? a.take<float>(2) : 0;
    size_t o_mp4[2]{}, o_mv4[2]{}, o_mnext[2]{};
    for (int k = 0; lazy && k < 2; ++k) {
        o_mp4[k] = a.take<float4>(nn);
        o_mv4[k] = a.take<float4>(nn);
        o_mnext[k] = a.take<uint32_t>(nn);
    }
    hipError_t e = hipMalloc(&h->res_arena, a.used);
    if (e != hipSuccess) {
        h->res_arena = nullptr;
        return hip_fail(e, "hipMalloc(resident arena)");
    }
    char *base = static_cast<char *>(h->res_arena);
    for (int q = 0; q < 2; ++q) {
        h->R[q].p4 = reinterpret_cast<float4 *>(base + o_p4[q]);
        h->R[q].v4 = reinterpret_cast<float4 *>(base + o_v4[q]);
        h->R[q].s4 = reinterpret_cast<float4 *>(base + o_s4[q]);
        h->R[q].spin = reinterpret_cast<float *>(base + o_spin[q]);
        h->R[q].id = reinterpret_cast<uint32_t *>(base + o_id[q]);
        h->R[q].key = reinterpret_cast<uint32_t *>(base + o_key[q]);
    }
    h->W.arank = reinterpret_cast<uint32_t *>(base + o_arank);
    h->W.mbits = reinterpret_cast<uint32_t *>(base + o_mbits);
    h->W.arr = reinterpret_cast<uint32_t *>(base + o_arr);
    h->W.counts = h->P.counts;
    h->W.abase = reinterpret_cast<uint32_t *>(base + o_abase);
    h->W.tsum = reinterpret_cast<uint32_t *>(base + o_tsum);
    h->W.tsum2 = reinterpret_cast<uint32_t *>(base + o_tsum2);
    if (flight) {
        h->F.p4 = reinterpret_cast<float4 *>(base + o_fp4);
        h->F.v4 = reinterpret_cast<float4 *>(base + o_fv4);
        h->F.spin = reinterpret_cast<float *>(base + o_fspin);
    }
    h->off[0] = h->P.offsets;
    h->off[1] = reinterpret_cast<uint32_t *>(base + o_off1);
    h->res_due = reinterpret_cast<uint32_t *>(base + o_due);
    if (lazy) {
        h->lazy = true;
        h->mheads = mb;
        h->mhead4 = reinterpret_cast<float4 *>(base + o_mhead);
        h->movf = reinterpret_cast<uint32_t *>(base + o_movf);
        h->mclock = reinterpret_cast<float *>(base + o_mclock);
        for (int k = 0; k < 2; ++k) {
            h->mp4[k] = reinterpret_cast<float4 *>(base + o_mp4[k]);
            h->mv4[k] = reinterpret_cast<float4 *>(base + o_mv4[k]);
            h->mnext[k] = reinterpret_cast<uint32_t *>(base + o_mnext[k]);
        }
    }
    // arrivals start at zero and every rebuild zeroes them again
    FH_HIP(hipMemsetAsync(h->res_arena, 0, a.used, h->stream));
    FH_HIP(hipStreamSynchronize(h->stream));
    return FH_OK;
}

// Bird-order state -> resident records: the reference grid of the current
// positions (hash, scan with the counts kept), then every bird to its slot.
static fh_status enter_resident(fh_flock *h) {
    fh_status st = alloc_resident(h);
    if (st) return st;
    if (h->counts_dirty) FH_HIP(launch_fill(h->P.counts, h->padded_buckets, 0u, h->stream));
    FH_HIP(launch_hash_count(h->geo, h->P, (h->prm.flags & FH_FLAG_WAVE_AGG) != 0, h->stream));
    Pools p = h->P;
    p.offsets = h->off[0];
    FH_HIP(launch_scan(h->geo, p, nullptr, 0u, 0.0f, h->stream, true));
    FH_HIP(launch_resident_enter(h->geo, p, h->R[0], h->stream));
    FH_HIP(launch_resident_tile_sums(h->geo, p, h->W, h->stream));
    // parity 0's counters may hold a step taken before the last leave
    FH_HIP(launch_fill(h->res_due, 4, 0u, h->stream));
    if (h->lazy) {
        FH_HIP(launch_fill(reinterpret_cast<uint32_t *>(h->mhead4), 5u * LAZY_PERIOD * h->mheads, MOVER_NIL, h->stream));
    }
    h->phase = 0;
    h->fq = 0;
    h->cpar = 0;
    h->parity = 0;
    h->resident = true;
    h->counts_dirty = false;  // the counts now describe the current layout
    h->stage = Stage::None;
    return FH_OK;
}

// The records of the current step with FH_FLAG_LAZY: R[parity]'s, with the
// positions, velocities and roll rates F's when `f`.
static Records lazy_view(const fh_flock *h, uint32_t f) {
    Records r = h->R[h->parity];
    if (f) {
        r.p4 = h->F.p4;
        r.v4 = h->F.v4;
        r.spin = h->F.spin;
    }
    return r;
}

// With FH_FLAG_LAZY between rebuilds: the layout of the current positions
// (mover accounting against the slots' buckets, then the rebuild without a
// clock tick), so the records and the grid are those of a rebuilt step.
static fh_status settle_lazy(fh_flock *h) {
    if (!h->lazy || !h->resident || h->phase == 0) return FH_OK;
    const uint32_t q = h->parity;
    Pools p = h->P;
    p.offsets = h->off[q];
    const Records cur = lazy_view(h, h->fq);
    FH_HIP(launch_resident_account(h->geo, p, cur, h->W, h->stream));
    FH_HIP(launch_resident_rebuild(h->geo, p, cur, h->R[1u - q], h->W, h->off[q], h->off[1u - q], nullptr, 0u,
                                   h->stream));
    FH_HIP(launch_fill(reinterpret_cast<uint32_t *>(h->mhead4), 5u * LAZY_PERIOD * h->mheads, MOVER_NIL, h->stream));
    h->parity = 1u - q;
    h->fq = 0;
    h->phase = 0;
    return FH_OK;
}

// Resident records -> bird-order state (before any call that reads it).
static fh_status leave_resident(fh_flock *h) {
    if (!h->resident) return FH_OK;
    fh_status st = settle_lazy(h);
    if (st) return st;
    FH_HIP(launch_resident_flush(h->geo, h->P, h->R[h->parity], h->stream));
    h->resident = false;
    h->counts_dirty = true;  // the counts hold the last layout, not a fresh histogram
    h->stage = Stage::None;
    return FH_OK;
}

// One resident step: decide over the current layout, banked flight (movers
// counted), then the grid of the new positions and the records moved into it.
// Events (per-pass timing): ev[0] | decide (its kernel: ev[6] to ev[7]) |
// ev[1] | integrate | ev[2] | scan and move | ev[3] = ev[4] = ev[5].
static fh_status resident_step_launches(fh_flock *h, hipEvent_t *ev) {
    const uint32_t q = h->parity;
    Pools p = h->P;
    p.offsets = h->off[q];
    p.due_count = h->res_due + q;
    p.work = h->res_due + 2 + q;
    const bool fused = (h->prm.flags & FH_FLAG_FLIGHT) != 0;
    if (ev) FH_HIP(hipEventRecord(ev[0], h->stream));
    FH_HIP(launch_decide_resident(h->geo, p, h->R[q], h->prm.decide_block, h->prm.birds_per_thread,
                                  h->prm.decide_group, walk_flags(h->prm.flags), h->res_due + (1u - q),
                                  h->res_due + 2 + (1u - q), h->W, h->F, h->stream,
                                  ev ? DecideTiming{ev[6], ev[7]} : DecideTiming{}));
    if (ev) FH_HIP(hipEventRecord(ev[1], h->stream));
    if (!fused) FH_HIP(launch_resident_integrate(h->geo, p, h->R[q], h->W, h->stream));
    if (ev) FH_HIP(hipEventRecord(ev[2], h->stream));
    // the move reads the flown positions, velocities and roll rates from F
    // when the flight was fused, everything else from the current records
    Records src = h->R[q];
    if (fused) {
        src.p4 = h->F.p4;
        src.v4 = h->F.v4;
        src.spin = h->F.spin;
    }
    FH_HIP(launch_resident_rebuild(h->geo, p, src, h->R[1u - q], h->W, h->off[q], h->off[1u - q], h->d_seg, 1u,
                                   h->stream));
    if (ev)
        for (int k = 3; k < 6; ++k) FH_HIP(hipEventRecord(ev[k], h->stream));
    return FH_OK;
}

// The work arrays of a lazy step at `phase` of the period: the mover lists
// (with_mover_lists) and the flock's double-buffered clock.
static ResWork lazy_work(const fh_flock *h, uint32_t phase) {
    MoverLists m;
    m.heads = h->mheads;
    m.head4 = h->mhead4;
    m.ovf = h->movf;
    for (int k = 0; k < 2; ++k) {
        m.p4[k] = h->mp4[k];
        m.v4[k] = h->mv4[k];
        m.next[k] = h->mnext[k];
    }
    ResWork w = with_mover_lists(h->W, m, phase);
    w.mclock_now = h->mclock + h->cpar;
    w.mclock_next = h->mclock + (1u - h->cpar);
    w.mtick = h->d_seg;
    return w;
}

// One step with FH_FLAG_LAZY: decide (the walks visit the mover lists too)
// with the fused flight from the current records into the other set. Between
// rebuilds the flight appends its movers to the next step's lists, and the
// decide launch commits the clock itself; every LAZY_PERIOD-th step the
// flight counts the movers against the layout instead, and the rebuild
// follows (it commits the clock). Events as resident_step_launches.
static fh_status lazy_step_launches(fh_flock *h, hipEvent_t *ev) {
    const uint32_t q = h->parity, c = h->cpar;
    const bool rebuild = h->phase + 1u == LAZY_PERIOD;
    Pools p = h->P;
    p.offsets = h->off[q];
    p.due_count = h->res_due + c;
    p.work = h->res_due + 2 + c;
    const Records cur = lazy_view(h, h->fq), flown = lazy_view(h, h->fq ^ 1u);
    const int wf = walk_flags(h->prm.flags) | WALK_MOVERS | (rebuild ? 0 : WALK_LAZY);
    if (ev) FH_HIP(hipEventRecord(ev[0], h->stream));
    FH_HIP(launch_decide_resident(h->geo, p, cur, h->prm.decide_block, h->prm.birds_per_thread, h->prm.decide_group,
                                  wf, h->res_due + (1u - c), h->res_due + 2 + (1u - c), lazy_work(h, h->phase), flown,
                                  h->stream, ev ? DecideTiming{ev[6], ev[7]} : DecideTiming{}));
    if (ev) {
        FH_HIP(hipEventRecord(ev[1], h->stream));
        FH_HIP(hipEventRecord(ev[2], h->stream));
    }
    if (rebuild) {
        FH_HIP(launch_resident_rebuild(h->geo, p, flown, h->R[1u - q], h->W, h->off[q], h->off[1u - q], h->d_seg,
                                       1u, h->stream));
        h->parity = 1u - q;
        h->fq = 0;
        h->phase = 0;
    } else {
        h->fq ^= 1u;
        h->phase += 1u;
    }
    h->cpar = 1u - c;
    if (ev)
        for (int k = 3; k < 6; ++k) FH_HIP(hipEventRecord(ev[k], h->stream));
    return FH_OK;
}

// Replaying a graph costs a gap of several microseconds before its first
// kernel (none between the kernels inside it), so fh_step replays runs of
// GRAPH_RUNS[i] steps as one graph where it can, longest first. Even, so that
// a resident run ends at the buffer parity it started from.
constexpr uint32_t GRAPH_RUNS[2] = {16, 4};
static_assert(GRAPH_RUNS[0] % LAZY_PERIOD == 0 && GRAPH_RUNS[1] % LAZY_PERIOD == 0, "a lazy run holds whole periods");

// Captures `body` (launches on h->stream) into *out, once.
template <class Body>
static fh_status capture_once(fh_flock *h, hipGraphExec_t *out, Body &&body) {
    if (*out) return FH_OK;
    hipGraph_t g = nullptr;
    FH_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    fh_status st = body();
    hipError_t e = hipStreamEndCapture(h->stream, &g);
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

// GRAPH_RUNS[i] resident steps from parity 0 (the parity is restored).
static fh_status capture_resident_run(fh_flock *h, int i) {
    return capture_once(h, &h->rgraph_run[i], [&]() -> fh_status {
        const uint32_t q0 = h->parity;
        fh_status st = FH_OK;
        for (uint32_t k = 0; k < GRAPH_RUNS[i] && !st; ++k) {
            st = resident_step_launches(h, nullptr);
            h->parity ^= 1u;
        }
        h->parity = q0;
        return st;
    });
}

// GRAPH_RUNS[i] lazy steps from phase 0 at the current counter parity, for
// both layout parities at once (a run whose rebuild count is odd ends at the
// other one, so the next run needs that graph too: capturing it here keeps
// the capture out of later timed steps). The host state is restored; the
// run's own effect is applied by the caller after each replay.
static fh_status capture_lazy_run(fh_flock *h, int i) {
    const uint32_t q0 = h->parity, c0 = h->cpar;
    fh_status st = FH_OK;
    for (uint32_t q = 0; q < 2 && !st; ++q) {
        h->parity = q;
        st = capture_once(h, &h->lgraph_run[i][q][c0], [&]() -> fh_status {
            fh_status s2 = FH_OK;
            for (uint32_t k = 0; k < GRAPH_RUNS[i] && !s2; ++k) s2 = lazy_step_launches(h, nullptr);
            h->parity = q;
            h->cpar = c0;
            h->phase = 0;
            h->fq = 0;
            return s2;
        });
    }
    h->parity = q0;
    return st;
}

// GRAPH_RUNS[i] fused steps from the Hashed stage (stage and counts state restored).
static fh_status capture_graph_run(fh_flock *h, int i) {
    return capture_once(h, &h->graph_run[i], [&]() -> fh_status {
        const Stage stage0 = h->stage;
        const bool dirty0 = h->counts_dirty;
        fh_status st = FH_OK;
        for (uint32_t k = 0; k < GRAPH_RUNS[i] && !st; ++k) st = step_launches(h, true, nullptr);
        h->stage = stage0;
        h->counts_dirty = dirty0;
        return st;
    });
}

static fh_status capture_resident_graph(fh_flock *h) {
    const uint32_t q = h->parity;
    if (h->rgraph[q]) return FH_OK;
    hipGraph_t g = nullptr;
    FH_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    fh_status st = resident_step_launches(h, nullptr);
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&h->rgraph[q], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        h->rgraph[q] = nullptr;
        return hip_fail(e, "hipGraphInstantiate");
    }
    return FH_OK;
}

extern "C" {

fh_status fh_create(uint32_t n, const fh_params *params, fh_flock **out) {
    clear_error();
    if (!out) return set_error(FH_ERR_INVALID, "fh_create: NULL out");
    *out = nullptr;
    fh_params prm;
    fh_status st = resolve_params(params, &prm, "fh_create");
    if (st) return st;
    if (n > (1u << 30)) return set_error(FH_ERR_INVALID, "fh_create: n = %u too large", n);
    st = require_device();
    if (st) return st;

    fh_flock *h = new fh_flock();
    h->n = n;
    h->prm = prm;
    h->buckets = flock_buckets(n, prm);
    const uint32_t ntiles = (h->buckets + SCAN_TILE - 1) / SCAN_TILE;
    h->padded_buckets = ntiles * SCAN_TILE;

    Arena a;
    const size_t nn = n ? n : 1;
    size_t o_pos = a.take<float4>(nn), o_vel = a.take<float4>(nn), o_steer = a.take<float4>(nn);
    size_t o_sbird = a.take<float4>(2 * nn);
    size_t o_spin = a.take<float>(nn), o_cen = a.take<float>(nn), o_nb = a.take<uint32_t>(nn);
    size_t o_key = a.take<uint32_t>(nn), o_rank = a.take<uint32_t>(nn), o_snext = a.take<float>(nn);
    size_t o_tmp = a.take<uint32_t>(nn);
    size_t o_counts = a.take<uint32_t>(h->padded_buckets);
    size_t o_offsets = a.take<uint32_t>((size_t)h->padded_buckets + 1);
    size_t o_tsum = a.take<uint32_t>(ntiles), o_tpre = a.take<uint32_t>(ntiles);
    size_t o_err = a.take<uint32_t>(4), o_due = a.take<uint32_t>(4), o_seg = a.take<Seg>(1);
    size_t o_bigc = a.take<uint32_t>(4), o_big = a.take<uint32_t>(big_list_cap((uint32_t)nn));
    size_t o_occ = a.take<uint32_t>(h->padded_buckets / 32);
    hipError_t e = hipMalloc(&h->arena, a.used);
    if (e != hipSuccess) {
        delete h;
        return hip_fail(e, "hipMalloc(flock arena)");
    }
    char *base = static_cast<char *>(h->arena);
    h->P.pos4 = reinterpret_cast<float4 *>(base + o_pos);
    h->P.vel4 = reinterpret_cast<float4 *>(base + o_vel);
    h->P.steer4 = reinterpret_cast<float4 *>(base + o_steer);
    h->P.sbird = reinterpret_cast<float4 *>(base + o_sbird);
    h->P.spin = reinterpret_cast<float *>(base + o_spin);
    h->P.cen = reinterpret_cast<float *>(base + o_cen);
    h->P.nb = reinterpret_cast<uint32_t *>(base + o_nb);
    h->P.key = reinterpret_cast<uint32_t *>(base + o_key);
    h->P.rank = reinterpret_cast<uint32_t *>(base + o_rank);
    h->P.snext = reinterpret_cast<float *>(base + o_snext);
    h->P.sorted_tmp = reinterpret_cast<uint32_t *>(base + o_tmp);
    h->P.counts = reinterpret_cast<uint32_t *>(base + o_counts);
    h->P.offsets = reinterpret_cast<uint32_t *>(base + o_offsets);
    h->P.tile_sums = reinterpret_cast<uint32_t *>(base + o_tsum);
    h->P.tile_prefix = reinterpret_cast<uint32_t *>(base + o_tpre);
    h->P.err = reinterpret_cast<uint32_t *>(base + o_err);
    h->P.due_count = reinterpret_cast<uint32_t *>(base + o_due);
    h->P.work = h->P.due_count + 1;
    h->P.big_count = reinterpret_cast<uint32_t *>(base + o_bigc);
    h->P.big = reinterpret_cast<uint32_t *>(base + o_big);
    h->P.big_cap = big_list_cap((uint32_t)nn);
    h->P.occ = reinterpret_cast<uint32_t *>(base + o_occ);
    h->d_seg = reinterpret_cast<Seg *>(base + o_seg);

    h->seg.bird_base = 0;
    h->seg.n = n;
    h->seg.bucket_base = 0;
    h->seg.mask = h->buckets - 1;
    h->seg.cell = prm.cell;
    h->seg.inv_cell = 1.0f / prm.cell;
    h->seg.time = 0.0f;
    h->seg.live = 1;
    h->seg.mode = (prm.flags & FH_FLAG_SPATIAL) ? SEG_SPATIAL : 0u;
    h->geo.segs = h->d_seg;
    h->geo.tile_seg = nullptr;
    h->geo.n_birds = n;
    h->geo.n_slots = n;
    h->geo.n_buckets = h->buckets;

    auto fail = [&](hipError_t err, const char *what) {
        fh_status s2 = hip_fail(err, what);
        if (h->stream) (void)hipStreamDestroy(h->stream);
        (void)hipFree(h->arena);
        delete h;
        return s2;
    };
    if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail(e, "hipStreamCreateWithFlags");
    // Zero everything once so every index array starts in range.
    if ((e = hipMemsetAsync(h->arena, 0, a.used, h->stream)) != hipSuccess) return fail(e, "hipMemsetAsync");
    if ((e = hipMemcpyAsync(h->d_seg, &h->seg, sizeof(Seg), hipMemcpyHostToDevice, h->stream)) != hipSuccess)
        return fail(e, "hipMemcpyAsync(seg)");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return fail(e, "hipStreamSynchronize");
    *out = h;
    return FH_OK;
}

fh_status fh_destroy(fh_flock *h) {
    clear_error();
    if (!h) return FH_OK;
    (void)hipStreamSynchronize(h->stream);
    if (h->graph) (void)hipGraphExecDestroy(h->graph);
    for (int i = 0; i < 2; ++i) {
        if (h->graph_run[i]) (void)hipGraphExecDestroy(h->graph_run[i]);
        if (h->rgraph_run[i]) (void)hipGraphExecDestroy(h->rgraph_run[i]);
    }
    for (auto &gx : h->rgraph)
        if (gx) (void)hipGraphExecDestroy(gx);
    for (auto &by_q : h->lgraph_run)
        for (auto &by_c : by_q)
            for (auto &gx : by_c)
                if (gx) (void)hipGraphExecDestroy(gx);
    if (h->res_arena) (void)hipFree(h->res_arena);
    if (h->d_nbr) (void)hipFree(h->d_nbr);
    (void)hipFree(h->arena);
    (void)hipStreamDestroy(h->stream);
    delete h;
    return FH_OK;
}

fh_status fh_get_params(const fh_flock *h, fh_params *out, uint32_t *buckets) {
    clear_error();
    fh_status st = check_handle(h, "fh_get_params");
    if (st) return st;
    if (out) *out = h->prm;
    if (buckets) *buckets = h->buckets;
    return FH_OK;
}

fh_status fh_upload_words(fh_flock *h, const float *words, uint32_t n, float time) {
    clear_error();
    fh_status st = check_handle(h, "fh_upload_words");
    if (st) return st;
    if (n != h->n) return set_error(FH_ERR_INVALID, "fh_upload_words: n = %u but the flock holds %u", n, h->n);
    if (n && !words) return set_error(FH_ERR_INVALID, "fh_upload_words: NULL words");
    std::vector<float4> pos(n), vel(n), steer(n);
    std::vector<float> spin(n), cen(n);
    std::vector<uint32_t> nb(n);
    for (uint32_t i = 0; i < n; ++i) {
        const float *w = words + (size_t)i * FH_WORDS;
        pos[i] = make_float4(w[PX], w[PY], w[PZ], w[NEXT]);
        vel[i] = make_float4(w[VX], w[VY], w[VZ], w[ROLL]);
        steer[i] = make_float4(w[SX], w[SY], w[SZ], w[RADIUS]);
        spin[i] = w[SPIN];
        std::memcpy(&nb[i], &w[NB], 4);
        cen[i] = w[CEN];
    }
    if (n) {
        FH_HIP(hipMemcpyAsync(h->P.pos4, pos.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream));
        FH_HIP(hipMemcpyAsync(h->P.vel4, vel.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream));
        FH_HIP(hipMemcpyAsync(h->P.steer4, steer.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream));
        FH_HIP(hipMemcpyAsync(h->P.spin, spin.data(), n * 4, hipMemcpyHostToDevice, h->stream));
        FH_HIP(hipMemcpyAsync(h->P.cen, cen.data(), n * 4, hipMemcpyHostToDevice, h->stream));
        FH_HIP(hipMemcpyAsync(h->P.nb, nb.data(), n * 4, hipMemcpyHostToDevice, h->stream));
    }
    h->time = time;
    st = sync_clock(h);
    if (st) return st;
    FH_HIP(hipStreamSynchronize(h->stream));
    if (h->resident) {  // the uploaded state replaces the resident records
        h->resident = false;
        h->counts_dirty = true;
    }
    h->stage = Stage::None;
    return FH_OK;
}

fh_status fh_download_words(fh_flock *h, float *words, float *time) {
    clear_error();
    fh_status st = check_handle(h, "fh_download_words");
    if (st) return st;
    if (h->n && !words) return set_error(FH_ERR_INVALID, "fh_download_words: NULL words");
    if ((st = leave_resident(h))) return st;
    const uint32_t n = h->n;
    std::vector<float4> pos(n), vel(n), steer(n);
    std::vector<float> spin(n), cen(n);
    std::vector<uint32_t> nb(n);
    if (n) {
        FH_HIP(hipMemcpyAsync(pos.data(), h->P.pos4, n * sizeof(float4), hipMemcpyDeviceToHost, h->stream));
        FH_HIP(hipMemcpyAsync(vel.data(), h->P.vel4, n * sizeof(float4), hipMemcpyDeviceToHost, h->stream));
        FH_HIP(hipMemcpyAsync(steer.data(), h->P.steer4, n * sizeof(float4), hipMemcpyDeviceToHost, h->stream));
        FH_HIP(hipMemcpyAsync(spin.data(), h->P.spin, n * 4, hipMemcpyDeviceToHost, h->stream));
        FH_HIP(hipMemcpyAsync(cen.data(), h->P.cen, n * 4, hipMemcpyDeviceToHost, h->stream));
        FH_HIP(hipMemcpyAsync(nb.data(), h->P.nb, n * 4, hipMemcpyDeviceToHost, h->stream));
    }
    FH_HIP(hipStreamSynchronize(h->stream));
    for (uint32_t i = 0; i < n; ++i) {
        float *w = words + (size_t)i * FH_WORDS;
        w[PX] = pos[i].x; w[PY] = pos[i].y; w[PZ] = pos[i].z;
        w[VX] = vel[i].x; w[VY] = vel[i].y; w[VZ] = vel[i].z;
        w[ROLL] = vel[i].w;
        w[SPIN] = spin[i];
        w[RADIUS] = steer[i].w;
        w[SX] = steer[i].x; w[SY] = steer[i].y; w[SZ] = steer[i].z;
        w[NEXT] = pos[i].w;
        std::memcpy(&w[NB], &nb[i], 4);
        w[CEN] = cen[i];
    }
    if (time) *time = h->time;
    return read_error_word(h);
}

fh_status fh_hash_count(fh_flock *h) {
    clear_error();
    fh_status st = check_handle(h, "fh_hash_count");
    if (!st) st = leave_resident(h);
    return st ? st : do_hash(h);
}

fh_status fh_scan(fh_flock *h) {
    clear_error();
    fh_status st = check_handle(h, "fh_scan");
    if (!st) st = leave_resident(h);
    if (st) return st;
    if (h->stage != Stage::Hashed) return set_error(FH_ERR_STATE, "fh_scan: no histogram of the current positions");
    return do_scan(h);
}

fh_status fh_scatter(fh_flock *h) {
    clear_error();
    fh_status st = check_handle(h, "fh_scatter");
    if (!st) st = leave_resident(h);
    if (st) return st;
    if (h->stage != Stage::Scanned) return set_error(FH_ERR_STATE, "fh_scatter: scan the histogram first");
    return do_scatter(h, false);
}

fh_status fh_decide(fh_flock *h, float time) {
    clear_error();
    fh_status st = check_handle(h, "fh_decide");
    if (!st) st = leave_resident(h);
    if (st) return st;
    if (h->stage != Stage::Scattered || !h->absolute_due)
        return set_error(FH_ERR_STATE, "fh_decide: build the grid (hash, scan, scatter) first");
    h->time = time;
    if ((st = sync_clock(h))) return st;
    return do_decide(h, time);
}

fh_status fh_integrate(fh_flock *h) {
    clear_error();
    fh_status st = check_handle(h, "fh_integrate");
    if (!st) st = leave_resident(h);
    return st ? st : do_integrate(h, false);
}

fh_status fh_integrate_hash(fh_flock *h) {
    clear_error();
    fh_status st = check_handle(h, "fh_integrate_hash");
    if (!st) st = leave_resident(h);
    return st ? st : do_integrate(h, true);
}

fh_status fh_sync(fh_flock *h) {
    clear_error();
    fh_status st = check_handle(h, "fh_sync");
    if (st) return st;
    FH_HIP(hipStreamSynchronize(h->stream));
    FH_HIP(hipGetLastError());
    return read_error_word(h);
}

fh_status fh_get_time(const fh_flock *h, float *time) {
    clear_error();
    fh_status st = check_handle(h, "fh_get_time");
    if (st) return st;
    if (!time) return set_error(FH_ERR_INVALID, "fh_get_time: NULL");
    *time = h->time;
    return FH_OK;
}

fh_status fh_step(fh_flock *h, uint32_t steps, uint32_t warmup, fh_pass_ms *out) {
    clear_error();
    fh_status st = check_handle(h, "fh_step");
    if (st) return st;
    if (warmup > steps) warmup = steps;  // cuda/flock.cu:389
    const bool fused = !(h->prm.flags & FH_FLAG_UNFUSED);
    const bool graph = fused && (h->prm.flags & FH_FLAG_GRAPH);
    const bool res = fused && (h->prm.flags & FH_FLAG_RESIDENT);
    if ((st = sync_clock(h))) return st;
    if (res && !h->resident && (st = enter_resident(h))) return st;
    if (res && h->lazy) FH_HIP(launch_set_clock(h->mclock + h->cpar, h->time, h->stream));
    const uint32_t timed = steps - warmup;
    std::vector<hipEvent_t> ev;
    const bool per_pass = out && !graph;
    if (out) {
        ev.resize(per_pass ? (size_t)timed * EV_PER_STEP : 2);
        for (auto &x : ev) FH_HIP(hipEventCreate(&x));
    }
    auto destroy_events = [&] {
        for (auto &x : ev) (void)hipEventDestroy(x);
    };
    std::chrono::steady_clock::time_point t0{}, t1{};
    for (uint32_t s = 0; s < steps; ++s) {
        if (s == warmup) {
            FH_HIP(hipStreamSynchronize(h->stream));
            t0 = std::chrono::steady_clock::now();
            if (out && graph) FH_HIP(hipEventRecord(ev[0], h->stream));
        }
        // A run of GRAPH_RUNS[i] steps as one graph, never across the warm-up
        // boundary (the timed steps start on a step of their own).
        const uint32_t seg_end = s < warmup ? warmup : steps;
        const bool lazy = res && h->lazy;
        int run = -1;
        if (graph && (lazy ? h->phase == 0 : (!res || h->parity == 0)))
            for (int i = 0; i < 2 && run < 0; ++i)
                if (seg_end - s >= GRAPH_RUNS[i]) run = i;
        if (run >= 0) {
            if (!res && h->stage != Stage::Hashed && (st = do_hash(h))) break;
            if ((st = lazy  ? capture_lazy_run(h, run)
                      : res ? capture_resident_run(h, run)
                            : capture_graph_run(h, run)))
                break;
            hipError_t e = hipGraphLaunch(lazy  ? h->lgraph_run[run][h->parity][h->cpar]
                                          : res ? h->rgraph_run[run]
                                                : h->graph_run[run],
                                          h->stream);
            if (e != hipSuccess) {
                st = hip_fail(e, "hipGraphLaunch");
                break;
            }
            for (uint32_t k = 0; k < GRAPH_RUNS[run]; ++k) h->time += DT;  // as one step at a time
            if (lazy) h->parity ^= (GRAPH_RUNS[run] / LAZY_PERIOD) & 1u;  // one rebuild per period
            if (!res) {
                h->stage = Stage::Hashed;
                h->counts_dirty = true;
                h->absolute_due = false;
            }
            s += GRAPH_RUNS[run] - 1;
            continue;
        }
        h->time += DT;  // host f32 clock (cuda/flock.cu:455); the device adds DT the same way
        if (lazy) {
            hipEvent_t *e6 = (per_pass && s >= warmup) ? &ev[(size_t)(s - warmup) * EV_PER_STEP] : nullptr;
            if ((st = lazy_step_launches(h, e6))) break;
        } else if (res) {
            if (graph) {
                if ((st = capture_resident_graph(h))) break;
                hipError_t e = hipGraphLaunch(h->rgraph[h->parity], h->stream);
                if (e != hipSuccess) {
                    st = hip_fail(e, "hipGraphLaunch");
                    break;
                }
            } else {
                hipEvent_t *e6 = (per_pass && s >= warmup) ? &ev[(size_t)(s - warmup) * EV_PER_STEP] : nullptr;
                if ((st = resident_step_launches(h, e6))) break;
            }
            h->parity ^= 1u;
        } else if (graph) {
            if (h->stage != Stage::Hashed) {
                if ((st = do_hash(h))) break;
            }
            if ((st = capture_graph(h))) break;
            hipError_t e = hipGraphLaunch(h->graph, h->stream);
            if (e != hipSuccess) {
                st = hip_fail(e, "hipGraphLaunch");
                break;
            }
            h->stage = Stage::Hashed;
            h->counts_dirty = true;
            h->absolute_due = false;
        } else {
            hipEvent_t *e6 = (per_pass && s >= warmup) ? &ev[(size_t)(s - warmup) * EV_PER_STEP] : nullptr;
            if ((st = step_launches(h, fused, e6))) break;
        }
    }
    if (st) {
        (void)hipStreamSynchronize(h->stream);
        destroy_events();
        return st;
    }
    if (out && graph && timed) FH_HIP(hipEventRecord(ev[1], h->stream));
    FH_HIP(hipStreamSynchronize(h->stream));
    t1 = std::chrono::steady_clock::now();
    if (out) {
        std::memset(out, 0, sizeof *out);
        out->timed_steps = timed;
        if (timed) {
            out->wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
            if (per_pass) {
                // iv[k]: ev[k] -> ev[k+1] for k < 5; iv[5]: decide's kernel; iv[6]: the step
                double iv[7] = {0, 0, 0, 0, 0, 0, 0};
                for (uint32_t s = 0; s < timed; ++s) {
                    hipEvent_t *e = &ev[(size_t)s * EV_PER_STEP];
                    float ms = 0.0f;
                    for (int k = 0; k < 5; ++k) {
                        FH_HIP(hipEventElapsedTime(&ms, e[k], e[k + 1]));
                        iv[k] += ms;
                    }
                    FH_HIP(hipEventElapsedTime(&ms, e[6], e[7]));
                    iv[5] += ms;
                    FH_HIP(hipEventElapsedTime(&ms, e[0], e[5]));
                    iv[6] += ms;
                }
                out->decide = iv[5] / timed;
                if (res) {  // decide, integrate, then the rebuild (scan and move, reported as scatter)
                    out->grid_histogram = 0.0;
                    out->grid_scan = 0.0;
                    out->grid_scatter = iv[2] / timed;
                    out->integrate = iv[1] / timed;
                } else {
                    out->grid_histogram = iv[0] / timed;
                    out->grid_scan = iv[1] / timed;
                    out->grid_scatter = iv[2] / timed;
                    out->integrate = iv[4] / timed;
                }
                out->step_total = iv[6] / timed;
            } else {
                float ms = 0.0f;
                FH_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
                out->step_total = ms / timed;
            }
        }
    }
    destroy_events();
    if (out) {
        uint32_t due = 0;
        const uint32_t last = h->lazy ? h->cpar ^ 1u : h->parity ^ 1u;  // the last step's counter parity
        FH_HIP(hipMemcpy(&due, res ? h->res_due + last : h->P.due_count, 4, hipMemcpyDeviceToHost));
        out->due_last = due;
    }
    return read_error_word(h);
}

fh_status fh_read_grid(fh_flock *h, uint32_t *keys, uint32_t *counts, uint32_t *offsets, uint32_t *sorted) {
    clear_error();
    fh_status st = check_handle(h, "fh_read_grid");
    if (st) return st;
    const uint32_t n = h->n, B = h->buckets;
    FH_HIP(hipStreamSynchronize(h->stream));
    if (h->resident) {
        // The resident layout is the grid of the current positions: keys by
        // bird, offsets, counts and the bird of each slot, read in place
        // (between lazy rebuilds, after the layout is rebuilt).
        if ((st = settle_lazy(h))) return st;
        FH_HIP(hipStreamSynchronize(h->stream));
        const Records &r = h->R[h->parity];
        std::vector<uint32_t> id(n), key(n), off((size_t)B + 1);
        if (n) {
            FH_HIP(hipMemcpy(id.data(), r.id, (size_t)n * 4, hipMemcpyDeviceToHost));
            FH_HIP(hipMemcpy(key.data(), r.key, (size_t)n * 4, hipMemcpyDeviceToHost));
        }
        FH_HIP(hipMemcpy(off.data(), h->off[h->parity], ((size_t)B + 1) * 4, hipMemcpyDeviceToHost));
        for (uint32_t s2 = 0; s2 < n; ++s2) {
            if (id[s2] >= n) return set_error(FH_ERR_HIP, "fh_read_grid: slot %u holds bird %u", s2, id[s2]);
            if (keys) keys[id[s2]] = key[s2];
            if (sorted) sorted[s2] = id[s2];
        }
        if (offsets) std::memcpy(offsets, off.data(), ((size_t)B + 1) * 4);
        if (counts)
            for (uint32_t b = 0; b < B; ++b) counts[b] = off[b + 1] - off[b];
        return read_error_word(h);
    }
    if (h->stage == Stage::None) return set_error(FH_ERR_STATE, "fh_read_grid: no grid built");
    if (keys && n) FH_HIP(hipMemcpy(keys, h->P.key, (size_t)n * 4, hipMemcpyDeviceToHost));
    const bool have_offsets = h->stage != Stage::Hashed;
    std::vector<uint32_t> off;
    if (offsets || (counts && have_offsets)) {
        if (!have_offsets) return set_error(FH_ERR_STATE, "fh_read_grid: offsets need fh_scan first");
        off.resize((size_t)B + 1);
        FH_HIP(hipMemcpy(off.data(), h->P.offsets, ((size_t)B + 1) * 4, hipMemcpyDeviceToHost));
        if (offsets) std::memcpy(offsets, off.data(), ((size_t)B + 1) * 4);
    }
    if (counts) {
        if (have_offsets) {
            for (uint32_t b = 0; b < B; ++b) counts[b] = off[b + 1] - off[b];
        } else {
            FH_HIP(hipMemcpy(counts, h->P.counts, (size_t)B * 4, hipMemcpyDeviceToHost));
        }
    }
    if (sorted) {
        if (h->stage != Stage::Scattered && h->stage != Stage::Decided)
            return set_error(FH_ERR_STATE, "fh_read_grid: sorted needs fh_scatter first");
        std::vector<float4> sp((size_t)n * 2);
        if (n) FH_HIP(hipMemcpy(sp.data(), h->P.sbird, (size_t)n * 2 * sizeof(float4), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; ++i) std::memcpy(&sorted[i], &sp[(size_t)i * 2].w, 4);
    }
    return read_error_word(h);
}

fh_status fh_neighbours(fh_flock *h, uint32_t bird, int walk, uint32_t *ids, uint32_t cap, uint32_t *count) {
    clear_error();
    fh_status st = check_handle(h, "fh_neighbours");
    if (!st) st = leave_resident(h);
    if (st) return st;
    if (bird >= h->n || (walk != 1 && walk != 2) || !count || (cap && !ids))
        return set_error(FH_ERR_INVALID, "fh_neighbours: bad bird/walk/buffer");
    if (h->stage != Stage::Scattered) return set_error(FH_ERR_STATE, "fh_neighbours: build the grid first");
    if (cap + 1 > h->nbr_cap) {
        if (h->d_nbr) FH_HIP(hipFree(h->d_nbr));
        h->d_nbr = nullptr;
        h->nbr_cap = 0;
        FH_HIP(hipMalloc(&h->d_nbr, ((size_t)cap + 1) * 4));
        h->nbr_cap = cap + 1;
    }
    FH_HIP(launch_neighbours(h->geo, h->P, bird, walk, h->d_nbr + 1, cap, h->d_nbr, h->stream));
    FH_HIP(hipStreamSynchronize(h->stream));
    uint32_t c = 0;
    FH_HIP(hipMemcpy(&c, h->d_nbr, 4, hipMemcpyDeviceToHost));
    *count = c;
    const uint32_t k = c < cap ? c : cap;
    if (k) FH_HIP(hipMemcpy(ids, h->d_nbr + 1, (size_t)k * 4, hipMemcpyDeviceToHost));
    std::vector<uint32_t> v(ids, ids + k);
    std::sort(v.begin(), v.end());
    std::memcpy(ids, v.data(), (size_t)k * 4);
    return FH_OK;
}

fh_status fh_kernel_info(const char *kernel, int *vgprs, int *lds_bytes, int *scratch_bytes) {
    clear_error();
    if (!kernel) return set_error(FH_ERR_INVALID, "fh_kernel_info: NULL name");
    fh_status st = require_device();
    if (st) return st;
    hipFuncAttributes at;
    hipError_t e = grid_kernel_attributes(kernel, &at);
    if (e == hipErrorInvalidValue) e = decide_kernel_attributes(kernel, &at);
    if (e == hipErrorInvalidValue) e = integrate_kernel_attributes(kernel, &at);
    if (e == hipErrorInvalidValue) return set_error(FH_ERR_INVALID, "fh_kernel_info: unknown kernel %s", kernel);
    if (e != hipSuccess) return hip_fail(e, "hipFuncGetAttributes");
    if (vgprs) *vgprs = at.numRegs;
    if (lds_bytes) *lds_bytes = (int)at.sharedSizeBytes;
    if (scratch_bytes) *scratch_bytes = (int)at.localSizeBytes;
    return FH_OK;
}

}  // extern "C"
