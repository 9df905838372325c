// fh_step_device_noalloc: the flock step for callers that cannot allocate
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc). The caller's device buffer of .flk records is stepped in place;
// every work array of the step (the bird-order state, the grid, the sorted
// copies, counters and the flock clock) lives in one file-scope device array,
// `scratch`, of NOALLOC_SCRATCH_BYTES, sized for FH_NOALLOC_MAX_BIRDS birds.
// Each GPU holds its own copy of the array until the process ends; its
// address on a device is looked up once (hipGetSymbolAddress, with that
// device current) and kept. Every call on a GPU reuses that GPU's array, so
// calls that overlap on one GPU (two streams, two host threads) race.
//
// The launches are those of fh_step's fused standard path in canonical order
// (cuda/flock.cu:452-470 with the histogram fused into the previous step's
// integrate), so the result equals fh_upload_words + fh_step + fh_download_words
// on a handle created with the same tunables and FH_FLAG_CANONICAL, bit for
// bit. No hipMalloc, hipFree, hipMemset or host-device copy is made: the
// state is converted on the device, and what must start at zero is cleared
// by fill kernels.
#include <mutex>

#include "host_util.h"
#include "layout.h"
#include "model.h"

namespace fh {
hipError_t launch_fill(uint32_t *dst, uint32_t count, uint32_t value, hipStream_t s);

constexpr uint32_t NOALLOC_MAX_BUCKETS = 1u << 17;  // buckets_for(FH_NOALLOC_MAX_BIRDS)
constexpr size_t NOALLOC_SCRATCH_BYTES = 8u << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

}  // namespace fh

static __device__ __align__(256) unsigned char scratch[fh::NOALLOC_SCRATCH_BYTES];

namespace fh {
namespace {

// Offsets of the work arrays in the scratch array for n birds and nb
// (padded) buckets, in the order and alignment of fh_create's arena.
struct Layout {
    size_t pos, vel, steer, sbird, spin, cen, nb, key, rank, snext, tmp, counts, offsets, tsum, tpre, err, due,
        seg, bigc, big, occ, used;
};

Layout carve(uint32_t n, uint32_t nb) {
    Arena a;
    Layout l;
    const size_t nn = n ? n : 1;
    const uint32_t ntiles = nb / SCAN_TILE;
    l.pos = a.take<float4>(nn);
    l.vel = a.take<float4>(nn);
    l.steer = a.take<float4>(nn);
    l.sbird = a.take<float4>(2 * nn);
    l.spin = a.take<float>(nn);
    l.cen = a.take<float>(nn);
    l.nb = a.take<uint32_t>(nn);
    l.key = a.take<uint32_t>(nn);
    l.rank = a.take<uint32_t>(nn);
    l.snext = a.take<float>(nn);
    l.tmp = a.take<uint32_t>(nn);
    l.counts = a.take<uint32_t>(nb);
    l.offsets = a.take<uint32_t>((size_t)nb + 1);
    l.tsum = a.take<uint32_t>(ntiles);
    l.tpre = a.take<uint32_t>(ntiles);
    l.err = a.take<uint32_t>(4);
    l.due = a.take<uint32_t>(4);
    l.seg = a.take<Seg>(1);
    l.bigc = a.take<uint32_t>(4);
    l.big = a.take<uint32_t>(big_list_cap((uint32_t)nn));
    l.occ = a.take<uint32_t>(nb / 32);
    l.used = a.used;
    return l;
}

// .flk records -> bird-order state (the conversion fh_upload_words does on the host).
__global__ __launch_bounds__(256) void k_words_in(const float *__restrict__ words, uint32_t n, Pools p) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float *w = words + (size_t)i * FH_WORDS;
    p.pos4[i] = make_float4(w[PX], w[PY], w[PZ], w[NEXT]);
    p.vel4[i] = make_float4(w[VX], w[VY], w[VZ], w[ROLL]);
    p.steer4[i] = make_float4(w[SX], w[SY], w[SZ], w[RADIUS]);
    p.spin[i] = w[SPIN];
    p.nb[i] = __float_as_uint(w[NB]);
    p.cen[i] = w[CEN];
}

// Bird-order state -> .flk records (fh_download_words).
__global__ __launch_bounds__(256) void k_words_out(float *__restrict__ words, uint32_t n, Pools p) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    float *w = words + (size_t)i * FH_WORDS;
    const float4 q = p.pos4[i], v = p.vel4[i], s = p.steer4[i];
    w[PX] = q.x;
    w[PY] = q.y;
    w[PZ] = q.z;
    w[VX] = v.x;
    w[VY] = v.y;
    w[VZ] = v.z;
    w[ROLL] = v.w;
    w[SPIN] = p.spin[i];
    w[RADIUS] = s.w;
    w[SX] = s.x;
    w[SY] = s.y;
    w[SZ] = s.z;
    w[NEXT] = q.w;
    w[NB] = __uint_as_float(p.nb[i]);
    w[CEN] = p.cen[i];
}

// The flock's segment record, written by a kernel (no host copy).
__global__ void k_seg_init(Seg *dst, Seg v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *dst = v;
}

std::mutex g_mu;
void *g_ptr[NOALLOC_MAX_DEVICES];

// The scratch array's address on the current device (looked up once per
// device, with that device current).
fh_status scratch_ptr(void **out) {
    int dev = 0;
    FH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return set_error(FH_ERR_INVALID, "fh_step_device_noalloc: device %d beyond %d", dev, NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return FH_OK;
}

}  // namespace
}  // namespace fh

using namespace fh;

extern "C" fh_status fh_step_device_noalloc(float *d_words, uint32_t n, float time, uint32_t steps,
                                            const fh_params *params, void *stream, float *time_out) {
    clear_error();
    static_assert(sizeof(Seg) <= 256, "one arena slot");
    if (n > FH_NOALLOC_MAX_BIRDS)
        return set_error(FH_ERR_INVALID, "fh_step_device_noalloc: %u birds exceed the maximum of %u", n,
                         (uint32_t)FH_NOALLOC_MAX_BIRDS);
    if (n && !d_words) return set_error(FH_ERR_INVALID, "fh_step_device_noalloc: NULL words");
    fh_params prm;
    fh_status st = resolve_params(params, &prm, "fh_step_device_noalloc");
    if (st) return st;
    constexpr uint32_t allowed = FH_FLAG_CANONICAL | FH_FLAG_WAVE_AGG | FH_FLAG_NO_TILE | FH_FLAG_OCC |
                                 FH_FLAG_SPATIAL | FH_FLAG_XCD | FH_FLAG_WINDOW;
    if (prm.flags & ~allowed)
        return set_error(FH_ERR_INVALID,
                         "fh_step_device_noalloc: flags 0x%x not supported (resident, graph, persistent, fused "
                         "flight and unfused layouts need a handle)",
                         prm.flags & ~allowed);
    prm.flags |= FH_FLAG_CANONICAL;
    const uint32_t buckets = flock_buckets(n, prm);
    if (buckets > NOALLOC_MAX_BUCKETS)
        return set_error(FH_ERR_INVALID, "fh_step_device_noalloc: %u buckets exceed the maximum of %u", buckets,
                         NOALLOC_MAX_BUCKETS);
    if ((st = require_device())) return st;
    const uint32_t nb = (buckets + SCAN_TILE - 1) / SCAN_TILE * SCAN_TILE;
    const Layout l = carve(n, nb);
    if (l.used > NOALLOC_SCRATCH_BYTES)
        return set_error(FH_ERR_INVALID, "fh_step_device_noalloc: %zu scratch bytes needed, %zu held", l.used,
                         NOALLOC_SCRATCH_BYTES);
    void *base_v = nullptr;
    if ((st = scratch_ptr(&base_v))) return st;
    char *base = static_cast<char *>(base_v);
    hipStream_t s = static_cast<hipStream_t>(stream);

    Pools P{};
    P.pos4 = reinterpret_cast<float4 *>(base + l.pos);
    P.vel4 = reinterpret_cast<float4 *>(base + l.vel);
    P.steer4 = reinterpret_cast<float4 *>(base + l.steer);
    P.sbird = reinterpret_cast<float4 *>(base + l.sbird);
    P.spin = reinterpret_cast<float *>(base + l.spin);
    P.cen = reinterpret_cast<float *>(base + l.cen);
    P.nb = reinterpret_cast<uint32_t *>(base + l.nb);
    P.key = reinterpret_cast<uint32_t *>(base + l.key);
    P.rank = reinterpret_cast<uint32_t *>(base + l.rank);
    P.snext = reinterpret_cast<float *>(base + l.snext);
    P.sorted_tmp = reinterpret_cast<uint32_t *>(base + l.tmp);
    P.counts = reinterpret_cast<uint32_t *>(base + l.counts);
    P.offsets = reinterpret_cast<uint32_t *>(base + l.offsets);
    P.tile_sums = reinterpret_cast<uint32_t *>(base + l.tsum);
    P.tile_prefix = reinterpret_cast<uint32_t *>(base + l.tpre);
    P.err = reinterpret_cast<uint32_t *>(base + l.err);
    P.due_count = reinterpret_cast<uint32_t *>(base + l.due);
    P.work = P.due_count + 1;
    P.big_count = reinterpret_cast<uint32_t *>(base + l.bigc);
    P.big = reinterpret_cast<uint32_t *>(base + l.big);
    P.big_cap = big_list_cap(n ? n : 1);
    P.occ = reinterpret_cast<uint32_t *>(base + l.occ);
    Seg *d_seg = reinterpret_cast<Seg *>(base + l.seg);

    Seg seg{};
    seg.bird_base = 0;
    seg.n = n;
    seg.bucket_base = 0;
    seg.mask = buckets - 1;
    seg.cell = prm.cell;
    seg.inv_cell = 1.0f / prm.cell;
    seg.time = time;
    seg.live = 1;
    seg.mode = (prm.flags & FH_FLAG_SPATIAL) ? SEG_SPATIAL : 0u;
    Geometry g{};
    g.segs = d_seg;
    g.tile_seg = nullptr;
    g.n_birds = n;
    g.n_slots = n;
    g.n_buckets = buckets;

    // The last call's contents stay in the array: the counts, the error word
    // and the step counters start at zero here.
    hipLaunchKernelGGL(k_seg_init, dim3(1), dim3(64), 0, s, d_seg, seg);
    FH_HIP(hipGetLastError());
    FH_HIP(launch_fill(P.counts, nb, 0u, s));
    FH_HIP(launch_fill(P.err, 4u, 0u, s));
    FH_HIP(launch_fill(P.due_count, 4u, 0u, s));
    FH_HIP(launch_fill(P.big_count, 4u, 0u, s));
    if (n) {
        hipLaunchKernelGGL(k_words_in, dim3((n + 255u) / 256u), dim3(256), 0, s, d_words, n, P);
        FH_HIP(hipGetLastError());
    }
    const bool wave_agg = (prm.flags & FH_FLAG_WAVE_AGG) != 0;
    float t = time;
    for (uint32_t k = 0; k < steps; ++k) {
        // as fh_step's fused standard path: the first step hashes, later ones
        // take the histogram the previous integrate left
        if (k == 0) FH_HIP(launch_hash_count(g, P, wave_agg, s));
        FH_HIP(launch_scan(g, P, d_seg, 1u, DT, s));  // the scan advances the device clock by DT
        FH_HIP(launch_scatter(g, P, true, true, s));
        FH_HIP(launch_decide(g, P, 0.0f, prm.decide_block, prm.birds_per_thread, prm.decide_group,
                             walk_flags(prm.flags), s));
        FH_HIP(launch_integrate(g, P, true, wave_agg, s));
        t = t + DT;  // the host clock, the same f32 additions (cuda/flock.cu:455)
    }
    if (n) {
        hipLaunchKernelGGL(k_words_out, dim3((n + 255u) / 256u), dim3(256), 0, s, d_words, n, P);
        FH_HIP(hipGetLastError());
    }
    if (time_out) *time_out = t;
    return FH_OK;
}
