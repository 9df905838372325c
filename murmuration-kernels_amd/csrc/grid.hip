// Spatial-hash counting sort for gfx950: cell keys + histogram, a three-kernel
// reduce/scan/apply exclusive scan built on a wave64 DPP scan, and the scatter
// into the sorted position/velocity buffers that decide walks.
//
// Reference passes: grid_histogram (cuda/flock.cu:118-124), grid_scan_blocks /
// totals / add (:144-172), grid_scatter (:175-179). Keys, counts and offsets
// are the reference's exactly; the scatter keeps the reference's arrival order
// within a bucket unless the canonical pass reorders each bucket by bird index.
#include "flight.h"
#include "layout.h"
#include "model.h"

namespace fh {

// ---- wave64 / workgroup scans --------------------------------------------

// Inclusive prefix sum across the 64 lanes of a wave with DPP row shifts and
// row broadcasts (rows of 16 lanes). All 64 lanes must be active.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

// Exclusive prefix of v over a workgroup of NWAVES waves; *total = sum.
template <int NWAVES>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t wave_tot[NWAVES];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t incl = wave_incl_scan(v);
    if (lane == 63u) wave_tot[wave] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < NWAVES; ++w) {
        uint32_t t = wave_tot[w];
        before += (uint32_t)w < wave ? t : 0u;
        all += t;
    }
    __syncthreads();  // wave_tot may be reused by the caller's next scan
    *total = all;
    return before + incl - v;
}

__device__ __forceinline__ const Seg &tile_segment(const Geometry &g, uint32_t tile, bool *skip) {
    uint32_t si = g.tile_seg ? g.tile_seg[tile] : 0u;
    *skip = si == NO_SEG;
    return g.segs[*skip ? 0u : si];
}

__device__ __forceinline__ uint32_t key_of(float x, float y, float z, const Seg &sg) {
    return bucket_of_cell(sg, cell_coord(x, sg.inv_cell), cell_coord(y, sg.inv_cell), cell_coord(z, sg.inv_cell));
}

// counts[key]++ returning the bird's arrival rank in its bucket. With
// WAVE_AGG the lanes that share a key add once per wave (leader lane) and
// take consecutive ranks; otherwise every lane issues its own atomic.
template <bool WAVE_AGG>
__device__ __forceinline__ uint32_t count_key(uint32_t *counts, uint32_t k, bool active) {
    if (!WAVE_AGG) return active ? atomicAdd(&counts[k], 1u) : 0u;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;
    uint64_t left = __ballot(active);
    uint32_t rank = 0;
    while (left) {
        const int leader = __ffsll((long long)left) - 1;
        const uint32_t lk = __shfl(k, leader);
        const uint64_t grp = __ballot(active && k == lk) & left;
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(&counts[lk], (uint32_t)__popcll(grp));
        base = __shfl(base, leader);
        if ((grp >> lane) & 1ull) rank = base + (uint32_t)__popcll(grp & below);
        left &= ~grp;
    }
    return rank;
}

// grid_histogram: key = bucket of the bird's cell; rank = arrival order.
template <bool WAVE_AGG>
__global__ __launch_bounds__(TILE) void k_hash_count(Geometry g, Pools p) {
    bool skip;
    const uint32_t tile = g.tile_begin + blockIdx.x;
    const Seg &sg = tile_segment(g, tile, &skip);
    if (skip) return;
    const uint32_t bird = tile * TILE + threadIdx.x;
    const bool active = bird < g.n_birds && bird - sg.bird_base < sg.n;
    uint32_t k = 0;
    if (active) {
        const float4 q = p.pos4[bird];
        k = key_of(q.x, q.y, q.z, sg);
        p.key[bird] = k;
    }
    const uint32_t r = count_key<WAVE_AGG>(p.counts, k, active);
    if (active) p.rank[bird] = r;
}

// Scan, level 1: the sum of each SCAN_TILE-bucket tile. With nseg > 0,
// workgroup 0 also advances the clock of every live segment by dt (the
// step's time += DT, which the scatter after this scan reads).
__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t *__restrict__ counts,
                                                     uint32_t *__restrict__ tile_sums, Seg *segs, uint32_t nseg,
                                                     float dt) {
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < nseg; i += 256u)
            if (segs[i].live) segs[i].time = segs[i].time + dt;
    const uint4 *c4 = reinterpret_cast<const uint4 *>(counts + (size_t)blockIdx.x * SCAN_TILE);
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint4 v = c4[k * 256 + threadIdx.x];
        sum += v.x + v.y + v.z + v.w;
    }
    uint32_t total;
    block_excl_scan<4>(sum, &total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Scan, level 2: one workgroup scans the tile sums; offsets[end] = grand total.
__global__ __launch_bounds__(1024) void k_scan_tiles(const uint32_t *__restrict__ tile_sums,
                                                     uint32_t *__restrict__ tile_prefix, uint32_t ntiles,
                                                     uint32_t *__restrict__ offsets_end,
                                                     uint32_t *__restrict__ due_count,
                                                     uint32_t *__restrict__ big_count, uint32_t *__restrict__ work) {
    if (threadIdx.x == 0) {  // a new grid starts a new step
        *due_count = 0;
        *big_count = 0;
        *work = 0;
    }
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += 4096u) {
        uint32_t v[4], sum = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint32_t t = base + threadIdx.x * 4u + (uint32_t)k;
            v[k] = t < ntiles ? tile_sums[t] : 0u;
            sum += v[k];
        }
        uint32_t total;
        uint32_t run = carry + block_excl_scan<16>(sum, &total);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint32_t t = base + threadIdx.x * 4u + (uint32_t)k;
            if (t < ntiles) tile_prefix[t] = run;
            run += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0) *offsets_end = carry;
}

// Scan, level 3: offsets of one tile = its prefix + local exclusive scan.
// The counts are zeroed for the next histogram as they are consumed, and the
// occupancy bitmap records which buckets hold a bird. With MERGED (at most
// SCAN_MERGE_TILES tiles) there is no level 2: each workgroup adds up the
// sums of the tiles before it itself (tile_in = tile sums; the last one
// writes the grand total and workgroup 0 zeroes the step counters);
// otherwise tile_in holds the prefixes of level 2.
constexpr uint32_t SCAN_MERGE_TILES = 2048;
template <bool MERGED>
__global__ __launch_bounds__(256) void k_scan_apply(uint32_t *__restrict__ counts,
                                                    const uint32_t *__restrict__ tile_in, uint32_t ntiles,
                                                    uint32_t *__restrict__ offsets,
                                                    uint32_t *__restrict__ occ, uint32_t *__restrict__ due_count,
                                                    uint32_t *__restrict__ big_count, uint32_t *__restrict__ work,
                                                    int keep_counts) {
    uint32_t prefix;
    if (MERGED) {
        uint32_t part = 0;
        for (uint32_t t = threadIdx.x; t < blockIdx.x; t += 256u) part += tile_in[t];
        block_excl_scan<4>(part, &prefix);
        if (blockIdx.x == 0 && threadIdx.x == 0) {  // a new grid starts a new step
            *due_count = 0;
            *big_count = 0;
            *work = 0;
        }
    } else {
        prefix = tile_in[blockIdx.x];
    }
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * 16u;
    uint4 *c4 = reinterpret_cast<uint4 *>(counts + base);
    uint32_t c[16];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint4 v = c4[k];
        c[4 * k] = v.x; c[4 * k + 1] = v.y; c[4 * k + 2] = v.z; c[4 * k + 3] = v.w;
    }
    uint32_t sum = 0, bits = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        sum += c[k];
        bits |= (c[k] != 0u ? 1u : 0u) << k;
    }
    // lanes 2i and 2i+1 hold the 32 buckets of one bitmap word
    const uint32_t hi = __shfl_xor(bits, 1);
    if ((threadIdx.x & 1u) == 0u) occ[base / 32u] = bits | (hi << 16);
    uint32_t total;
    uint32_t run = prefix + block_excl_scan<4>(sum, &total);
    if (MERGED && blockIdx.x == ntiles - 1u && threadIdx.x == 0) offsets[(size_t)ntiles * SCAN_TILE] = prefix + total;
    uint32_t o[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        o[k] = run;
        run += c[k];
    }
    uint4 *o4 = reinterpret_cast<uint4 *>(offsets + base);  // base is a multiple of 16
#pragma unroll
    for (int k = 0; k < 4; ++k) o4[k] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
    if (!keep_counts) {
        const uint4 z = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
        for (int k = 0; k < 4; ++k) c4[k] = z;
    }
}

// grid_scatter: slot = offsets[key] + rank. Writes the sorted copies directly,
// or (canonical mode) only the slot -> bird map for the reordering pass.
__global__ __launch_bounds__(TILE) void k_scatter(Geometry g, Pools p, int canonical, int relative) {
    bool skip;
    const uint32_t tile = g.tile_begin + blockIdx.x;
    const Seg &sg = tile_segment(g, tile, &skip);
    if (skip) return;
    const uint32_t bird = tile * TILE + threadIdx.x;
    if (bird >= g.n_birds || bird - sg.bird_base >= sg.n) return;
    const uint32_t s = p.offsets[p.key[bird]] + p.rank[bird];
    if (s >= g.n_slots) {
        atomicOr(p.err, ERR_SLOT);
        return;
    }
    if (canonical) {
        p.sorted_tmp[s] = bird;
        return;
    }
    const float4 q = p.pos4[bird];
    const float4 v = p.vel4[bird];
    const f3 h = hit_heading(mk3(v.x, v.y, v.z));
    p.sbird[2u * s] = make_float4(q.x, q.y, q.z, __uint_as_float(bird));
    p.sbird[2u * s + 1u] = make_float4(h.x, h.y, h.z, p.steer4[bird].w);
    p.snext[s] = relative ? q.w - sg.time : q.w;
}

// The sorted copies of `bird` at slot c (bird < g.n_birds, c < g.n_slots).
__device__ __forceinline__ void write_sorted(const Geometry &g, const Pools &p, uint32_t c, uint32_t bird,
                                             int relative) {
    const float4 q = p.pos4[bird];
    const float4 v = p.vel4[bird];
    float t = 0.0f;
    if (relative) t = g.segs[g.tile_seg ? g.tile_seg[bird / TILE] : 0u].time;
    const f3 h = hit_heading(mk3(v.x, v.y, v.z));
    p.sbird[2u * c] = make_float4(q.x, q.y, q.z, __uint_as_float(bird));
    p.sbird[2u * c + 1u] = make_float4(h.x, h.y, h.z, p.steer4[bird].w);
    p.snext[c] = relative ? q.w - t : q.w;
}

// Canonical order: each bucket lists its birds in ascending bird index (the
// order of a stable counting sort over bird index). A bird of a bucket with
// at most CANON_CAP birds moves to bucket start + number of birds of its
// bucket with a smaller index. Larger buckets are put on the big-bucket list
// (by their first slot) and left to k_canonical_big.
__global__ __launch_bounds__(256) void k_canonical(Geometry g, Pools p, int relative) {
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= g.n_slots) return;
    const uint32_t bird = p.sorted_tmp[s];
    if (bird >= g.n_birds) {
        atomicOr(p.err, ERR_SLOT);
        return;
    }
    const uint32_t k = p.key[bird];
    const uint2 r = bucket_range(p.offsets, k);
    const uint32_t o0 = r.x, o1 = min(r.y, g.n_slots);
    if (!(o0 <= s && s < o1)) {  // inconsistent grid: keep the slot
        atomicOr(p.err, ERR_SLOT);
        write_sorted(g, p, s, bird, relative);
        return;
    }
    if (o1 - o0 > CANON_CAP) {
        if (s == o0) {
            // Listed buckets are disjoint slot ranges of more than CANON_CAP
            // slots each, so big_cap entries always suffice.
            const uint32_t i = atomicAdd(p.big_count, 1u);
            if (i < p.big_cap) p.big[i] = k;
            else atomicOr(p.err, ERR_SLOT);
        }
        return;
    }
    uint32_t c = o0;
    for (uint32_t t = o0; t < o1; ++t) c += p.sorted_tmp[t] < bird;
    write_sorted(g, p, c, bird, relative);
}

// Canonical order of the big buckets: one workgroup per listed bucket (a
// grid-stride loop over the list) loads the bucket's bird indices into LDS,
// bitonic-sorts them and writes the sorted copies in that order. Buckets of
// more than CANON_BIG_CAP birds keep the arrival order and raise ERR_CANON.
constexpr uint32_t BIG_BLOCK = 1024;
__global__ __launch_bounds__(BIG_BLOCK) void k_canonical_big(Geometry g, Pools p, int relative) {
    __shared__ uint32_t ids[CANON_BIG_CAP];
    const uint32_t nbig = min(*p.big_count, p.big_cap);
    for (uint32_t i = blockIdx.x; i < nbig; i += gridDim.x) {
        const uint32_t k = p.big[i];
        const uint2 r = bucket_range(p.offsets, k);
        const uint32_t o0 = r.x, o1 = min(r.y, g.n_slots);
        const uint32_t m = o1 > o0 ? o1 - o0 : 0u;
        if (m > CANON_BIG_CAP) {
            if (threadIdx.x == 0) atomicOr(p.err, ERR_CANON);
            for (uint32_t j = threadIdx.x; j < m; j += BIG_BLOCK) {
                const uint32_t bird = p.sorted_tmp[o0 + j];
                if (bird < g.n_birds) write_sorted(g, p, o0 + j, bird, relative);
                else atomicOr(p.err, ERR_SLOT);
            }
            continue;
        }
        uint32_t size = 2u;
        while (size < m) size <<= 1;
        for (uint32_t j = threadIdx.x; j < size; j += BIG_BLOCK) ids[j] = j < m ? p.sorted_tmp[o0 + j] : 0xffffffffu;
        __syncthreads();
        for (uint32_t span = 2u; span <= size; span <<= 1) {
            for (uint32_t stride = span >> 1; stride > 0u; stride >>= 1) {
                // pair j = (lo, lo + stride) where lo is j with a zero bit inserted at log2(stride)
                for (uint32_t j = threadIdx.x; j < (size >> 1); j += BIG_BLOCK) {
                    const uint32_t lo = 2u * j - (j & (stride - 1u));
                    const uint32_t a = ids[lo], b = ids[lo + stride];
                    const bool ascending = (lo & span) == 0u;
                    if ((a > b) == ascending) {
                        ids[lo] = b;
                        ids[lo + stride] = a;
                    }
                }
                __syncthreads();
            }
        }
        for (uint32_t j = threadIdx.x; j < m; j += BIG_BLOCK) {
            const uint32_t bird = ids[j];
            if (bird < g.n_birds) write_sorted(g, p, o0 + j, bird, relative);
            else atomicOr(p.err, ERR_SLOT);
        }
        __syncthreads();  // ids is refilled for the next bucket
    }
}

// Fill kernel (used instead of hipMemsetAsync so a captured graph holds only
// kernel nodes).
__global__ __launch_bounds__(256) void k_fill(uint32_t *dst, uint32_t count, uint32_t value) {
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < count; i += gridDim.x * 256u) dst[i] = value;
}

hipError_t launch_fill(uint32_t *dst, uint32_t count, uint32_t value, hipStream_t s) {
    if (count == 0) return hipSuccess;
    const uint32_t blocks = min((count + 255u) / 256u, 8192u);
    hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, s, dst, count, value);
    return hipGetLastError();
}

__global__ void k_set_time(Seg *segs, uint32_t nseg, float time, float dt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseg) return;
    if (dt != 0.0f) {
        if (segs[i].live) segs[i].time = segs[i].time + dt;
    } else {
        segs[i].time = time;
    }
}

__global__ void k_set_clock(float *clock, float time) {
    if (threadIdx.x == 0) *clock = time;
}

hipError_t launch_set_clock(float *clock, float time, hipStream_t s) {
    hipLaunchKernelGGL(k_set_clock, dim3(1), dim3(64), 0, s, clock, time);
    return hipGetLastError();
}

static inline uint32_t launch_tiles(const Geometry &g) {
    return g.tile_count ? g.tile_count : (g.n_birds + TILE - 1) / TILE;
}

hipError_t launch_hash_count(const Geometry &g, const Pools &p, bool wave_agg, hipStream_t s) {
    if (g.n_birds == 0) return hipSuccess;
    if (wave_agg)
        hipLaunchKernelGGL(k_hash_count<true>, dim3(launch_tiles(g)), dim3(TILE), 0, s, g, p);
    else
        hipLaunchKernelGGL(k_hash_count<false>, dim3(launch_tiles(g)), dim3(TILE), 0, s, g, p);
    return hipGetLastError();
}

hipError_t launch_scan(const Geometry &g, const Pools &p, Seg *tick_segs, uint32_t tick_nseg, float dt,
                       hipStream_t s, bool keep_counts) {
    const int keep = keep_counts ? 1 : 0;
    const uint32_t ntiles = (g.n_buckets + SCAN_TILE - 1) / SCAN_TILE;
    if (ntiles == 0) {
        if (tick_nseg) return launch_set_time(tick_segs, tick_nseg, 0.0f, dt, s);
        return hipSuccess;
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3(ntiles), dim3(256), 0, s, p.counts, p.tile_sums, tick_segs, tick_nseg, dt);
    if (ntiles <= SCAN_MERGE_TILES) {
        hipLaunchKernelGGL(k_scan_apply<true>, dim3(ntiles), dim3(256), 0, s, p.counts, p.tile_sums, ntiles,
                           p.offsets, p.occ, p.due_count, p.big_count, p.work, keep);
    } else {
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, s, p.tile_sums, p.tile_prefix, ntiles,
                           p.offsets + (size_t)ntiles * SCAN_TILE, p.due_count, p.big_count, p.work);
        hipLaunchKernelGGL(k_scan_apply<false>, dim3(ntiles), dim3(256), 0, s, p.counts, p.tile_prefix, ntiles,
                           p.offsets, p.occ, p.due_count, p.big_count, p.work, keep);
    }
    return hipGetLastError();
}

hipError_t launch_scatter(const Geometry &g, const Pools &p, bool canonical, bool relative_time,
                          hipStream_t s) {
    if (g.n_birds == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter, dim3(launch_tiles(g)), dim3(TILE), 0, s, g, p, canonical ? 1 : 0,
                       relative_time ? 1 : 0);
    if (canonical && g.n_slots) {
        hipLaunchKernelGGL(k_canonical, dim3((g.n_slots + 255u) / 256u), dim3(256), 0, s, g, p,
                           relative_time ? 1 : 0);
        // the list length is known only on the device: a fixed grid walks it
        const uint32_t blocks = min(big_list_cap(g.n_slots), 256u);
        hipLaunchKernelGGL(k_canonical_big, dim3(blocks), dim3(BIG_BLOCK), 0, s, g, p, relative_time ? 1 : 0);
    }
    return hipGetLastError();
}

// ---- resident layout (one flock) ------------------------------------------

// Records from the bird-order state: slot = offsets[key] + rank, the same slot
// the scatter gives the bird.
__global__ __launch_bounds__(TILE) void k_r_enter(Geometry g, Pools p, Records rec, uint32_t bird_lo,
                                                  uint32_t bird_hi) {
    const uint32_t b = bird_lo + blockIdx.x * TILE + threadIdx.x;
    if (b >= bird_hi || b >= g.n_birds) return;
    if (g.tile_seg) {  // a batch: padding birds and dead flocks hold no slot
        const uint32_t si = g.tile_seg[b / TILE];
        if (si == NO_SEG || !g.segs[si].live || b - g.segs[si].bird_base >= g.segs[si].n) return;
    }
    const uint32_t k = p.key[b];
    const uint32_t s = p.offsets[k] + p.rank[b];
    if (s >= g.n_slots) {
        atomicOr(p.err, ERR_SLOT);
        return;
    }
    rec.p4[s] = p.pos4[b];
    rec.v4[s] = p.vel4[b];
    rec.s4[s] = p.steer4[b];
    rec.spin[s] = p.spin[b];
    rec.id[s] = b;
    rec.key[s] = k;
}

// Records back to the bird-order state.
__global__ __launch_bounds__(TILE) void k_r_flush(Geometry g, Pools p, Records rec, uint32_t slot_lo,
                                                  uint32_t slot_hi) {
    const uint32_t s = slot_lo + blockIdx.x * TILE + threadIdx.x;
    if (s >= slot_hi || s >= g.n_slots) return;
    const uint32_t b = rec.id[s];
    if (b >= g.n_birds) {
        atomicOr(p.err, ERR_SLOT);
        return;
    }
    p.pos4[b] = rec.p4[s];
    p.vel4[b] = rec.v4[s];
    p.steer4[b] = rec.s4[s];
    p.spin[b] = rec.spin[s];
    p.key[b] = rec.key[s];
}

// Flocks joining or leaving a resident batch: each record keeps its order and
// moves by its flock's shift; a leaving flock's records are dropped.
__global__ __launch_bounds__(TILE) void k_r_splice(Geometry g, Pools p, Records cur, Records nxt,
                                                   const int32_t *__restrict__ shift, uint32_t n_cur,
                                                   uint32_t n_nxt) {
    const uint32_t s = blockIdx.x * TILE + threadIdx.x;
    if (s >= n_cur) return;
    const uint32_t b = cur.id[s];
    const uint32_t si = b >= g.n_birds ? NO_SEG : g.tile_seg ? g.tile_seg[b / TILE] : 0u;
    if (si == NO_SEG) {
        atomicOr(p.err, ERR_SLOT);
        return;
    }
    const int32_t d = shift[si];
    if (d == SPLICE_DROP) return;
    const uint32_t ns = (uint32_t)((int64_t)s + d);
    if (ns >= n_nxt) {
        atomicOr(p.err, ERR_SLOT);
        return;
    }
    nxt.p4[ns] = cur.p4[s];
    nxt.v4[ns] = cur.v4[s];
    nxt.s4[ns] = cur.s4[s];
    nxt.spin[ns] = cur.spin[s];
    nxt.id[ns] = b;
    nxt.key[ns] = cur.key[s];
}

// Each resident scan tile's count total (RES_TILE buckets per workgroup).
__global__ __launch_bounds__(256) void k_r_tile_sums(const uint32_t *__restrict__ counts, uint32_t *__restrict__ tsum) {
    const uint4 c = reinterpret_cast<const uint4 *>(counts + (size_t)blockIdx.x * RES_TILE)[threadIdx.x];
    uint32_t total;
    block_excl_scan<4>(c.x + c.y + c.z + c.w, &total);
    if (threadIdx.x == 0) tsum[blockIdx.x] = total;
}

// Totals of RES_SUPER consecutive resident tiles (one workgroup each).
__global__ __launch_bounds__(256) void k_r_super_sums(const uint32_t *__restrict__ tsum, uint32_t ntiles,
                                                      uint32_t *__restrict__ tsum2) {
    static_assert(RES_SUPER == 256, "one tile total per lane");
    const uint32_t t = blockIdx.x * RES_SUPER + threadIdx.x;
    uint32_t total;
    block_excl_scan<4>(t < ntiles ? tsum[t] : 0u, &total);
    if (threadIdx.x == 0) tsum2[blockIdx.x] = total;
}

// Rebuild scan (RES_TILE buckets per workgroup, 4 per lane): the new offsets
// from the counts the flight kept current, and from the totals of the tiles
// before this tile (summed directly, or with tsum2 through the super tiles'
// totals); the first arrival slot of each bucket (its stayers come first);
// the occupancy bits. Arrivals are zeroed for the next step. The grid also
// commits the clocks and budgets of the `tick` segments (nothing in this
// launch reads them).
__global__ __launch_bounds__(256) void k_r_scan_apply(const uint32_t *__restrict__ counts,
                                                      uint32_t *__restrict__ arr,
                                                      const uint32_t *__restrict__ tsum,
                                                      const uint32_t *__restrict__ tsum2, uint32_t ntiles,
                                                      uint32_t *__restrict__ noff, uint32_t *__restrict__ abase,
                                                      uint32_t *__restrict__ occ, Seg *tick, uint32_t ntick,
                                                      uint32_t tick_steps) {
    commit_clocks(tick, ntick, blockIdx.x * 256u + threadIdx.x, gridDim.x * 256u, tick_steps);
    uint32_t part = 0;
    if (tsum2) {
        const uint32_t sup = blockIdx.x / RES_SUPER;
        for (uint32_t u = threadIdx.x; u < sup; u += 256u) part += tsum2[u];
        const uint32_t t = sup * RES_SUPER + threadIdx.x;  // RES_SUPER == 256: one tile per lane
        if (t < blockIdx.x) part += tsum[t];
    } else {
        for (uint32_t t = threadIdx.x; t < blockIdx.x; t += 256u) part += tsum[t];
    }
    uint32_t prefix;
    block_excl_scan<4>(part, &prefix);
    const size_t base = (size_t)blockIdx.x * RES_TILE + (size_t)threadIdx.x * 4u;
    uint4 *a4 = reinterpret_cast<uint4 *>(arr + base);
    const uint4 v = reinterpret_cast<const uint4 *>(counts + base)[0], w = *a4;
    const uint32_t c[4] = {v.x, v.y, v.z, v.w};
    const uint32_t a[4] = {w.x, w.y, w.z, w.w};
    uint32_t bits = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) bits |= (c[k] != 0u ? 1u : 0u) << k;
    // one occupancy word per 8 lanes (32 buckets)
    bits <<= 4u * (threadIdx.x & 7u);
    bits |= __shfl_xor(bits, 1);
    bits |= __shfl_xor(bits, 2);
    bits |= __shfl_xor(bits, 4);
    if ((threadIdx.x & 7u) == 0u) occ[base / 32u] = bits;
    uint32_t total;
    uint32_t run = prefix + block_excl_scan<4>(c[0] + c[1] + c[2] + c[3], &total);
    if (blockIdx.x == ntiles - 1u && threadIdx.x == 0) noff[(size_t)ntiles * RES_TILE] = prefix + total;
    uint32_t o[4], ab[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        o[k] = run;
        ab[k] = run + c[k] - a[k];
        run += c[k];
    }
    reinterpret_cast<uint4 *>(noff + base)[0] = make_uint4(o[0], o[1], o[2], o[3]);
    reinterpret_cast<uint4 *>(abase + base)[0] = make_uint4(ab[0], ab[1], ab[2], ab[3]);
    *a4 = make_uint4(0u, 0u, 0u, 0u);
}

// Movers among old slots [lo, hi) (lo <= hi), from the mover bits.
__device__ __forceinline__ uint32_t movers_between(const uint32_t *mbits, uint32_t lo, uint32_t hi) {
    if (lo >= hi) return 0u;
    const uint32_t wl = lo >> 5, wh = hi >> 5;
    const uint32_t low_mask = ~0u << (lo & 31u);        // bits >= lo in word wl
    const uint32_t high_mask = (1u << (hi & 31u)) - 1u;  // bits < hi in word wh
    if (wl == wh) return (uint32_t)__popc(mbits[wl] & low_mask & high_mask);
    uint32_t n = (uint32_t)__popc(mbits[wl] & low_mask);
    for (uint32_t x = wl + 1u; x < wh; ++x) n += (uint32_t)__popc(mbits[x]);
    if (high_mask) n += (uint32_t)__popc(mbits[wh] & high_mask);
    return n;
}

// Rebuild, level 3: every record moves to its slot of the new layout. A bird
// that kept its bucket keeps its order among the bucket's stayers; a mover is
// placed after the stayers of its new bucket, by arrival rank.
__global__ __launch_bounds__(TILE) void k_r_move(Geometry g, Pools p, Records cur, Records nxt, ResWork w,
                                                 const uint32_t *__restrict__ off_cur,
                                                 const uint32_t *__restrict__ noff) {
    const uint32_t s = blockIdx.x * TILE + threadIdx.x;
    if (s >= g.n_slots) return;
    // the bucket of the new position, by its flock's hash (as the flight found it)
    const float4 p4 = cur.p4[s];
    const Seg &sg = g.segs[slot_segment(g, cur, s)];
    const uint32_t ko = cur.key[s],
                   kn = bucket_of_cell(sg, cell_coord(p4.x, sg.inv_cell), cell_coord(p4.y, sg.inv_cell),
                                       cell_coord(p4.z, sg.inv_cell));
    uint32_t ns;
    if (kn == ko) {
        const uint32_t o = off_cur[ko];
        if (o > s) {
            atomicOr(p.err, ERR_SLOT);
            return;
        }
        ns = noff[ko] + (s - o) - movers_between(w.mbits, o, s);
    } else {
        ns = w.abase[kn] + w.arank[s];
    }
    if (ns >= g.n_slots) {
        atomicOr(p.err, ERR_SLOT);
        return;
    }
    nxt.p4[ns] = p4;
    nxt.v4[ns] = cur.v4[s];
    nxt.s4[ns] = cur.s4[s];
    nxt.spin[ns] = cur.spin[s];
    nxt.id[ns] = cur.id[s];
    nxt.key[ns] = kn;
}

// The rebuild's mover accounting for the current records without a flight:
// each bird against the bucket of its slot (as fly_loaded counts after
// flying). One lane per slot, whole waves (the mover bits are ballots).
__global__ __launch_bounds__(TILE) void k_r_account(Geometry g, Records cur, ResWork w) {
    const uint32_t s = blockIdx.x * TILE + threadIdx.x;
    bool mover = false;
    if (s < g.n_slots) {
        const float4 p4 = cur.p4[s];
        const Seg &sg = g.segs[slot_segment(g, cur, s)];
        const uint32_t ko = cur.key[s];
        const uint32_t kn = bucket_of_cell(sg, cell_coord(p4.x, sg.inv_cell), cell_coord(p4.y, sg.inv_cell),
                                           cell_coord(p4.z, sg.inv_cell));
        if (kn != ko) {
            count_mover(w, s, kn, ko);
            mover = true;
        }
    }
    write_mover_bits(g, w, s, mover);
}

hipError_t launch_resident_account(const Geometry &g, const Pools &p, const Records &cur, const ResWork &w,
                                   hipStream_t s) {
    (void)p;
    if (g.n_slots == 0) return hipSuccess;
    hipLaunchKernelGGL(k_r_account, dim3((g.n_slots + TILE - 1) / TILE), dim3(TILE), 0, s, g, cur, w);
    return hipGetLastError();
}

hipError_t launch_resident_enter(const Geometry &g, const Pools &p, const Records &rec, hipStream_t s) {
    if (g.n_birds == 0) return hipSuccess;
    hipLaunchKernelGGL(k_r_enter, dim3((g.n_birds + TILE - 1) / TILE), dim3(TILE), 0, s, g, p, rec, 0u, g.n_birds);
    return hipGetLastError();
}

hipError_t launch_resident_enter_range(const Geometry &g, const Pools &p, const Records &rec, uint32_t bird_lo,
                                       uint32_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_r_enter, dim3((count + TILE - 1) / TILE), dim3(TILE), 0, s, g, p, rec, bird_lo,
                       bird_lo + count);
    return hipGetLastError();
}

hipError_t launch_resident_splice(const Geometry &g, const Pools &p, const Records &cur, const Records &nxt,
                                  const int32_t *shift, uint32_t n_cur, uint32_t n_nxt, hipStream_t s) {
    if (n_cur == 0) return hipSuccess;
    hipLaunchKernelGGL(k_r_splice, dim3((n_cur + TILE - 1) / TILE), dim3(TILE), 0, s, g, p, cur, nxt, shift, n_cur,
                       n_nxt);
    return hipGetLastError();
}

hipError_t launch_resident_tile_sums(const Geometry &g, const Pools &p, const ResWork &w, hipStream_t s) {
    const uint32_t ntiles = (g.n_buckets + RES_TILE - 1) / RES_TILE;
    if (ntiles) hipLaunchKernelGGL(k_r_tile_sums, dim3(ntiles), dim3(256), 0, s, p.counts, w.tsum);
    return hipGetLastError();
}

hipError_t launch_resident_flush(const Geometry &g, const Pools &p, const Records &rec, hipStream_t s) {
    if (g.n_slots == 0) return hipSuccess;
    hipLaunchKernelGGL(k_r_flush, dim3((g.n_slots + TILE - 1) / TILE), dim3(TILE), 0, s, g, p, rec, 0u, g.n_slots);
    return hipGetLastError();
}

hipError_t launch_resident_flush_range(const Geometry &g, const Pools &p, const Records &rec, uint32_t slot_lo,
                                       uint32_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_r_flush, dim3((count + TILE - 1) / TILE), dim3(TILE), 0, s, g, p, rec, slot_lo,
                       slot_lo + count);
    return hipGetLastError();
}

// Commits the clocks of a layout with no buckets (every flock left a batch).
__global__ void k_commit(Seg *tick, uint32_t ntick, uint32_t tick_steps) {
    commit_clocks(tick, ntick, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, tick_steps);
}

hipError_t launch_resident_rebuild(const Geometry &g, const Pools &p, const Records &cur, const Records &nxt,
                                   const ResWork &w, const uint32_t *off_cur, uint32_t *off_nxt, Seg *tick,
                                   uint32_t ntick, hipStream_t s, uint32_t tick_steps) {
    const uint32_t ntiles = (g.n_buckets + RES_TILE - 1) / RES_TILE;
    if (ntiles == 0) {
        if (tick && ntick) hipLaunchKernelGGL(k_commit, dim3(1), dim3(64), 0, s, tick, ntick, tick_steps);
        return hipGetLastError();
    }
    const uint32_t *tsum2 = nullptr;
    if (ntiles > RES_DIRECT_TILES) {
        hipLaunchKernelGGL(k_r_super_sums, dim3((ntiles + RES_SUPER - 1) / RES_SUPER), dim3(256), 0, s, w.tsum,
                           ntiles, w.tsum2);
        tsum2 = w.tsum2;
    }
    hipLaunchKernelGGL(k_r_scan_apply, dim3(ntiles), dim3(256), 0, s, w.counts, w.arr, w.tsum, tsum2, ntiles,
                       off_nxt, w.abase, p.occ, tick, tick ? ntick : 0u, tick_steps);
    if (g.n_slots)
        hipLaunchKernelGGL(k_r_move, dim3((g.n_slots + TILE - 1) / TILE), dim3(TILE), 0, s, g, p, cur, nxt, w,
                           off_cur, off_nxt);
    return hipGetLastError();
}

hipError_t launch_set_time(Seg *segs, uint32_t nseg, float time, float dt, hipStream_t s) {
    if (nseg == 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_time, dim3((nseg + 255u) / 256u), dim3(256), 0, s, segs, nseg, time, dt);
    return hipGetLastError();
}

hipError_t grid_kernel_attributes(const char *name, hipFuncAttributes *out) {
    if (!__builtin_strcmp(name, "hash_count"))
        return hipFuncGetAttributes(out, reinterpret_cast<const void *>(k_hash_count<false>));
    if (!__builtin_strcmp(name, "scan"))
        return hipFuncGetAttributes(out, reinterpret_cast<const void *>(k_scan_apply<true>));
    if (!__builtin_strcmp(name, "scatter"))
        return hipFuncGetAttributes(out, reinterpret_cast<const void *>(k_scatter));
    if (!__builtin_strcmp(name, "canonical"))
        return hipFuncGetAttributes(out, reinterpret_cast<const void *>(k_canonical));
    return hipErrorInvalidValue;
}

}  // namespace fh
