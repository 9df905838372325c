// decide's kernels and launchers, instantiated for instances without a window
// (decide.hip) and with one (decide_win.hip), so both build in parallel.
//
// decide for gfx950: each workgroup takes a run of sorted slots, compacts the
// birds that are due into LDS, and gives every lane one due bird. Birds are
// visited in slot order, so lanes of a wave that share a cell walk the same
// buckets. The candidates come either from the sorted copies the scatter
// writes (32-byte records: position, then velocity) or, in the resident
// layout, from the state itself kept in slot order (positions and velocities
// in their own arrays); a candidate is known by its slot.
//
// Mirrors crates/flock-ref/src/model.rs:53-145 over the grid of
// cuda/flock.cu:190-290. The reference walks twice, once over the box of
// radius r (sphere test r^2) and once over the box of 2r (sphere test 4r^2).
// Here one walk covers the 2r box: every cell of the r walk is a cell of the
// 2r walk (floor is monotone and r^2 <= 4r^2) and z-y-x order restricted to
// the r box is the r walk's own order, so each accumulator receives the same
// candidates in the same order as in the reference; walk-1 membership is the
// oracle's own test d^2 <= r^2.
#pragma once

#include <hip/hip_ext.h>

#include <cstdio>
#include <type_traits>

#include "flight.h"
#include "layout.h"
#include "model.h"

namespace fh {

struct Decision {
    f3 force;
    float radius;
    uint32_t n_all;
    float centrality;
};

// Where a walk reads candidate t: position pos[STRIDE t], velocity
// vel[STRIDE t] (STRIDE 2: interleaved sorted records; 1: resident arrays).
// With a window, the positions and velocities of slots [wlo, wlo + wn) are
// also held in LDS (winp, winv) and read from there.
// The workgroup's dynamic LDS: the due-slot list, then (with a window) the
// window's positions and velocities. Window reads index this array, so the
// compiler sees LDS loads, never a generic pointer into LDS.
extern __shared__ uint32_t decide_lds[];

struct Cand {
    const float4 *pos, *vel;
    uint32_t wpos = 0, wvel = 0;  // word offsets of the window's positions and velocities in decide_lds
    uint32_t wlo = 0, wn = 0;
};

// WIN: the instance stages a window (FH_FLAG_WINDOW). Without one, a
// candidate is a plain global load; with one, each read chooses between LDS
// and global memory (a generic load, slower when the window is off: the
// instances are separate).
template <int STRIDE, bool WIN>
__device__ __forceinline__ float4 cand_pos(const Cand &c, uint32_t t) {
    if constexpr (WIN) {
        const uint32_t o = t - c.wlo;  // wraps for t < wlo
        return o < c.wn ? reinterpret_cast<const float4 *>(decide_lds + c.wpos)[o] : c.pos[(size_t)STRIDE * t];
    } else {
        return c.pos[(size_t)STRIDE * t];
    }
}

template <int STRIDE, bool WIN>
__device__ __forceinline__ float4 cand_vel(const Cand &c, uint32_t t) {
    if constexpr (WIN) {
        const uint32_t o = t - c.wlo;
        return o < c.wn ? reinterpret_cast<const float4 *>(decide_lds + c.wvel)[o] : c.vel[(size_t)STRIDE * t];
    } else {
        return c.vel[(size_t)STRIDE * t];
    }
}

// Per-bird walk. `visit(slot)` sees each candidate slot that passes the
// self and exact-cell checks and d^2 <= 4r^2, in walk order.
struct Walk {
    f3 p;
    float r, r2, c2;
    int32_t lo[3], hi[3];
    bool ok;
};

// A candidate lies in cell (fx, fy, fz), the cell's integer coordinates as
// floats, when the floors of its scaled coordinates equal them: the test of
// cell_coord without the conversion, exact while |cell| < 2^24 (the walk's
// bound), and false for a NaN coordinate (the oracle's d^2 test rejects such
// a candidate too).
constexpr int32_t FLOAT_CELL_MAX = 1 << 24;
__device__ __forceinline__ bool in_cell(float4 c, float inv, float fx, float fy, float fz) {
    return (int)(floorf(c.x * inv) == fx) & (int)(floorf(c.y * inv) == fy) & (int)(floorf(c.z * inv) == fz);
}

// Whether every cell of the walk's box widened by one cell per side falls in
// a bucket of its own, so that the cell test can be left out: a candidate in
// the bucket of a visited cell is then either in that cell or in a cell
// outside the widened box, more than 2r away on some axis (the margin covers
// the rounding of the cell coordinates, under 1/16 cell while they stay
// below 2^20). It holds under the spatial hash when the widened box lies in
// one block of 1024 cells per axis (the hash adds one constant for the block
// to the Morton code of the low 10 bits) and spans at most 2^(b/3) cells per
// axis, b the bits of the bucket mask: the masked Morton code keeps the low
// b/3 bits of every axis, which differ for two cells closer than 2^(b/3).
constexpr int32_t DISTINCT_CELL_MAX = 1 << 20;
__device__ __forceinline__ bool box_buckets_distinct(const Walk &w, const Seg &sg) {
    if (!(sg.mode & SEG_SPATIAL) || (sg.mask & (sg.mask + 1u))) return false;
    const int32_t span = 1 << (__popc(sg.mask) / 3);
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int32_t lo = w.lo[a] - 1, hi = w.hi[a] + 1;  // |w.lo|, |w.hi| < 2^24 when w.ok
        ok = ok && hi - lo < span && (lo >> 10) == (hi >> 10) && lo > -DISTINCT_CELL_MAX && hi < DISTINCT_CELL_MAX;
    }
    return ok;
}

__device__ __forceinline__ Walk make_walk(f3 p, float r, const Seg &sg) {
    Walk w;
    w.p = p;
    w.r = r;
    w.r2 = r * r;
    const float rr = 2.0f * r;
    w.c2 = rr * rr;
    w.lo[0] = cell_coord(p.x - rr, sg.inv_cell);
    w.hi[0] = cell_coord(p.x + rr, sg.inv_cell);
    w.lo[1] = cell_coord(p.y - rr, sg.inv_cell);
    w.hi[1] = cell_coord(p.y + rr, sg.inv_cell);
    w.lo[2] = cell_coord(p.z - rr, sg.inv_cell);
    w.hi[2] = cell_coord(p.z + rr, sg.inv_cell);
    // A finite bird with r <= R_MAX spans at most 4 R_MAX / cell + 2 cells per
    // axis; anything wider comes from a radius or position outside the model
    // and is not walked (the bird then decides with no neighbours), as is a
    // box beyond 2^24 cells from the origin (67,000 km at 4 m cells).
    const int64_t max_extent = (int64_t)(4.0f * R_MAX * sg.inv_cell) + 3;
    w.ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        w.ok = w.ok && (int64_t)w.hi[a] - (int64_t)w.lo[a] <= max_extent && w.hi[a] >= w.lo[a] &&
               w.hi[a] < FLOAT_CELL_MAX && w.lo[a] > -FLOAT_CELL_MAX;
    return w;
}

__device__ __forceinline__ bool bucket_occupied(const Pools &p, uint32_t b) {
    return (p.occ[b >> 5] >> (b & 31u)) & 1u;
}

// Cells of the box are visited in z-y-x order (x fastest), as the nested
// loops of cuda/flock.cu:209-211. A lane visits every G-th cell starting at
// `first` (G = 1: every cell). Its cells are taken PF at a time: the bucket
// ranges of all PF cells are loaded together, and their candidates then form
// one run, loaded AHEAD at a time across bucket ends, so a wave waits one
// memory latency per AHEAD candidates of its fullest lane's run instead of
// once more for every bucket of every lane. With `use_occ` the occupancy bits
// of the PF buckets are read first (a small, cache-resident table) and empty
// buckets never touch the offsets table. Within a lane the visit order is the
// walk order.
//
// With MG (FH_FLAG_LAZY) the PF cells' mover lists are visited too, after
// their buckets: `visit_m(t, d, d2)` sees list entry t. A mover's record
// in its old bucket is never taken (its position lies in another cell), so
// each bird is seen once.
template <int G, int PF, int STRIDE, bool WIN, bool MG, bool CT = true, class Visit, class VisitM>
__device__ __forceinline__ void walk_cells(const Walk &w, const Seg &sg, const Pools &p, Cand cand, uint32_t n_slots,
                                           uint32_t self, int32_t first, bool use_occ, Visit &&visit,
                                           const ResWork &mw, VisitM &&visit_m) {
    if (!w.ok) return;
    constexpr int AHEAD = 4;  // 8 spills at the register budgets of G 2-8
    const int32_t nx = w.hi[0] - w.lo[0] + 1, ny = w.hi[1] - w.lo[1] + 1, nz = w.hi[2] - w.lo[2] + 1;
    const int32_t ncell = nx * ny * nz;
    int32_t x = w.lo[0], y = w.lo[1], z = w.lo[2];
    auto advance = [&](int32_t steps) {
        for (int32_t s = 0; s < steps; ++s)
            if (++x > w.hi[0]) {
                x = w.lo[0];
                if (++y > w.hi[1]) {
                    y = w.lo[1];
                    ++z;
                }
            }
    };
    advance(first);
    for (int32_t base = first; base < ncell; base += G * PF) {
        int32_t cx[PF], cy[PF], cz[PF];
        uint32_t s0[PF], s1[PF], bk[PF];
        bool want[PF];
        [[maybe_unused]] float4 mh[PF];
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            cx[k] = x;
            cy[k] = y;
            cz[k] = z;
            s0[k] = 0u;
            s1[k] = 0u;
            want[k] = base + k * G < ncell && cell_meets_sphere(x, y, z, sg.cell, w.p, w.c2);
            bk[k] = bucket_of_cell(sg, x, y, z);
            if constexpr (MG)
                mh[k] = want[k] && mw.mr_head4 ? mw.mr_head4[bk[k] & mw.mmask]
                                               : make_float4(0.f, 0.f, 0.f, __uint_as_float(MOVER_NIL));
#pragma unroll
            for (int s = 0; s < G; ++s)
                if (++x > w.hi[0]) {
                    x = w.lo[0];
                    if (++y > w.hi[1]) {
                        y = w.lo[1];
                        ++z;
                    }
                }
        }
        if (use_occ) {
            uint32_t word[PF];
#pragma unroll
            for (int k = 0; k < PF; ++k) word[k] = want[k] ? p.occ[bk[k] >> 5] : 0u;
#pragma unroll
            for (int k = 0; k < PF; ++k) want[k] = (word[k] >> (bk[k] & 31u)) & 1u;
        }
#pragma unroll
        for (int k = 0; k < PF; ++k)
            if (want[k]) {
                const uint2 r = bucket_range(p.offsets, bk[k]);
                s0[k] = r.x;
                s1[k] = r.y;
            }
        // The candidates of the PF buckets are taken as one run, in walk order,
        // AHEAD loads at a time across bucket ends: a wave's rounds follow its
        // lanes' total counts over the PF buckets (one memory latency each),
        // not the fullest lane of every bucket in turn.
        uint32_t fend[PF], foff[PF];
        float fcx[PF], fcy[PF], fcz[PF];
        uint32_t total = 0u;
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            s1[k] = min(s1[k], n_slots);
            foff[k] = s0[k] - total;  // slot = run index + foff[k] within bucket k (wrapping)
            total += s0[k] < s1[k] ? s1[k] - s0[k] : 0u;
            fend[k] = total;
            fcx[k] = (float)cx[k];
            fcy[k] = (float)cy[k];
            fcz[k] = (float)cz[k];
        }
        for (uint32_t j0 = 0; j0 < total; j0 += AHEAD) {
            float4 c[AHEAD];
            uint32_t tt[AHEAD];
#pragma unroll
            for (int u = 0; u < AHEAD; ++u) {
                // past the run's end a lane re-reads its last candidate (every
                // load in bounds, no branch); the predicate drops it
                const uint32_t j = min(j0 + (uint32_t)u, total - 1u);
                uint32_t o = foff[PF - 1];
#pragma unroll
                for (int k = PF - 2; k >= 0; --k) o = j < fend[k] ? foff[k] : o;
                tt[u] = j + o;
                c[u] = cand_pos<STRIDE, WIN>(cand, tt[u]);
            }
#pragma unroll
            for (int u = 0; u < AHEAD; ++u) {
                const uint32_t j = j0 + (uint32_t)u;
                float fx = fcx[PF - 1], fy = fcy[PF - 1], fz = fcz[PF - 1];
#pragma unroll
                for (int k = PF - 2; k >= 0; --k)
                    if (j < fend[k]) {
                        fx = fcx[k];
                        fy = fcy[k];
                        fz = fcz[k];
                    }
                const f3 d = mk3(c[u].x - w.p.x, c[u].y - w.p.y, c[u].z - w.p.z);
                const float d2 = dot3(d, d);
                // one predicate, no branch per test: inside the run, not the
                // bird itself, in this cell, within 2r
                if ((int)(j < total) & (int)(tt[u] != self) & (int)(!CT || in_cell(c[u], sg.inv_cell, fx, fy, fz)) &
                    (int)(d2 <= w.c2))
                    visit(tt[u], d, d2);
            }
        }
        if constexpr (MG) {
#pragma unroll
            for (int k = 0; k < PF; ++k) {
                const uint32_t hw = __float_as_uint(mh[k].w);
                if (hw != MOVER_NIL) {
                    uint32_t t = hw & ~MOVER_OVF;  // the head's bird, then the overflow chain
                    float4 c = mh[k];
                    uint32_t next = (hw & MOVER_OVF) ? mw.mr_ovf[bk[k] & mw.mmask] : MOVER_NIL;
                    // a list holds each slot once at most: n_slots entries bound any walk of it.
                    // Lists are shared by the flocks of a batch: a listed bird counts only for
                    // walks of its own flock (its slot in the flock's slot range).
                    for (uint32_t it = 0; t < n_slots && it < n_slots; ++it) {
                        const f3 d = mk3(c.x - w.p.x, c.y - w.p.y, c.z - w.p.z);
                        const float d2 = dot3(d, d);
                        if ((int)(t != self) & (int)in_cell(c, sg.inv_cell, fcx[k], fcy[k], fcz[k]) &
                            (int)(d2 <= w.c2) & (int)(t - sg.slot_base < sg.n))
                            visit_m(t, d, d2);
                        t = next;
                        if (t >= n_slots) break;
                        c = mw.mr_p4[t];
                        next = mw.mr_next[t];
                    }
                }
            }
        }
    }
}

// Candidate-split walk for dense cells: the G lanes of a group walk the same
// cells in walk order, and lane j takes candidates s0 + j, s0 + j + G, ... of
// each bucket, so one bucket's candidates are read with coalesced loads and
// spread over the group. PF cells' bucket ranges are loaded together (every
// lane of the group loads the same words); a lane then loads UNROLL of its
// candidates (position and velocity) before it tests the first, so one
// memory latency covers UNROLL candidates. `visit(d, d2, velocity)`.
// With MG, each cell's mover list follows its own bucket, taken by the
// group's first lane (lists hold a few birds at most).
template <int G, int PF, int STRIDE, bool WIN, bool MG, bool CT, class Visit>
__device__ __forceinline__ void walk_cells_split(const Walk &w, const Seg &sg, const Pools &p, Cand cand,
                                                 uint32_t n_slots, uint32_t self, uint32_t lane, bool use_occ,
                                                 Visit &&visit, const ResWork &mw) {
    if (!w.ok) return;
    const int32_t nx = w.hi[0] - w.lo[0] + 1, ny = w.hi[1] - w.lo[1] + 1, nz = w.hi[2] - w.lo[2] + 1;
    const int32_t ncell = nx * ny * nz;
    int32_t x = w.lo[0], y = w.lo[1], z = w.lo[2];
    for (int32_t base = 0; base < ncell; base += PF) {
        int32_t cx[PF], cy[PF], cz[PF];
        uint32_t s0[PF], s1[PF], bk[PF];
        bool want[PF];
        [[maybe_unused]] float4 mh[PF];
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            cx[k] = x;
            cy[k] = y;
            cz[k] = z;
            s0[k] = 0u;
            s1[k] = 0u;
            want[k] = base + k < ncell && cell_meets_sphere(x, y, z, sg.cell, w.p, w.c2);
            bk[k] = bucket_of_cell(sg, x, y, z);
            if constexpr (MG)
                mh[k] = want[k] && lane == 0u && mw.mr_head4 ? mw.mr_head4[bk[k] & mw.mmask]
                                                             : make_float4(0.f, 0.f, 0.f, __uint_as_float(MOVER_NIL));
            if (++x > w.hi[0]) {
                x = w.lo[0];
                if (++y > w.hi[1]) {
                    y = w.lo[1];
                    ++z;
                }
            }
        }
        if (use_occ) {
            uint32_t word[PF];
#pragma unroll
            for (int k = 0; k < PF; ++k) word[k] = want[k] ? p.occ[bk[k] >> 5] : 0u;
#pragma unroll
            for (int k = 0; k < PF; ++k) want[k] = (word[k] >> (bk[k] & 31u)) & 1u;
        }
#pragma unroll
        for (int k = 0; k < PF; ++k)
            if (want[k]) {
                const uint2 r = bucket_range(p.offsets, bk[k]);
                s0[k] = r.x;
                s1[k] = min(r.y, n_slots);
            }
        constexpr int UNROLL = 2;
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            for (uint32_t t0 = s0[k] + lane; t0 < s1[k]; t0 += UNROLL * G) {
                float4 c[UNROLL], v[UNROLL];
#pragma unroll
                for (int u = 0; u < UNROLL; ++u) {
                    // past the bucket's end a lane re-reads its last candidate
                    // (every load in bounds, no branch); the predicate drops it
                    const uint32_t t = min(t0 + (uint32_t)u * G, s1[k] - 1u);
                    c[u] = cand_pos<STRIDE, WIN>(cand, t);
                    v[u] = cand_vel<STRIDE, WIN>(cand, t);
                }
                const float fx = (float)cx[k], fy = (float)cy[k], fz = (float)cz[k];
#pragma unroll
                for (int u = 0; u < UNROLL; ++u) {
                    const uint32_t t = t0 + (uint32_t)u * G;
                    const f3 d = mk3(c[u].x - w.p.x, c[u].y - w.p.y, c[u].z - w.p.z);
                    const float d2 = dot3(d, d);
                    // inside the bucket, not the bird itself, in this cell, within 2r
                    if ((int)(t < s1[k]) & (int)(t != self) & (int)(!CT || in_cell(c[u], sg.inv_cell, fx, fy, fz)) &
                        (int)(d2 <= w.c2))
                        visit(d, d2, v[u]);
                }
            }
            if constexpr (MG) {
                const uint32_t hw = __float_as_uint(mh[k].w);
                if (hw != MOVER_NIL) {
                    const float fx = (float)cx[k], fy = (float)cy[k], fz = (float)cz[k];
                    uint32_t t = hw & ~MOVER_OVF;
                    float4 c = mh[k];
                    uint32_t next = (hw & MOVER_OVF) ? mw.mr_ovf[bk[k] & mw.mmask] : MOVER_NIL;
                    for (uint32_t it = 0; t < n_slots && it < n_slots; ++it) {
                        const f3 d = mk3(c.x - w.p.x, c.y - w.p.y, c.z - w.p.z);
                        const float d2 = dot3(d, d);
                        if ((int)(t != self) & (int)in_cell(c, sg.inv_cell, fx, fy, fz) & (int)(d2 <= w.c2) &
                            (int)(t - sg.slot_base < sg.n))
                            visit(d, d2, mw.mr_v4[t]);
                        t = next;
                        if (t >= n_slots) break;
                        c = mw.mr_p4[t];
                        next = mw.mr_next[t];
                    }
                }
            }
        }
    }
}

// The running sums of one bird's walk (model.rs:62-115).
struct Acc {
    f3 sep, coh, align, csum;
    uint32_t n_all, n_star, n_align, n_g;
};

__device__ __forceinline__ void acc_zero(Acc &a) {
    a.sep = mk3(0.f, 0.f, 0.f);
    a.coh = a.sep;
    a.align = a.sep;
    a.csum = a.sep;
    a.n_all = a.n_star = a.n_align = a.n_g = 0;
}

// Per-hit arithmetic. A walk's hits (candidates within 2r) take the hardware
// square root, reciprocal, exponential and reciprocal square root (1 ulp
// each) instead of a correctly rounded square root and divisions: the unit
// offset is d times the reciprocal of the distance, the separation weight
// the hardware exponential, a neighbour's heading its velocity times the
// reciprocal square root of its length (hit_heading: per aligned hit on
// resident records, once per bird and step by the scatter for the sorted
// copies, the same bits). Neighbour sets and counts are
// unchanged (they follow d^2 alone, computed as the oracle does); forces move
// by a few roundings, far inside the reference's bars. Measured on MI355X
// against correctly rounded arithmetic: the 1M-bird step 6.8% faster, the
// 256-flock batch 3.2%, the clumped 10M-bird step 30% (139.5 -> 97.3 ms),
// where hundreds to thousands of hits per bird make the arithmetic the bound.
// (Queueing the hits in LDS and summing them after the walk, which executes
// the sums of 60% fewer hits with exact arithmetic, made decide 3-12%
// slower: LDS traffic and occupancy cost more than the masked sums.)

// walk 2 (model.rs:104-115) for a candidate with d^2 <= 4r^2: csum += d /
// max(sqrt(d2), 1e-6); returns the unit offset and the distance.
__device__ __forceinline__ void accumulate_2r(Acc &a, f3 d, float d2, f3 &dhat, float &dist) {
    dist = __builtin_amdgcn_sqrtf(d2);
    dhat = d * __builtin_amdgcn_rcpf(fmaxf(dist, 1e-6f));
    a.csum = a.csum + dhat;
    ++a.n_g;
}

// walk 1 (model.rs:67-93) for a candidate within r; `velocity()` gives its
// velocity (read only for an aligned neighbour), or with HEAD its heading
// hit_heading(velocity) as the sorted copies hold it (the same bits).
template <bool HEAD, class Velocity>
__device__ __forceinline__ void accumulate_r(Acc &a, f3 f_i, f3 dhat, float dist, Velocity &&velocity) {
    ++a.n_all;
    const float g = dist <= R_H ? 1.0f : __expf(-((dist - R_H) * (dist - R_H)) * (1.0f / SEP_DENOM));
    a.sep = a.sep - dhat * g;
    if (dist > R_H) {
        a.coh = a.coh + dhat;
        ++a.n_star;
    }
    if (dot3(f_i, dhat) >= BLIND_COS) {
        const float4 vj = velocity();
        const f3 v = mk3(vj.x, vj.y, vj.z);
        if constexpr (HEAD)
            a.align = a.align + v;
        else
            a.align = a.align + hit_heading(v);
        ++a.n_align;
    }
}

// One candidate with offset d and d^2 <= 4r^2.
template <bool HEAD, class Velocity>
__device__ __forceinline__ void accumulate(Acc &a, f3 f_i, float r2, f3 d, float d2, Velocity &&velocity) {
    f3 dhat;
    float dist;
    accumulate_2r(a, d, d2, dhat, dist);
    if (d2 <= r2) accumulate_r<HEAD>(a, f_i, dhat, dist, velocity);
}

template <int G, int PF, int STRIDE, bool WIN, bool MG>
__device__ __forceinline__ Acc walk_sums(const Pools &p, Cand cand, const Seg &sg, uint32_t n_slots, uint32_t self,
                                         const Walk &w, f3 f_i, int32_t first, bool use_occ, const ResWork &mw) {
    constexpr bool HEAD = STRIDE == 2;  // sorted copies hold headings, resident records velocities
    Acc a;
    acc_zero(a);
    // Between lazy rebuilds a bucket's slots also hold birds that have left
    // its cell (listed again where they are now): those walks keep the cell
    // test. Others leave it out where the box's buckets are all distinct.
    auto walk = [&](auto ct) {
        constexpr bool CT = decltype(ct)::value;
        if constexpr (G <= 8)
            walk_cells<G, PF, STRIDE, WIN, MG, CT>(
                w, sg, p, cand, n_slots, self, first, use_occ,
                [&](uint32_t t, f3 d, float d2) {
                    accumulate<HEAD>(a, f_i, w.r2, d, d2, [&] { return cand_vel<STRIDE, WIN>(cand, t); });
                },
                mw,
                [&](uint32_t t, f3 d, float d2) { accumulate<false>(a, f_i, w.r2, d, d2, [&] { return mw.mr_v4[t]; }); });
        else
            walk_cells_split<G, PF, STRIDE, WIN, MG, CT>(w, sg, p, cand, n_slots, self, (uint32_t)first, use_occ,
                                                         [&](f3 d, float d2, float4 vj) {
                                                             accumulate<HEAD>(a, f_i, w.r2, d, d2, [&] { return vj; });
                                                         },
                                                         mw);
    };
    if (MG || !box_buckets_distinct(w, sg))
        walk(std::true_type{});
    else if constexpr (!MG)
        walk(std::false_type{});
    return a;
}

__device__ __forceinline__ f3 shfl_xor3(f3 v, int m) {
    return mk3(__shfl_xor(v.x, m), __shfl_xor(v.y, m), __shfl_xor(v.z, m));
}

// Sums the partial walks of the G lanes of a group (butterfly; every lane
// ends with bit-identical totals because float addition is commutative).
template <int G>
__device__ __forceinline__ void reduce_group(Acc &a) {
#pragma unroll
    for (int m = 1; m < G; m <<= 1) {
        a.sep = a.sep + shfl_xor3(a.sep, m);
        a.coh = a.coh + shfl_xor3(a.coh, m);
        a.align = a.align + shfl_xor3(a.align, m);
        a.csum = a.csum + shfl_xor3(a.csum, m);
        a.n_all += __shfl_xor(a.n_all, m);
        a.n_star += __shfl_xor(a.n_star, m);
        a.n_align += __shfl_xor(a.n_align, m);
        a.n_g += __shfl_xor(a.n_g, m);
    }
}

// combine (model.rs:95-137, cuda/flock.cu:242-283)
__device__ __forceinline__ Decision finish(const Acc &a, f3 pos, float speed, f3 f_i, float r) {
    f3 force = mk3(0.f, 0.f, 0.f);
    if (a.n_all > 0) force = force + a.sep * (W_SEP / (float)a.n_all);
    const float centrality = a.n_g > 0 ? sqrtf(dot3(a.csum, a.csum)) / (float)a.n_g : 0.0f;
    if (a.n_star > 0) force = force + normalize_or_zero(a.coh) * (W_COH * centrality);
    const f3 al = a.n_align > 0 ? a.align / (float)a.n_align - f_i : mk3(0.f, 0.f, 0.f);
    force = force + al * W_ALIGN;
    force = force + mk3(0.f, 0.f, 1.f) * (W_ROOST_V * clampf_((Z0 - pos.z) / H_ROOST_V, -1.0f, 1.0f));
    force = force + f_i * ((MASS / TAU) * (V0 - speed));
    Decision dcs;
    dcs.force = force;
    dcs.radius = radius_step(r, a.n_all);
    dcs.n_all = a.n_all;
    dcs.centrality = centrality;
    return dcs;
}

// Cells of the union box the tiled path may walk, and how much larger than
// the largest single box the union may be, before the per-lane walk is used.
constexpr int64_t TILED_MAX_CELLS = 4096;
constexpr int64_t TILED_MAX_RATIO = 8;
constexpr uint32_t NO_BIRD = 0xffffffffu;

// BLOCK lanes per workgroup, G lanes per due bird. With G = 1 each bird's sums
// follow the walk order exactly (the order of cuda/flock.cu:209-238); with
// G = 2..8 lane j sums cells j, j+G, ...; with G = 16 or 64 lane j sums
// candidates j, j+G, ... of every bucket; the G partial sums are then added.
//
// Tiled path (G = 1): when the due birds a workgroup holds at once share
// their neighbourhood (a dense cell: they sit together in slot order), the
// workgroup walks the union of their boxes in z-y-x order. For each cell some
// lane needs, the bucket's candidates are staged in LDS BLOCK at a time with
// coalesced loads and every lane that needs the cell reads them from LDS
// (broadcast). A lane sees the same (cell, slot) sequence as in its own walk,
// so its sums are bit-identical to the per-lane path.
//
// Persistent form (WALK_PERSIST): a grid sized to the machine rather than to
// the flock; each workgroup takes chunks from *p.work until they run out, so
// a workgroup that drew a sparse chunk moves on at once and the tail of the
// launch is the last chunk's walk, not the slowest workgroup's share.
// Waves per SIMD the compiler keeps room for (1: no floor). With G >= 4 the
// walk fits 95 VGPRs without spilling, so 5 waves per SIMD instead of the 4
// that its natural 105-115 VGPRs allow (measured: -2% to -9% decide time).
// G = 2 is held to 128 VGPRs (4 waves): it fits without the fused flight, and
// with it (136 VGPRs, 3 waves, left alone) one spilled register outside the
// walk buys the fourth wave (fused step at 1M birds: 0.127 -> 0.099 ms).
// The split walks (G >= 16) load 2 candidates ahead instead of 4 and run 6
// waves in 80 VGPRs (dense clumps at 10M birds: -14% per step; reference
// density at 1M: -3% decide). G = 1 would spill inside the walk.
constexpr int decide_min_waves(int g) { return g >= 16 ? 6 : g >= 4 ? 5 : g == 2 ? 4 : 1; }

// MG: the walks also visit the mover grid (FH_FLAG_LAZY; resident records with
// the fused flight only), and the flight counts movers into it unless the step
// rebuilds (WALK_LAZY).
template <int BLOCK, int G, bool RES, bool FLY, bool WIN, bool MG>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(decide_min_waves(G), 8))) void k_decide(
    Geometry g, Pools p, Records rec, float time, uint32_t chunk, int walk_flags, uint32_t *due_clear,
    uint32_t *work_clear, ResWork w, Records flown) {
    static_assert(!MG || (RES && FLY && !WIN), "the mover grid goes with resident records and the fused flight");
    const ResWork &movers = w;  // inside the walk loop, `w` is the bird's Walk
    const bool allow_tiled = !MG && (walk_flags & WALK_TILED) != 0, use_occ = (walk_flags & WALK_OCC) != 0;
    const bool persist = (walk_flags & WALK_PERSIST) != 0;
    // Cells whose bucket ranges a lane loads at once. The split walks (G >= 16)
    // take one cell at a time: their buckets are long, and the smaller code
    // (the candidate loop is unrolled per cell) runs faster. A lane of G = 4-8
    // walks one or two cells of a typical box and loads its candidates 4
    // ahead, so one cell at a time keeps it within 95 VGPRs (5 waves per
    // SIMD) without spilling. Fewer registers for the prefetch mean more
    // waves (G = 2 with 4 cells instead of 8: -20% decide time at 1M birds).
    constexpr int PF = G >= 4 ? 1 : 4;
    constexpr int STRIDE = RES ? 1 : 2;
    Cand cand = RES ? Cand{rec.p4, rec.v4} : Cand{p.sbird, p.sbird + 1};
    // decide_lds: [chunk] due slots, then with a window float4 positions and
    // velocities [window_slots(chunk)] each
    uint32_t *const due_slots = decide_lds;
    float4 *const winp = reinterpret_cast<float4 *>(decide_lds + chunk);
    float4 *const winv = winp + window_slots(chunk);
    __shared__ uint32_t n_due;
    __shared__ int32_t u_lo[3], u_hi[3];
    __shared__ uint32_t u_seg_lo, u_seg_hi, u_vol, u_any;
    __shared__ float4 tile[G == 1 ? BLOCK : 1];  // the tiled path (G = 1 only)
    __shared__ uint32_t cur_chunk;
    if (RES && blockIdx.x == 0 && threadIdx.x == 0) {
        if (due_clear) *due_clear = 0;
        if (work_clear) *work_clear = 0;
    }
    [[maybe_unused]] float now = 0.0f;  // the step's clock (one lazy flock: from the lazy clock)
    if constexpr (MG) {
        // One lazy flock reads its clock from mclock_now and commits the
        // next; a batch's flocks take theirs from the segment table (seg_step).
        if (w.mclock_now) now = *w.mclock_now;
        if (w.mclock_now && blockIdx.x == 0 && threadIdx.x == 0) {
            // The next step's clock; between rebuilds also the segment's (its
            // clock field is read by no launch of this step).
            *w.mclock_next = now + DT;
            if ((walk_flags & WALK_LAZY) && w.mtick->live && w.mtick->left) w.mtick->time = now + DT;
        }
        // mover lists a later step of the period writes start empty
        for (int j = 0; j < 2; ++j)
            if (w.mc_head4[j])
                for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i <= w.mmask; i += gridDim.x * BLOCK) {
                    reinterpret_cast<uint32_t *>(&w.mc_head4[j][i])[3] = MOVER_NIL;
                    w.mc_ovf[j][i] = MOVER_NIL;
                }
    }
    const uint32_t n_chunks = (g.n_slots + chunk - 1) / chunk;
    for (uint32_t round = 0;; ++round) {  // one chunk per trip; every lane leaves together
    if (threadIdx.x == 0) {
        n_due = 0;
        uint32_t c = round ? n_chunks : blockIdx.x;
        if (persist) {
            c = atomicAdd(p.work, 1u);
        } else if (!round && (walk_flags & WALK_XCD)) {
            // The workgroups of XCD x (blockIdx % XCDS == x) take the x-th
            // contiguous run of chunks in order, so the halo cells two
            // neighbouring chunks share are read through one L2.
            const uint32_t per = n_chunks / XCDS, rem = n_chunks % XCDS;
            const uint32_t x = blockIdx.x % XCDS, i = blockIdx.x / XCDS;
            c = x * per + min(x, rem) + i;
        }
        cur_chunk = c;
    }
    __syncthreads();
    const uint32_t c_idx = cur_chunk;
    if (c_idx >= n_chunks) break;
    const uint32_t start = c_idx * chunk;
    const uint32_t end = min(start + chunk, g.n_slots);
    if constexpr (WIN) {
        // Stage the positions and velocities of the chunk and a halo on each
        // side with coalesced loads; with the spatial hash most candidates
        // of the chunk's due birds lie in this window of slots.
        const uint32_t halo = window_halo(chunk);
        const uint32_t wlo = start > halo ? start - halo : 0u;
        const uint32_t whi = min(min(end + halo, g.n_slots), wlo + window_slots(chunk));
        for (uint32_t t = wlo + threadIdx.x; t < whi; t += BLOCK) {
            winp[t - wlo] = cand.pos[(size_t)STRIDE * t];
            winv[t - wlo] = cand.vel[(size_t)STRIDE * t];
        }
        cand.wpos = chunk;
        cand.wvel = chunk + 4u * window_slots(chunk);
        cand.wlo = wlo;
        cand.wn = whi - wlo;
        __syncthreads();
    }
    // Sorted copies: time is relative (0) and snext holds next - time. Resident
    // records: this step's clock is the bird's flock clock + DT (committed by
    // the pass that follows) and p4.w holds next_decision.
    for (uint32_t s = start + threadIdx.x; s < end; s += BLOCK) {
        bool due;
        if (RES) {
            const Seg &sg = g.segs[slot_segment(g, rec, s)];
            if (MG && !w.mclock_now) {  // a batch's lazy step: each flock's clock mphase steps on
                float t;
                const bool live = seg_step(sg, w.mphase, t);
                due = live && t + DT >= cand_pos<STRIDE, WIN>(cand, s).w;
            } else {
                due = sg.left && (MG ? now : sg.time) + DT >= cand_pos<STRIDE, WIN>(cand, s).w;  // frozen flocks never decide
            }
        } else {
            due = time >= p.snext[s];
        }
        if (due) due_slots[atomicAdd(&n_due, 1u)] = s;
    }
    __syncthreads();
    const uint32_t nd = n_due;
    if (threadIdx.x == 0 && nd) atomicAdd(p.due_count, nd);
    constexpr uint32_t groups = BLOCK / G;
    const uint32_t lane_g = threadIdx.x % G;
    for (uint32_t base = 0; base < nd; base += groups) {  // same trip count for every lane
        const uint32_t k = base + threadIdx.x / G;
        bool active = k < nd;
        uint32_t me = 0, si = 0, slot = 0;
        f3 pos = mk3(0.f, 0.f, 0.f), f_i = pos;
        float r = 0.0f, speed = 0.0f;  // speed: |velocity|
        Acc a;
        acc_zero(a);
        Walk w;
        w.ok = false;
        if (active) {
            slot = due_slots[k];
            const float4 q = cand_pos<STRIDE, WIN>(cand, slot);
            // sorted copies: the heading and the radius; resident records: the velocity
            const float4 v = cand_vel<STRIDE, WIN>(cand, slot);
            me = RES ? rec.id[slot] : __float_as_uint(q.w);
            if (me >= g.n_birds) {
                if (lane_g == 0) atomicOr(p.err, ERR_SLOT);
                active = false;
            } else {
                si = g.tile_seg ? g.tile_seg[me / TILE] : 0u;
                pos = mk3(q.x, q.y, q.z);
                // the bird's own heading is exact (forward_of); sorted copies
                // hold the hit heading, so the velocity comes from vel4
                const float4 vb = RES ? v : p.vel4[me];
                f_i = forward_and_length(mk3(vb.x, vb.y, vb.z), speed);
                r = RES ? rec.s4[slot].w : v.w;
                w = make_walk(pos, r, g.segs[si]);
            }
        }
        bool tiled = false;
        if (G == 1) {
            if (threadIdx.x == 0) {
                u_lo[0] = u_lo[1] = u_lo[2] = 2147483647;
                u_hi[0] = u_hi[1] = u_hi[2] = -2147483647 - 1;
                u_seg_lo = 0xffffffffu;
                u_seg_hi = 0u;
                u_vol = 0u;
                u_any = 0u;
            }
            __syncthreads();
            if (active && w.ok) {
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    atomicMin(&u_lo[ax], w.lo[ax]);
                    atomicMax(&u_hi[ax], w.hi[ax]);
                }
                atomicMin(&u_seg_lo, si);
                atomicMax(&u_seg_hi, si);
                const int64_t vol = (int64_t)(w.hi[0] - w.lo[0] + 1) * (w.hi[1] - w.lo[1] + 1) *
                                    (w.hi[2] - w.lo[2] + 1);
                atomicMax(&u_vol, (uint32_t)min<int64_t>(vol, 0x7fffffff));
                u_any = 1u;
            }
            __syncthreads();
            const int64_t ex = (int64_t)u_hi[0] - u_lo[0] + 1, ey = (int64_t)u_hi[1] - u_lo[1] + 1,
                          ez = (int64_t)u_hi[2] - u_lo[2] + 1;
            tiled = allow_tiled && u_any && u_seg_lo == u_seg_hi && ex <= TILED_MAX_CELLS && ey <= TILED_MAX_CELLS &&
                    ez <= TILED_MAX_CELLS && ex * ey * ez <= TILED_MAX_CELLS &&
                    ex * ey * ez <= TILED_MAX_RATIO * (int64_t)u_vol;
        }
        if (tiled) {
            const Seg &sg = g.segs[u_seg_lo];
            const int32_t lx = u_lo[0], ly = u_lo[1], lz = u_lo[2];
            const int32_t hx = u_hi[0], hy = u_hi[1], hz = u_hi[2];
            for (int32_t cz = lz; cz <= hz; ++cz)
                for (int32_t cy = ly; cy <= hy; ++cy)
                    for (int32_t cx = lx; cx <= hx; ++cx) {
                        const bool need = active && w.ok && cx >= w.lo[0] && cx <= w.hi[0] && cy >= w.lo[1] &&
                                          cy <= w.hi[1] && cz >= w.lo[2] && cz <= w.hi[2] &&
                                          cell_meets_sphere(cx, cy, cz, sg.cell, w.p, w.c2);
                        if (!__syncthreads_or(need)) continue;
                        const uint32_t b = bucket_of_cell(sg, cx, cy, cz);
                        if (use_occ && !bucket_occupied(p, b)) continue;  // same b for every lane
                        const uint2 rg = bucket_range(p.offsets, b);
                        const uint32_t s0 = rg.x;
                        const uint32_t s1 = min(rg.y, g.n_slots);
                        for (uint32_t t0 = s0; t0 < s1; t0 += BLOCK) {
                            const uint32_t t = t0 + threadIdx.x;
                            if (t < s1) {
                                const float4 c = cand_pos<STRIDE, WIN>(cand, t);
                                const bool in_cell = cell_coord(c.x, sg.inv_cell) == cx &&
                                                     cell_coord(c.y, sg.inv_cell) == cy &&
                                                     cell_coord(c.z, sg.inv_cell) == cz;
                                tile[threadIdx.x] = make_float4(c.x, c.y, c.z, __uint_as_float(in_cell ? t : NO_BIRD));
                            }
                            __syncthreads();
                            const uint32_t cnt = min((uint32_t)BLOCK, s1 - t0);
                            if (need) {
                                for (uint32_t j = 0; j < cnt; ++j) {
                                    const float4 c = tile[j];
                                    const uint32_t ts = __float_as_uint(c.w);
                                    if (ts == NO_BIRD || ts == slot) continue;
                                    const f3 d = mk3(c.x - w.p.x, c.y - w.p.y, c.z - w.p.z);
                                    const float d2 = dot3(d, d);
                                    if (d2 > w.c2) continue;
                                    accumulate<STRIDE == 2>(a, f_i, w.r2, d, d2,
                                                            [&] { return cand_vel<STRIDE, WIN>(cand, ts); });
                                }
                            }
                            __syncthreads();
                        }
                    }
        } else if (active) {
            a = walk_sums<G, PF, STRIDE, WIN, MG>(p, cand, g.segs[si], g.n_slots, slot, w, f_i, (int32_t)lane_g,
                                                  use_occ, movers);
        }
        if (G == 1) __syncthreads();  // the union box is reset by the next iteration
        if (G > 1) reduce_group<G>(a);
        if (active && lane_g == 0) {
            const Decision d = finish(a, pos, speed, f_i, r);
            if (RES) {
                rec.s4[slot] = make_float4(d.force.x, d.force.y, d.force.z, d.radius);
                rec.p4[slot].w = rec.p4[slot].w + DU;
            } else {
                p.steer4[me] = make_float4(d.force.x, d.force.y, d.force.z, d.radius);
                p.pos4[me].w = p.pos4[me].w + DU;
            }
            p.nb[me] = d.n_all;
            p.cen[me] = d.centrality;
        }
    }
    if constexpr (RES && FLY) {
        // Fused flight: the chunk's decisions are in rec (visible to the whole
        // workgroup after the barrier); every bird of the chunk flies into
        // `flown`, so the walks of other workgroups keep reading the positions
        // of this step from rec. A separate instance: the flight's code would
        // otherwise cost the G = 1-2 walks a wave per SIMD.
        __syncthreads();
        for (uint32_t b0 = start; b0 < end; b0 += 2u * BLOCK)
            resident_flight2<MG>(g, rec, flown, w, b0 + threadIdx.x, b0 + BLOCK + threadIdx.x, end, start,
                                 (walk_flags & WALK_LAZY) != 0);
    }
    __syncthreads();  // due_slots and n_due are reused by the next chunk
    }
}

// Workgroups of one persistent decide launch: as many as the chip holds at
// once (occupancy of this instance at this LDS size, times the CUs), and no
// more than there are chunks. The occupancy is a property of the compiled
// kernel; it is looked up once per instance and chunk size.
template <int BLOCK, int G, bool RES, bool FLY, bool WIN, bool MG>
static uint32_t persistent_blocks(uint32_t chunk, uint32_t n_chunks, size_t lds) {
    // by log2(chunk / BLOCK) (1..16 birds per thread), without and with a window
    static int per_cu[2][5] = {{0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}};
    const uint32_t bpt = chunk / BLOCK;
    const int idx = bpt <= 1 ? 0 : bpt <= 2 ? 1 : bpt <= 4 ? 2 : bpt <= 8 ? 3 : 4;
    int *slot = &per_cu[lds > chunk * sizeof(uint32_t) ? 1 : 0][idx];
    int occ = __atomic_load_n(slot, __ATOMIC_RELAXED);
    if (occ <= 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_decide<BLOCK, G, RES, FLY, WIN, MG>, BLOCK, lds) != hipSuccess ||
            occ <= 0)
            occ = 1;
        __atomic_store_n(slot, occ, __ATOMIC_RELAXED);
    }
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    return min(n_chunks, (uint32_t)occ * (uint32_t)cus);
}

// With timing events, the launch itself records the kernel's start and end
// (hipExtLaunchKernelGGL), so the measured interval is the kernel alone.
template <int BLOCK, int G, bool RES, bool FLY, bool WIN, bool MG>
static void launch_decide_block(const Geometry &g, const Pools &p, const Records &rec, float time, uint32_t chunk,
                                int walk_flags, uint32_t *due_clear, uint32_t *work_clear, const ResWork &w,
                                const Records &flown, const DecideTiming &tm, hipStream_t s) {
    const uint32_t n_chunks = (g.n_slots + chunk - 1) / chunk;
    const size_t lds = decide_lds_bytes(chunk, walk_flags);
    const uint32_t blocks =
        (walk_flags & WALK_PERSIST) ? persistent_blocks<BLOCK, G, RES, FLY, WIN, MG>(chunk, n_chunks, lds) : n_chunks;
    if (tm.start)
        hipExtLaunchKernelGGL((k_decide<BLOCK, G, RES, FLY, WIN, MG>), dim3(blocks), dim3(BLOCK), lds, s, tm.start, tm.stop, 0u, g,
                              p, rec, time, chunk, walk_flags, due_clear, work_clear, w, flown);
    else
        hipLaunchKernelGGL((k_decide<BLOCK, G, RES, FLY, WIN, MG>), dim3(blocks), dim3(BLOCK), lds, s, g, p, rec, time, chunk,
                           walk_flags, due_clear, work_clear, w, flown);
}

template <int BLOCK, bool RES, bool FLY, bool WIN, bool MG>
static hipError_t launch_decide_group(const Geometry &g, const Pools &p, const Records &rec, float time,
                                      uint32_t chunk, uint32_t group, int walk_flags, uint32_t *due_clear,
                                      uint32_t *work_clear, const ResWork &w, const Records &flown,
                                      const DecideTiming &tm, hipStream_t s) {
#define FH_DECIDE_G(GV)                                                                                    \
    case GV:                                                                                               \
        launch_decide_block<BLOCK, GV, RES, FLY, WIN, MG>(g, p, rec, time, chunk, walk_flags, due_clear, work_clear, w, flown, tm, \
                                            s);                                                                    \
        break;
    switch (group) {
        FH_DECIDE_G(1)
        FH_DECIDE_G(2)
        FH_DECIDE_G(4)
        FH_DECIDE_G(8)
        FH_DECIDE_G(16)
        FH_DECIDE_G(64)
        default: return hipErrorInvalidValue;
    }
#undef FH_DECIDE_G
    return hipSuccess;
}

template <bool RES, bool FLY, bool WIN, bool MG = false>
static hipError_t launch_decide_any(const Geometry &g, const Pools &p, const Records &rec, float time,
                                    uint32_t block, uint32_t birds_per_thread, uint32_t group, int walk_flags,
                                    uint32_t *due_clear, uint32_t *work_clear, const ResWork &w,
                                    const Records &flown, const DecideTiming &tm, hipStream_t s) {
    if (g.n_slots == 0) return hipSuccess;
    const uint32_t chunk = block * birds_per_thread;
    auto by_block = [&](auto block_c) {
        constexpr int B = decltype(block_c)::value;
        return launch_decide_group<B, RES, FLY, WIN, MG>(g, p, rec, time, chunk, group, walk_flags, due_clear, work_clear, w, flown,
                                           tm, s);
    };
    hipError_t e;
    switch (block) {
        case 64: e = by_block(std::integral_constant<int, 64>{}); break;
        case 128: e = by_block(std::integral_constant<int, 128>{}); break;
        case 256: e = by_block(std::integral_constant<int, 256>{}); break;
        case 512: e = by_block(std::integral_constant<int, 512>{}); break;
        default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

// decide over sorted copies (resident = false) or resident records, with or
// without the fused flight, for instances with (WIN) or without a window.
template <bool WIN>
hipError_t decide_dispatch(const Geometry &g, const Pools &p, const Records &rec, float time, uint32_t block,
                           uint32_t birds_per_thread, uint32_t group, int walk_flags, uint32_t *due_clear,
                           uint32_t *work_clear, const ResWork &w, const Records &flown, const DecideTiming &tm,
                           hipStream_t s, bool resident) {
    if (!resident)
        return launch_decide_any<false, false, WIN>(g, p, Records{}, time, block, birds_per_thread, group,
                                                    walk_flags & ~WALK_FLIGHT, nullptr, nullptr, ResWork{}, Records{},
                                                    tm, s);
    if (walk_flags & WALK_FLIGHT)
        return launch_decide_any<true, true, WIN>(g, p, rec, 0.0f, block, birds_per_thread, group, walk_flags, due_clear,
                                                  work_clear, w, flown, tm, s);
    return launch_decide_any<true, false, WIN>(g, p, rec, 0.0f, block, birds_per_thread, group, walk_flags, due_clear,
                                               work_clear, w, flown, tm, s);
}

// decide over resident records with the fused flight whose walks also visit
// the mover grid (FH_FLAG_LAZY; instances without a window, decide_lazy.hip).
hipError_t decide_dispatch_movers(const Geometry &g, const Pools &p, const Records &rec, uint32_t block,
                                  uint32_t birds_per_thread, uint32_t group, int walk_flags, uint32_t *due_clear,
                                  uint32_t *work_clear, const ResWork &w, const Records &flown,
                                  const DecideTiming &tm, hipStream_t s);

extern template hipError_t decide_dispatch<false>(const Geometry &, const Pools &, const Records &, float, uint32_t,
                                                  uint32_t, uint32_t, int, uint32_t *, uint32_t *, const ResWork &,
                                                  const Records &, const DecideTiming &, hipStream_t, bool);
extern template hipError_t decide_dispatch<true>(const Geometry &, const Pools &, const Records &, float, uint32_t,
                                                 uint32_t, uint32_t, int, uint32_t *, uint32_t *, const ResWork &,
                                                 const Records &, const DecideTiming &, hipStream_t, bool);

}  // namespace fh
