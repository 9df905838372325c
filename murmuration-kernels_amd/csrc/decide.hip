// decide for gfx950: each workgroup takes a run of sorted slots, compacts the
// birds that are due into LDS, and gives every lane one due bird. Birds are
// visited in slot order, so lanes of a wave that share a cell walk the same
// buckets.
//
// Mirrors crates/flock-ref/src/model.rs:53-145 over the grid of
// cuda/flock.cu:190-290. The reference walks twice, once over the box of
// radius r (sphere test r^2) and once over the box of 2r (sphere test 4r^2).
// Here one walk covers the 2r box: every cell of the r walk is a cell of the
// 2r walk (floor is monotone and r^2 <= 4r^2) and z-y-x order restricted to
// the r box is the r walk's own order, so each accumulator receives the same
// candidates in the same order as in the reference; walk-1 membership is the
// oracle's own test d^2 <= r^2.
#include "layout.h"
#include "model.h"

namespace fh {

struct Decision {
    f3 force;
    float radius;
    uint32_t n_all;
    float centrality;
};

// Per-bird walk. `visit(slot)` sees each candidate slot that passes the
// self and exact-cell checks and d^2 <= 4r^2, in walk order.
struct Walk {
    f3 p;
    float r, r2, c2;
    int32_t lo[3], hi[3];
    bool ok;
};

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
    // and is not walked (the bird then decides with no neighbours).
    const int64_t max_extent = (int64_t)(4.0f * R_MAX * sg.inv_cell) + 3;
    w.ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        w.ok = w.ok && (int64_t)w.hi[a] - (int64_t)w.lo[a] <= max_extent && w.hi[a] >= w.lo[a] &&
               w.hi[a] < 2147483647;
    return w;
}

// Cells of the box are visited in z-y-x order (x fastest), as the nested
// loops of cuda/flock.cu:209-211. They are taken PREFETCH at a time: the
// bucket ranges of all PREFETCH cells are loaded together, then the first
// candidate of each non-empty bucket, so one lane has up to 2 * PREFETCH
// independent loads in flight instead of a chain of dependent ones. The visit
// order of candidates is unchanged.
constexpr int PREFETCH = 8;

template <class Visit>
__device__ __forceinline__ void walk_cells(const Walk &w, const Seg &sg, const Pools &p, uint32_t n_slots,
                                           uint32_t me, Visit &&visit) {
    if (!w.ok) return;
    const int32_t nx = w.hi[0] - w.lo[0] + 1, ny = w.hi[1] - w.lo[1] + 1, nz = w.hi[2] - w.lo[2] + 1;
    const int32_t ncell = nx * ny * nz;
    int32_t cx0 = w.lo[0], cy0 = w.lo[1], cz0 = w.lo[2];  // cell of index `base`
    for (int32_t base = 0; base < ncell; base += PREFETCH) {
        int32_t cx[PREFETCH], cy[PREFETCH], cz[PREFETCH];
        uint32_t s0[PREFETCH], s1[PREFETCH];
        {
            int32_t x = cx0, y = cy0, z = cz0;
#pragma unroll
            for (int k = 0; k < PREFETCH; ++k) {
                cx[k] = x;
                cy[k] = y;
                cz[k] = z;
                s0[k] = 0u;
                s1[k] = 0u;
                if (base + k < ncell && cell_meets_sphere(x, y, z, sg.cell, w.p, w.c2)) {
                    const uint32_t b = sg.bucket_base + (bucket_hash(x, y, z) & sg.mask);
                    s0[k] = p.offsets[b];
                    s1[k] = p.offsets[b + 1];
                }
                if (++x > w.hi[0]) {
                    x = w.lo[0];
                    if (++y > w.hi[1]) {
                        y = w.lo[1];
                        ++z;
                    }
                }
            }
            cx0 = x;
            cy0 = y;
            cz0 = z;
        }
        float4 first[PREFETCH];
#pragma unroll
        for (int k = 0; k < PREFETCH; ++k) {
            s1[k] = min(s1[k], n_slots);
            first[k] = s0[k] < s1[k] ? p.spos[s0[k]] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int k = 0; k < PREFETCH; ++k) {
            for (uint32_t t = s0[k]; t < s1[k]; ++t) {
                const float4 c = t == s0[k] ? first[k] : p.spos[t];
                if (__float_as_uint(c.w) == me) continue;
                if (cell_coord(c.x, sg.inv_cell) != cx[k] || cell_coord(c.y, sg.inv_cell) != cy[k] ||
                    cell_coord(c.z, sg.inv_cell) != cz[k])
                    continue;
                const f3 d = mk3(c.x - w.p.x, c.y - w.p.y, c.z - w.p.z);
                const float d2 = dot3(d, d);
                if (d2 > w.c2) continue;
                visit(t, d, d2);
            }
        }
    }
}

__device__ __forceinline__ Decision decide_bird(const Pools &p, const Seg &sg, uint32_t n_slots, uint32_t me,
                                                f3 pos, f3 vel, float r) {
    const f3 f_i = forward_of(vel);
    f3 sep = mk3(0.f, 0.f, 0.f), coh = sep, align = sep, csum = sep;
    uint32_t n_all = 0, n_star = 0, n_align = 0, n_g = 0;
    const Walk w = make_walk(pos, r, sg);
    walk_cells(w, sg, p, n_slots, me, [&](uint32_t t, f3 d, float d2) {
        const float dist = sqrtf(d2);
        const f3 dhat = d / fmaxf(dist, 1e-6f);
        // walk 2 (model.rs:104-115): csum += d / max(sqrt(d2), 1e-6)
        csum = csum + dhat;
        ++n_g;
        if (d2 > w.r2) return;
        // walk 1 (model.rs:67-93)
        ++n_all;
        const float g = dist <= R_H ? 1.0f : expf(-((dist - R_H) * (dist - R_H)) / SEP_DENOM);
        sep = sep - dhat * g;
        if (dist > R_H) {
            coh = coh + dhat;
            ++n_star;
        }
        if (dot3(f_i, dhat) >= BLIND_COS) {
            const float4 vj = p.svel[t];
            align = align + forward_of(mk3(vj.x, vj.y, vj.z));
            ++n_align;
        }
    });
    // combine (model.rs:95-137, cuda/flock.cu:242-283)
    f3 force = mk3(0.f, 0.f, 0.f);
    if (n_all > 0) force = force + sep * (W_SEP / (float)n_all);
    const float centrality = n_g > 0 ? sqrtf(dot3(csum, csum)) / (float)n_g : 0.0f;
    if (n_star > 0) force = force + normalize_or_zero(coh) * (W_COH * centrality);
    const f3 a = n_align > 0 ? align / (float)n_align - f_i : mk3(0.f, 0.f, 0.f);
    force = force + a * W_ALIGN;
    force = force + mk3(0.f, 0.f, 1.f) * (W_ROOST_V * clampf_((Z0 - pos.z) / H_ROOST_V, -1.0f, 1.0f));
    const float speed = sqrtf(dot3(vel, vel));
    force = force + f_i * ((MASS / TAU) * (V0 - speed));
    Decision dcs;
    dcs.force = force;
    dcs.radius = radius_step(r, n_all);
    dcs.n_all = n_all;
    dcs.centrality = centrality;
    return dcs;
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_decide(Geometry g, Pools p, float time, uint32_t chunk) {
    extern __shared__ uint32_t due_slots[];
    __shared__ uint32_t n_due;
    if (threadIdx.x == 0) n_due = 0;
    __syncthreads();
    const uint32_t start = blockIdx.x * chunk;
    const uint32_t end = min(start + chunk, g.n_slots);
    for (uint32_t s = start + threadIdx.x; s < end; s += BLOCK)
        if (time >= p.snext[s]) due_slots[atomicAdd(&n_due, 1u)] = s;
    __syncthreads();
    const uint32_t nd = n_due;
    if (threadIdx.x == 0 && nd) atomicAdd(p.due_count, nd);
    for (uint32_t k = threadIdx.x; k < nd; k += BLOCK) {
        const uint32_t s = due_slots[k];
        const float4 q = p.spos[s];
        const uint32_t me = __float_as_uint(q.w);
        if (me >= g.n_birds) {
            atomicOr(p.err, ERR_SLOT);
            continue;
        }
        const Seg &sg = g.segs[g.tile_seg ? g.tile_seg[me / TILE] : 0u];
        const float4 v = p.svel[s];
        const float r = p.steer4[me].w;
        const Decision d = decide_bird(p, sg, g.n_slots, me, mk3(q.x, q.y, q.z), mk3(v.x, v.y, v.z), r);
        p.steer4[me] = make_float4(d.force.x, d.force.y, d.force.z, d.radius);
        p.pos4[me].w = p.pos4[me].w + DU;
        p.nb[me] = d.n_all;
        p.cen[me] = d.centrality;
    }
}

// Debug walk for one bird: the ids the grid walk finds within r (walk 1) or
// 2r (walk 2), in walk order. One lane; used by the parity tests only.
__global__ void k_neighbours(Geometry g, Pools p, uint32_t bird, int walk, uint32_t *ids, uint32_t cap,
                             uint32_t *count) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const Seg &sg = g.segs[g.tile_seg ? g.tile_seg[bird / TILE] : 0u];
    const float4 q = p.pos4[bird];
    const float r = p.steer4[bird].w;
    const Walk w = make_walk(mk3(q.x, q.y, q.z), r, sg);
    uint32_t c = 0;
    walk_cells(w, sg, p, g.n_slots, bird, [&](uint32_t t, f3, float d2) {
        if (walk == 1 && d2 > w.r2) return;
        if (c < cap) ids[c] = __float_as_uint(p.spos[t].w);
        ++c;
    });
    *count = c;
}

template <int BLOCK>
static void launch_decide_block(const Geometry &g, const Pools &p, float time, uint32_t chunk, hipStream_t s) {
    const uint32_t blocks = (g.n_slots + chunk - 1) / chunk;
    hipLaunchKernelGGL(k_decide<BLOCK>, dim3(blocks), dim3(BLOCK), chunk * sizeof(uint32_t), s, g, p, time,
                       chunk);
}

hipError_t launch_decide(const Geometry &g, const Pools &p, float time, uint32_t block,
                         uint32_t birds_per_thread, hipStream_t s) {
    if (g.n_slots == 0) return hipSuccess;
    const uint32_t chunk = block * birds_per_thread;
    switch (block) {
        case 64: launch_decide_block<64>(g, p, time, chunk, s); break;
        case 128: launch_decide_block<128>(g, p, time, chunk, s); break;
        case 256: launch_decide_block<256>(g, p, time, chunk, s); break;
        case 512: launch_decide_block<512>(g, p, time, chunk, s); break;
        case 1024: launch_decide_block<1024>(g, p, time, chunk, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_neighbours(const Geometry &g, const Pools &p, uint32_t bird, int walk, uint32_t *ids,
                             uint32_t cap, uint32_t *count, hipStream_t s) {
    hipLaunchKernelGGL(k_neighbours, dim3(1), dim3(64), 0, s, g, p, bird, walk, ids, cap, count);
    return hipGetLastError();
}

hipError_t decide_kernel_attributes(const char *name, hipFuncAttributes *out) {
    if (!__builtin_strcmp(name, "decide"))
        return hipFuncGetAttributes(out, reinterpret_cast<const void *>(k_decide<256>));
    return hipErrorInvalidValue;
}

}  // namespace fh
