// integrate for gfx950: banked flight for every bird (crates/flock-ref/src/
// model.rs:149-176, cuda/flock.cu:293-328), optionally fused with the next
// step's grid_histogram so the new position is hashed while it is still in
// registers instead of being read back by a separate pass.
#include "layout.h"
#include "model.h"

namespace fh {

template <bool WAVE_AGG>
__device__ __forceinline__ uint32_t count_key_i(uint32_t *counts, uint32_t k, bool active) {
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

// Banked flight of one bird (model.rs:149-176): new (pos, next), (vel, roll)
// and the roll rate, from its (pos, next), (vel, roll) and (steering, radius).
__device__ __forceinline__ void integrate_bird(float4 &p4, float4 &v4, const float4 s4, float &rate_out) {
    f3 vel = mk3(v4.x, v4.y, v4.z);
    const float roll = v4.w;
    const f3 steering = mk3(s4.x, s4.y, s4.z);

    const f3 fwd = forward_of(vel);
    const f3 side = side_of(fwd);
    const float speed = fmaxf(sqrtf(dot3(vel, vel)), 1e-6f);

    const float a_lat = dot3(steering, side) / MASS;
    const float target = clampf_(atanf(a_lat / G), -MAX_BANK, MAX_BANK);
    const float rate = clampf_((target - roll) / TAU_ROLL, -MAX_ROLL_RATE, MAX_ROLL_RATE);
    const float beta = clampf_(roll + rate * DT, -MAX_BANK, MAX_BANK);

    // body_up (model.rs:34-38): (side x f) cos b + side sin b
    const f3 up0 = cross3(side, fwd);
    const float cb = cosf(beta);
    const f3 up = up0 * cb + side * sinf(beta);
    const float cos_beta = fmaxf(cb, 0.2f);
    const float sr = speed / V0;
    const float lift = sr * sr * L0 / cos_beta;
    const float drag = sr * sr * D0;
    const f3 flight = ((up * lift - fwd * drag) + fwd * D0) - (mk3(0.f, 0.f, 1.f) * MASS) * G;

    const f3 control = steering - side * dot3(steering, side);
    vel = vel + ((control + flight) / MASS) * DT;
    const float sp = sqrtf(dot3(vel, vel));
    if (sp > V_MAX)
        vel = vel * (V_MAX / sp);
    else if (sp < V_MIN && sp > 1e-6f)
        vel = vel * (V_MIN / sp);
    const f3 pos = mk3(p4.x, p4.y, p4.z) + vel * DT;

    p4 = make_float4(pos.x, pos.y, pos.z, p4.w);
    v4 = make_float4(vel.x, vel.y, vel.z, beta);
    rate_out = rate;
}

template <bool HASH, bool WAVE_AGG>
__global__ __launch_bounds__(TILE) void k_integrate(Geometry g, Pools p) {
    const uint32_t tile = g.tile_begin + blockIdx.x;
    const uint32_t si = g.tile_seg ? g.tile_seg[tile] : 0u;
    if (si == NO_SEG) return;
    const Seg &sg = g.segs[si];
    const uint32_t bird = tile * TILE + threadIdx.x;
    const bool active = bird < g.n_birds && bird - sg.bird_base < sg.n;
    uint32_t k = 0;
    if (active) {
        float4 p4 = p.pos4[bird];
        float4 v4 = p.vel4[bird];
        float rate;
        integrate_bird(p4, v4, p.steer4[bird], rate);
        p.pos4[bird] = p4;
        p.vel4[bird] = v4;
        p.spin[bird] = rate;
        if (HASH) {
            k = bucket_of_cell(sg, cell_coord(p4.x, sg.inv_cell), cell_coord(p4.y, sg.inv_cell),
                               cell_coord(p4.z, sg.inv_cell));
            p.key[bird] = k;
        }
    }
    if (HASH) {
        const uint32_t r = count_key_i<WAVE_AGG>(p.counts, k, active);
        if (active) p.rank[bird] = r;
    }
}

// Banked flight over resident records (slot order, one flock). A bird whose
// new cell falls in another bucket is a mover: it counts itself out of its
// old bucket and into its new one (keeping its arrival rank), and out of and
// into the scan tiles' totals when the tile changes; every slot records its
// new bucket and its mover bit. Workgroup 0 commits the step's clock (segment
// time += DT, the time decide used).
__global__ __launch_bounds__(TILE) void k_r_integrate(Geometry g, Records rec, ResWork w, Seg *seg) {
    const uint32_t s = blockIdx.x * TILE + threadIdx.x;
    const bool active = s < g.n_slots;
    const Seg sg = g.segs[0];
    bool mover = false;
    if (active) {
        float4 p4 = rec.p4[s];
        float4 v4 = rec.v4[s];
        float rate;
        integrate_bird(p4, v4, rec.s4[s], rate);
        rec.p4[s] = p4;
        rec.v4[s] = v4;
        rec.spin[s] = rate;
        const uint32_t kn = bucket_of_cell(sg, cell_coord(p4.x, sg.inv_cell), cell_coord(p4.y, sg.inv_cell),
                                           cell_coord(p4.z, sg.inv_cell));
        const uint32_t ko = rec.key[s];
        w.knew[s] = kn;
        if (kn != ko) {
            mover = true;
            w.arank[s] = atomicAdd(&w.arr[kn], 1u);
            atomicAdd(&w.dep[ko], 1u);
            if (kn / SCAN_TILE != ko / SCAN_TILE) {
                atomicAdd(&w.tsum[kn / SCAN_TILE], 1u);
                atomicSub(&w.tsum[ko / SCAN_TILE], 1u);
            }
        }
    }
    const uint64_t bits = __ballot(mover);
    const uint32_t lane = threadIdx.x & 63u;
    if (lane == 0u) w.mbits[s >> 5] = (uint32_t)bits;  // s: first slot of this wave
    if (lane == 32u) w.mbits[s >> 5] = (uint32_t)(bits >> 32);
    if (blockIdx.x == 0 && threadIdx.x == 0) seg->time = sg.time + DT;
}

hipError_t launch_resident_integrate(const Geometry &g, const Pools &, const Records &rec, const ResWork &w,
                                     Seg *seg, hipStream_t s) {
    if (g.n_slots == 0) return hipSuccess;
    hipLaunchKernelGGL(k_r_integrate, dim3((g.n_slots + TILE - 1) / TILE), dim3(TILE), 0, s, g, rec, w, seg);
    return hipGetLastError();
}

hipError_t launch_integrate(const Geometry &g, const Pools &p, bool hash, bool wave_agg, hipStream_t s) {
    if (g.n_birds == 0) return hipSuccess;
    const dim3 grid(g.tile_count ? g.tile_count : (g.n_birds + TILE - 1) / TILE), block(TILE);
    if (!hash)
        hipLaunchKernelGGL((k_integrate<false, false>), grid, block, 0, s, g, p);
    else if (wave_agg)
        hipLaunchKernelGGL((k_integrate<true, true>), grid, block, 0, s, g, p);
    else
        hipLaunchKernelGGL((k_integrate<true, false>), grid, block, 0, s, g, p);
    return hipGetLastError();
}

hipError_t integrate_kernel_attributes(const char *name, hipFuncAttributes *out) {
    if (!__builtin_strcmp(name, "integrate"))
        return hipFuncGetAttributes(out, reinterpret_cast<const void *>(k_integrate<true, false>));
    return hipErrorInvalidValue;
}

}  // namespace fh
