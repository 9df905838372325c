// integrate for gfx950: banked flight for every bird (crates/flock-ref/src/
// model.rs:149-176, cuda/flock.cu:293-328), optionally fused with the next
// step's grid_histogram so the new position is hashed while it is still in
// registers instead of being read back by a separate pass.
#include "flight.h"
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

// Banked flight over resident records (slot order), in place; see
// resident_flight. Every workgroup reads its flocks' step budgets here, so
// the clocks and budgets are committed by the rebuild scan that follows, never
// inside this launch.
__global__ __launch_bounds__(TILE) void k_r_integrate(Geometry g, Records rec, ResWork w) {
    const uint32_t s = blockIdx.x * TILE + threadIdx.x;
    resident_flight(g, rec, rec, w, s, s < g.n_slots);
}

hipError_t launch_resident_integrate(const Geometry &g, const Pools &, const Records &rec, const ResWork &w,
                                     hipStream_t s) {
    if (g.n_slots == 0) return hipSuccess;
    hipLaunchKernelGGL(k_r_integrate, dim3((g.n_slots + TILE - 1) / TILE), dim3(TILE), 0, s, g, rec, w);
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
