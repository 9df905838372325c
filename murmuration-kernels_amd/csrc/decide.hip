// decide's public launchers and the instances without a window (the kernels
// are in decide_kernels.h; the instances with a window are built by
// decide_win.hip). See decide_kernels.h for the walk.
#include "decide_kernels.h"

namespace fh {

template hipError_t decide_dispatch<false>(const Geometry &, const Pools &, const Records &, float, uint32_t, uint32_t,
                                           uint32_t, int, uint32_t *, uint32_t *, const ResWork &, const Records &,
                                           const DecideTiming &, hipStream_t, bool);

// Debug walk for one bird (sorted copies): the ids the grid walk finds within
// r (walk 1) or 2r (walk 2), in walk order. One lane; used by the parity
// tests only. The bird's own slot is looked up in its bucket first.
__global__ void k_neighbours(Geometry g, Pools p, uint32_t bird, int walk, uint32_t *ids, uint32_t cap,
                             uint32_t *count) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const Seg &sg = g.segs[g.tile_seg ? g.tile_seg[bird / TILE] : 0u];
    const float4 q = p.pos4[bird];
    const float r = p.steer4[bird].w;
    const Walk w = make_walk(mk3(q.x, q.y, q.z), r, sg);
    const uint2 own = bucket_range(p.offsets, p.key[bird]);
    uint32_t self = 0xffffffffu;
    for (uint32_t t = own.x; t < min(own.y, g.n_slots); ++t)
        if (__float_as_uint(p.sbird[2u * t].w) == bird) self = t;
    uint32_t c = 0;
    walk_cells<1, 8, 2, false, false>(
        w, sg, p, Cand{p.sbird, p.sbird + 1}, g.n_slots, self, 0, true,
        [&](uint32_t t, f3, float d2) {
            if (walk == 1 && d2 > w.r2) return;
            if (c < cap) ids[c] = __float_as_uint(p.sbird[2u * t].w);
            ++c;
        },
        ResWork{}, [](uint32_t, f3, float) {});
    *count = c;
}

hipError_t launch_decide(const Geometry &g, const Pools &p, float time, uint32_t block,
                         uint32_t birds_per_thread, uint32_t group, int walk_flags, hipStream_t s,
                         const DecideTiming &tm) {
    if (walk_flags & WALK_WINDOW)
        return decide_dispatch<true>(g, p, Records{}, time, block, birds_per_thread, group, walk_flags, nullptr,
                                     nullptr, ResWork{}, Records{}, tm, s, false);
    return decide_dispatch<false>(g, p, Records{}, time, block, birds_per_thread, group, walk_flags, nullptr, nullptr,
                                  ResWork{}, Records{}, tm, s, false);
}

hipError_t launch_decide_resident(const Geometry &g, const Pools &p, const Records &rec, uint32_t block,
                                  uint32_t birds_per_thread, uint32_t group, int walk_flags, uint32_t *due_clear,
                                  uint32_t *work_clear, const ResWork &w, const Records &flown, hipStream_t s,
                                  const DecideTiming &tm) {
    if (walk_flags & WALK_MOVERS) {
        if (!(walk_flags & WALK_FLIGHT) || (walk_flags & WALK_WINDOW)) return hipErrorInvalidValue;
        return decide_dispatch_movers(g, p, rec, block, birds_per_thread, group, walk_flags, due_clear, work_clear, w,
                                      flown, tm, s);
    }
    if (walk_flags & WALK_WINDOW)
        return decide_dispatch<true>(g, p, rec, 0.0f, block, birds_per_thread, group, walk_flags, due_clear, work_clear,
                                     w, flown, tm, s, true);
    return decide_dispatch<false>(g, p, rec, 0.0f, block, birds_per_thread, group, walk_flags, due_clear, work_clear, w,
                                  flown, tm, s, true);
}

hipError_t launch_neighbours(const Geometry &g, const Pools &p, uint32_t bird, int walk, uint32_t *ids,
                             uint32_t cap, uint32_t *count, hipStream_t s) {
    hipLaunchKernelGGL(k_neighbours, dim3(1), dim3(64), 0, s, g, p, bird, walk, ids, cap, count);
    return hipGetLastError();
}

template <int BLOCK, bool RES, bool FLY>
static const void *decide_fn_group(uint32_t group) {
    switch (group) {
        case 1: return reinterpret_cast<const void *>(k_decide<BLOCK, 1, RES, FLY, false, false>);
        case 2: return reinterpret_cast<const void *>(k_decide<BLOCK, 2, RES, FLY, false, false>);
        case 4: return reinterpret_cast<const void *>(k_decide<BLOCK, 4, RES, FLY, false, false>);
        case 8: return reinterpret_cast<const void *>(k_decide<BLOCK, 8, RES, FLY, false, false>);
        case 16: return reinterpret_cast<const void *>(k_decide<BLOCK, 16, RES, FLY, false, false>);
        case 64: return reinterpret_cast<const void *>(k_decide<BLOCK, 64, RES, FLY, false, false>);
        default: return nullptr;
    }
}

template <bool RES, bool FLY>
static const void *decide_fn(uint32_t block, uint32_t group) {
    switch (block) {
        case 64: return decide_fn_group<64, RES, FLY>(group);
        case 128: return decide_fn_group<128, RES, FLY>(group);
        case 256: return decide_fn_group<256, RES, FLY>(group);
        case 512: return decide_fn_group<512, RES, FLY>(group);
        default: return nullptr;
    }
}

const void *decide_movers_fn(uint32_t block, uint32_t group);

// "decide" is the default instance (block 256, G 1, sorted copies);
// "decide:BLOCK:G:K" names the instance for that workgroup size and lane
// group over sorted copies (K = s), resident records (r), resident records
// with the fused flight (f), or those with the mover grid too (m).
hipError_t decide_kernel_attributes(const char *name, hipFuncAttributes *out) {
    if (!__builtin_strcmp(name, "decide"))
        return hipFuncGetAttributes(out, reinterpret_cast<const void *>(k_decide<256, 1, false, false, false, false>));
    unsigned block = 0, group = 0;
    char kind = 0;
    if (std::sscanf(name, "decide:%u:%u:%c", &block, &group, &kind) == 3) {
        const void *f = kind == 's'   ? decide_fn<false, false>(block, group)
                        : kind == 'r' ? decide_fn<true, false>(block, group)
                        : kind == 'f' ? decide_fn<true, true>(block, group)
                        : kind == 'm' ? decide_movers_fn(block, group)
                                      : nullptr;
        if (f) return hipFuncGetAttributes(out, f);
    }
    return hipErrorInvalidValue;
}

}  // namespace fh
