// decide's instances whose walks also visit the mover grid (FH_FLAG_LAZY:
// resident records, fused flight, no window), built apart from the others so
// the three sets compile in parallel.
#include "decide_kernels.h"

namespace fh {

hipError_t decide_dispatch_movers(const Geometry &g, const Pools &p, const Records &rec, uint32_t block,
                                  uint32_t birds_per_thread, uint32_t group, int walk_flags, uint32_t *due_clear,
                                  uint32_t *work_clear, const ResWork &w, const Records &flown,
                                  const DecideTiming &tm, hipStream_t s) {
    return launch_decide_any<true, true, false, true>(g, p, rec, 0.0f, block, birds_per_thread, group, walk_flags,
                                                      due_clear, work_clear, w, flown, tm, s);
}

template <int BLOCK>
static const void *movers_fn_group(uint32_t group) {
    switch (group) {
        case 1: return reinterpret_cast<const void *>(k_decide<BLOCK, 1, true, true, false, true>);
        case 2: return reinterpret_cast<const void *>(k_decide<BLOCK, 2, true, true, false, true>);
        case 4: return reinterpret_cast<const void *>(k_decide<BLOCK, 4, true, true, false, true>);
        case 8: return reinterpret_cast<const void *>(k_decide<BLOCK, 8, true, true, false, true>);
        case 16: return reinterpret_cast<const void *>(k_decide<BLOCK, 16, true, true, false, true>);
        case 64: return reinterpret_cast<const void *>(k_decide<BLOCK, 64, true, true, false, true>);
        default: return nullptr;
    }
}

// The instance of a workgroup size and lane group, for resource reports.
const void *decide_movers_fn(uint32_t block, uint32_t group) {
    switch (block) {
        case 64: return movers_fn_group<64>(group);
        case 128: return movers_fn_group<128>(group);
        case 256: return movers_fn_group<256>(group);
        case 512: return movers_fn_group<512>(group);
        default: return nullptr;
    }
}

}  // namespace fh
