// decide's instances that stage an LDS window of positions and velocities
// (FH_FLAG_WINDOW), built apart from the others so both compile in parallel.
#include "decide_kernels.h"

namespace fh {

template hipError_t decide_dispatch<true>(const Geometry &, const Pools &, const Records &, float, uint32_t, uint32_t,
                                          uint32_t, int, uint32_t *, uint32_t *, const ResWork &, const Records &,
                                          const DecideTiming &, hipStream_t, bool);

}  // namespace fh
