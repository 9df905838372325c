// Host helpers shared by the library's C ABI translation units: the
// thread-local error message, HIP error checks and the .flk record layout.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/flock_hip.h"

namespace fh {

fh_status set_error(fh_status st, const char *fmt, ...);
void clear_error();

// Records a HIP failure (with the call site) and returns FH_ERR_HIP.
fh_status hip_fail(hipError_t e, const char *what);

#define FH_HIP(call)                                                  \
    do {                                                              \
        hipError_t e_ = (call);                                       \
        if (e_ != hipSuccess) return ::fh::hip_fail(e_, #call);       \
    } while (0)

// Checks that device 0 is a usable gfx950 GPU (once per process).
fh_status require_device();

// Fills defaults into `in` (NULL = all defaults) and validates the result.
fh_status resolve_params(const fh_params *in, fh_params *out, const char *who);

// Buckets of a flock of n birds under resolved params: the override, or
// buckets_for(n) >> bucket_shift, at least 64.
inline uint32_t flock_buckets(uint32_t n, const fh_params &p) {
    if (p.buckets) return p.buckets;
    uint32_t b = 1;
    while ((uint64_t)b < (uint64_t)n * 2) b <<= 1;
    b = b < (1u << 16) ? (1u << 16) : (b > (1u << 22) ? (1u << 22) : b);
    b >>= p.bucket_shift;
    return b < 64u ? 64u : b;
}

// decide's walk switches (WALK_* of layout.h) for the FH_FLAG_* flags.
inline int walk_flags(uint32_t flags) {
    return ((flags & FH_FLAG_NO_TILE) ? 0 : 1) | ((flags & FH_FLAG_OCC) ? 2 : 0) |
           ((flags & FH_FLAG_PERSISTENT) ? 4 : 0) | ((flags & FH_FLAG_XCD) ? 8 : 0) |
           ((flags & FH_FLAG_WINDOW) ? 16 : 0) | ((flags & FH_FLAG_FLIGHT) ? 32 : 0);
}

// .flk record words (crates/flock-ref/src/state.rs:4-7).
enum Word { PX = 0, PY, PZ, VX, VY, VZ, ROLL, SPIN, RADIUS, SX, SY, SZ, NEXT, NB, CEN };

// Carves one device allocation into 256-byte aligned sub-buffers.
struct Arena {
    size_t used = 0;
    template <class T>
    size_t take(size_t count) {
        size_t at = used;
        used += (count * sizeof(T) + 255) & ~size_t(255);
        return at;
    }
};

}  // namespace fh
