// Host-only parts of the C ABI: errors, constants, .flk files and the input
// generators (no device work here).
#include <cmath>
#include <cstring>
#include <mutex>

#include "host_util.h"
#include "model.h"

namespace fh {

static thread_local std::string g_error;

fh_status set_error(fh_status st, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return st;
}

void clear_error() { g_error.clear(); }

fh_status hip_fail(hipError_t e, const char *what) {
    return set_error(e == hipErrorOutOfMemory ? FH_ERR_NOMEM : FH_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

fh_status require_device() {
    static std::once_flag once;
    static fh_status status = FH_OK;
    static std::string why;
    std::call_once(once, [] {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
            status = FH_ERR_NODEVICE;
            why = "no HIP device visible";
            return;
        }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
            status = FH_ERR_NODEVICE;
            why = "hipGetDeviceProperties failed";
            return;
        }
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
            status = FH_ERR_NODEVICE;
            why = std::string("device 0 is ") + prop.gcnArchName + ", this library is built for gfx950 only";
        }
    });
    if (status != FH_OK) return set_error(status, "%s", why.c_str());
    return FH_OK;
}

fh_status resolve_params(const fh_params *in, fh_params *out, const char *who) {
    fh_params prm;
    fh_params_default(&prm);
    if (in) {
        if (in->cell != 0.0f) prm.cell = in->cell;
        prm.buckets = in->buckets;
        if (in->decide_block) prm.decide_block = in->decide_block;
        if (in->birds_per_thread) prm.birds_per_thread = in->birds_per_thread;
        if (in->decide_group) prm.decide_group = in->decide_group;
        prm.bucket_shift = in->bucket_shift;
        prm.flags = in->flags;
    }
    if (prm.bucket_shift > 8) return set_error(FH_ERR_INVALID, "%s: bucket_shift %u > 8", who, prm.bucket_shift);
    if ((prm.flags & FH_FLAG_RESIDENT) && (prm.flags & FH_FLAG_CANONICAL))
        return set_error(FH_ERR_INVALID, "%s: FH_FLAG_RESIDENT does not combine with FH_FLAG_CANONICAL", who);
    if ((prm.flags & FH_FLAG_LAZY) &&
        (!(prm.flags & FH_FLAG_RESIDENT) || !(prm.flags & FH_FLAG_FLIGHT) || (prm.flags & FH_FLAG_WINDOW) ||
         (prm.flags & FH_FLAG_UNFUSED)))
        return set_error(FH_ERR_INVALID,
                         "%s: FH_FLAG_LAZY needs FH_FLAG_RESIDENT and FH_FLAG_FLIGHT, without FH_FLAG_WINDOW or "
                         "FH_FLAG_UNFUSED", who);
    if (!(prm.cell >= 0.5f) || !(prm.cell <= 1000.0f))
        return set_error(FH_ERR_INVALID, "%s: cell %g outside [0.5, 1000] m", who, prm.cell);
    if (prm.buckets && (prm.buckets & (prm.buckets - 1)))
        return set_error(FH_ERR_INVALID, "%s: buckets %u is not a power of two", who, prm.buckets);
    if (prm.buckets > (1u << 28)) return set_error(FH_ERR_INVALID, "%s: buckets %u > 2^28", who, prm.buckets);
    const uint32_t blk = prm.decide_block;
    if (blk != 64 && blk != 128 && blk != 256 && blk != 512)
        return set_error(FH_ERR_INVALID, "%s: decide_block %u not in {64,128,256,512}", who, blk);
    if (prm.birds_per_thread > 64 || blk * prm.birds_per_thread > 16384)
        return set_error(FH_ERR_INVALID, "%s: decide chunk %u x %u too large", who, blk, prm.birds_per_thread);
    const uint32_t grp = prm.decide_group;
    if (grp != 1 && grp != 2 && grp != 4 && grp != 8 && grp != 16 && grp != 64)
        return set_error(FH_ERR_INVALID, "%s: decide_group %u not in {1,2,4,8,16,64}", who, grp);
    *out = prm;
    return FH_OK;
}

// SplitMix64 (Steele, Lea, Flood 2014), as crates/flock-ref/src/state.rs:58-73.
struct SplitMix {
    uint64_t s;
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform in [0, 1) with 24 random bits
    float unit() { return (float)(next() >> 40) / (float)(1ull << 24); }
};

static inline float *rec(float *words, uint32_t i) { return words + (size_t)i * FH_WORDS; }

}  // namespace fh

using namespace fh;

extern "C" {

const char *fh_version(void) { return "flock_hip 0.1 (gfx950)"; }

const char *fh_last_error(void) { return g_error.c_str(); }

int fh_constants(float *out, int cap) {
    const float c[] = {DT,    DU,     V0,          MASS,     CL_CD,     G,         L0,        D0,
                       TAU,   TAU_ROLL, MAX_ROLL_RATE, MAX_BANK, V_MIN,   V_MAX,     N_C,       R_MAX,
                       RADIUS_LERP, R_H, SIGMA,       W_SEP,    W_ALIGN,   W_COH,     BLIND_COS, W_ROOST_V,
                       H_ROOST_V, Z0,   GRID_CELL,   RADIUS_EXP};
    const int k = (int)(sizeof c / sizeof c[0]);
    for (int i = 0; i < k && out && i < cap; ++i) out[i] = c[i];
    return k;
}

uint32_t fh_buckets_for(uint32_t n) { return buckets_for(n); }

fh_status fh_read_flk(const char *path, uint32_t *n, float *time, float *words, uint32_t cap) {
    clear_error();
    if (!path || !n || !time) return set_error(FH_ERR_INVALID, "fh_read_flk: NULL argument");
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return set_error(FH_ERR_IO, "%s: cannot open", path);
    uint32_t head[4];
    fh_status st = FH_OK;
    if (std::fread(head, 4, 4, fp) != 4) {
        st = set_error(FH_ERR_FLK, "%s: truncated header", path);
    } else if (std::memcmp(head, "FLK1", 4) != 0) {
        st = set_error(FH_ERR_FLK, "%s: not a FLK1 file", path);
    } else if (head[2] != FH_WORDS) {
        st = set_error(FH_ERR_FLK, "%s: %u words per bird, expected %d", path, head[2], FH_WORDS);
    } else {
        *n = head[1];
        std::memcpy(time, &head[3], 4);
        if (words) {
            if (head[1] > cap) {
                st = set_error(FH_ERR_INVALID, "%s: %u birds exceed the buffer of %u", path, head[1], cap);
            } else {
                const size_t want = (size_t)head[1] * FH_WORDS;
                if (std::fread(words, 4, want, fp) != want) st = set_error(FH_ERR_FLK, "%s: truncated", path);
            }
        }
    }
    std::fclose(fp);
    return st;
}

fh_status fh_write_flk(const char *path, uint32_t n, float time, const float *words) {
    clear_error();
    if (!path || (n && !words)) return set_error(FH_ERR_INVALID, "fh_write_flk: NULL argument");
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return set_error(FH_ERR_IO, "%s: cannot create", path);
    uint32_t head[4] = {0x314B4C46u, n, (uint32_t)FH_WORDS, 0u};
    std::memcpy(&head[3], &time, 4);
    bool ok = std::fwrite(head, 4, 4, fp) == 4 &&
              std::fwrite(words, 4, (size_t)n * FH_WORDS, fp) == (size_t)n * FH_WORDS;
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? FH_OK : set_error(FH_ERR_IO, "%s: write failed", path);
}

// Flock::scatter, crates/flock-ref/src/state.rs:79-101: cube of density 0.5
// about (0, 0, Z0), heading +x with 0.05 jitter, draws in the order
// offset xyz, jitter xyz, next_decision.
fh_status fh_gen_reference(uint32_t n, uint64_t seed, float *words) {
    clear_error();
    if (n && !words) return set_error(FH_ERR_INVALID, "fh_gen_reference: NULL words");
    const float side = cbrtf((float)n / 0.5f);
    SplitMix rng{seed};
    for (uint32_t i = 0; i < n; ++i) {
        const float ox = (rng.unit() - 0.5f) * side;
        const float oy = (rng.unit() - 0.5f) * side;
        const float oz = (rng.unit() - 0.5f) * side;
        const float jx = (rng.unit() - 0.5f) * 0.05f;
        const float jy = (rng.unit() - 0.5f) * 0.05f;
        const float jz = (rng.unit() - 0.5f) * 0.05f;
        // (X + jitter).normalize() * V0, glam: v * (1 / |v|)
        const f3 h = mk3(1.0f + jx, 0.0f + jy, 0.0f + jz);
        const f3 v = (h * (1.0f / sqrtf(dot3(h, h)))) * V0;
        const float next = rng.unit() * DU;
        float *r = rec(words, i);
        std::memset(r, 0, FH_WORDS * sizeof(float));
        r[PX] = 0.0f + ox;
        r[PY] = 0.0f + oy;
        r[PZ] = Z0 + oz;
        r[VX] = v.x;
        r[VY] = v.y;
        r[VZ] = v.z;
        r[RADIUS] = START_RADIUS;
        r[NEXT] = next;
    }
    return FH_OK;
}

// Per-bird streams so a flock of any size is reproducible bird by bird:
// bird i draws from SplitMix64 seeded with seed + i * 0xD1B54A32D192ED03.
static inline SplitMix bird_stream(uint64_t seed, uint32_t i) {
    return SplitMix{seed + (uint64_t)i * 0xD1B54A32D192ED03ull};
}

static inline f3 unit_sphere(SplitMix &rng) {
    const float cz = 2.0f * rng.unit() - 1.0f;
    const float phi = 6.28318530717958647692f * rng.unit();
    const float sxy = sqrtf(fmaxf(0.0f, 1.0f - cz * cz));
    return mk3(sxy * cosf(phi), sxy * sinf(phi), cz);
}

// Uniform cube, isotropic headings, speed 10 +- 1 m/s. Draw order per bird:
// x, y, z, cos(theta), phi, speed noise, next_decision.
fh_status fh_gen_cube(uint32_t n, float side, uint64_t seed, float *words) {
    clear_error();
    if (n && !words) return set_error(FH_ERR_INVALID, "fh_gen_cube: NULL words");
    if (!(side > 0.0f)) return set_error(FH_ERR_INVALID, "fh_gen_cube: side must be > 0");
    for (uint32_t i = 0; i < n; ++i) {
        SplitMix rng = bird_stream(seed, i);
        const float x = (rng.unit() - 0.5f) * side;
        const float y = (rng.unit() - 0.5f) * side;
        const float z = (rng.unit() - 0.5f) * side;
        const f3 dir = unit_sphere(rng);
        const float speed = V0 + (2.0f * rng.unit() - 1.0f);
        const float next = rng.unit() * DU;
        float *r = rec(words, i);
        std::memset(r, 0, FH_WORDS * sizeof(float));
        r[PX] = x;
        r[PY] = y;
        r[PZ] = Z0 + z;
        r[VX] = dir.x * speed;
        r[VY] = dir.y * speed;
        r[VZ] = dir.z * speed;
        r[RADIUS] = START_RADIUS;
        r[NEXT] = next;
    }
    return FH_OK;
}

// Gaussian clumps. Clump c's centre and heading come from the stream of
// seed ^ 0xC1A3B5 (index c); bird i picks its clump, then draws a Gaussian
// offset (Box-Muller) and a heading in the 10 degree cap around the clump's.
fh_status fh_gen_clumps(uint32_t n, uint32_t clumps, float sigma, float cube, uint64_t seed, float *words) {
    clear_error();
    if (n && !words) return set_error(FH_ERR_INVALID, "fh_gen_clumps: NULL words");
    if (clumps == 0 || clumps > 4096 || !(sigma >= 0.0f) || !(cube >= 0.0f))
        return set_error(FH_ERR_INVALID, "fh_gen_clumps: bad clumps/sigma/cube");
    f3 centre[4096], heading[4096];
    for (uint32_t c = 0; c < clumps; ++c) {
        SplitMix rng = bird_stream(seed ^ 0xC1A3B5ull, c);
        centre[c] = mk3((rng.unit() - 0.5f) * cube, (rng.unit() - 0.5f) * cube, Z0 + (rng.unit() - 0.5f) * cube);
        heading[c] = unit_sphere(rng);
    }
    const float cos_cap = cosf(10.0f * PI_F32 / 180.0f);
    const float two_pi = 6.28318530717958647692f;
    for (uint32_t i = 0; i < n; ++i) {
        SplitMix rng = bird_stream(seed, i);
        uint32_t c = (uint32_t)(rng.unit() * (float)clumps);
        if (c >= clumps) c = clumps - 1;
        float g[4];
        for (int k = 0; k < 4; k += 2) {
            const float u1 = 1.0f - rng.unit();  // (0, 1]
            const float u2 = rng.unit();
            const float rad = sqrtf(-2.0f * logf(u1));
            g[k] = rad * cosf(two_pi * u2);
            g[k + 1] = rad * sinf(two_pi * u2);
        }
        const f3 h = heading[c];
        const f3 a = fabsf(h.x) < 0.9f ? mk3(1.f, 0.f, 0.f) : mk3(0.f, 1.f, 0.f);
        f3 e1 = cross3(a, h);
        e1 = e1 * (1.0f / sqrtf(dot3(e1, e1)));
        const f3 e2 = cross3(h, e1);
        const float ct = cos_cap + (1.0f - cos_cap) * rng.unit();
        const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
        const float psi = two_pi * rng.unit();
        const f3 dir = h * ct + (e1 * cosf(psi) + e2 * sinf(psi)) * st;
        const float next = rng.unit() * DU;
        float *r = rec(words, i);
        std::memset(r, 0, FH_WORDS * sizeof(float));
        r[PX] = centre[c].x + sigma * g[0];
        r[PY] = centre[c].y + sigma * g[1];
        r[PZ] = centre[c].z + sigma * g[2];
        r[VX] = dir.x * V0;
        r[VY] = dir.y * V0;
        r[VZ] = dir.z * V0;
        r[RADIUS] = START_RADIUS;
        r[NEXT] = next;
    }
    return FH_OK;
}

fh_status fh_gen_request(uint32_t n, uint64_t seed, float birds_per_cell, float cell, float *words) {
    clear_error();
    if (!(birds_per_cell > 0.0f) || !(cell > 0.0f))
        return set_error(FH_ERR_INVALID, "fh_gen_request: birds_per_cell and cell must be > 0");
    const float side = cell * cbrtf((float)(n ? n : 1) / birds_per_cell);
    return fh_gen_cube(n, side, seed, words);
}

fh_status fh_params_default(fh_params *p) {
    clear_error();
    if (!p) return set_error(FH_ERR_INVALID, "fh_params_default: NULL");
    std::memset(p, 0, sizeof *p);
    p->cell = GRID_CELL;
    p->buckets = 0;
    p->decide_block = 256;
    p->birds_per_thread = 8;
    p->flags = 0;
    p->decide_group = 1;
    return FH_OK;
}

fh_status fh_device_info(char *name, int cap, int *cus, size_t *hbm) {
    clear_error();
    fh_status st = require_device();
    if (st) return st;
    hipDeviceProp_t prop;
    FH_HIP(hipGetDeviceProperties(&prop, 0));
    if (name && cap > 0) std::snprintf(name, (size_t)cap, "%s (%s)", prop.name, prop.gcnArchName);
    if (cus) *cus = prop.multiProcessorCount;
    if (hbm) *hbm = prop.totalGlobalMem;
    return FH_OK;
}

}  // extern "C"
