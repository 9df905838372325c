// Flock model constants and per-bird maths shared by the gfx950 kernels and
// the library's host code.
//
// Constants restate crates/flock-ref/src/params.rs:6-58 (the CUDA step keeps
// a second copy in cuda/params.cuh:5-33); fh_constants() exports them so a
// test can compare them with the oracle's copy word for word.
//
// Every expression keeps the reference's operation order
// (crates/flock-ref/src/model.rs; cuda/flock.cu:36-91, 182-186), and the
// library is compiled with -ffp-contract=off and correctly rounded
// division/sqrt, so the only differences from the CPU oracle come from the
// device expf/powf/atanf/sinf/cosf.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define FH_HD __host__ __device__ __forceinline__

namespace fh {

constexpr float DT = 0.005f;
constexpr float DU = 0.05f;
constexpr float V0 = 10.0f;
constexpr float MASS = 0.08f;
constexpr float CL_CD = 3.3f;
constexpr float G = 9.81f;
constexpr float L0 = MASS * G;
constexpr float D0 = L0 / CL_CD;
constexpr float TAU = 1.0f;
constexpr float TAU_ROLL = 0.15f;
constexpr float MAX_ROLL_RATE = 6.0f;
constexpr float PI_F32 = 3.14159265358979323846264338327950288f;
constexpr float MAX_BANK = 80.0f * PI_F32 / 180.0f;
constexpr float V_MIN = 0.5f * V0;
constexpr float V_MAX = 2.0f * V0;
constexpr float N_C = 6.5f;
constexpr float R_MAX = 100.0f;
constexpr float RADIUS_LERP = 0.1f;
constexpr float R_H = 0.2f;
constexpr float SIGMA = 1.37f;
constexpr float W_SEP = 1.0f;
constexpr float W_ALIGN = 0.5f;
constexpr float W_COH = 0.65f;
constexpr float BLIND_COS = -0.707106781186547524400844362104849039f;
constexpr float W_ROOST_V = 0.2f;
constexpr float H_ROOST_V = 40.0f;
constexpr float Z0 = 100.0f;
constexpr float GRID_CELL = 4.0f;
constexpr float RADIUS_EXP = RADIUS_LERP / 3.0f;
constexpr float SEP_DENOM = 2.0f * SIGMA * SIGMA;
constexpr float START_RADIUS = 1.5f;  // state.rs:54

struct f3 {
    float x, y, z;
};
FH_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
FH_HD f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
FH_HD f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
FH_HD f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
FH_HD f3 operator/(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
// glam Vec3::dot: (x*x') + (y*y') + (z*z'), left to right.
FH_HD float dot3(f3 a, f3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
FH_HD f3 cross3(f3 a, f3 b) {
    return mk3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
FH_HD float clampf_(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

// x / c, correctly rounded, for c one of the model constants the flight
// divides by (MASS, G, TAU_ROLL, V0), with rc = 1 / c rounded: the product
// q = x rc, then one correction q + (x - q c) rc, both steps fused
// multiply-adds (the remainder x - q c is exact). Compared with the division
// for every float x with 2^-104 <= |x| <= 2^100 and each of the four
// constants, not one bit differs (tests/div_check.cpp, exhaustive mode).
// Other x divide: zero (the correction would lose the sign of -0), subnormal
// and tiny values, huge values, infinities and NaN.
FH_HD float div_model_const(float x, float c, float rc) {
    float q = x * rc;
    const float ax = fabsf(x);
    if (ax >= 0x1p-104f && ax <= 0x1p100f)
        q = fmaf(fmaf(-q, c, x), rc, q);
    else
        q = x / c;
    return q;
}
FH_HD f3 div_model_const(f3 v, float c, float rc) {
    return mk3(div_model_const(v.x, c, rc), div_model_const(v.y, c, rc), div_model_const(v.z, c, rc));
}

// glam normalize: v * (1 / |v|); forward/side_of fallbacks from model.rs:15-31.
FH_HD f3 forward_of(f3 v) {
    float l2 = dot3(v, v);
    return l2 > 1e-12f ? v * (1.0f / sqrtf(l2)) : mk3(1.f, 0.f, 0.f);
}
// forward_of(v), also returning |v| = sqrt(v.v) (the speed), from one square root.
FH_HD f3 forward_and_length(f3 v, float &len) {
    const float l2 = dot3(v, v);
    len = sqrtf(l2);
    return l2 > 1e-12f ? v * (1.0f / len) : mk3(1.f, 0.f, 0.f);
}
FH_HD f3 side_of(f3 f) {
    f3 s = cross3(f, mk3(0.f, 0.f, 1.f));
    float l2 = dot3(s, s);
    return l2 > 1e-8f ? s * (1.0f / sqrtf(l2)) : mk3(1.f, 0.f, 0.f);
}
// glam normalize_or_zero: keep v * rcp only for a finite, positive rcp.
FH_HD f3 normalize_or_zero(f3 v) {
    float rcp = 1.0f / sqrtf(dot3(v, v));
    return (rcp > 0.0f && rcp < __builtin_huge_valf()) ? v * rcp : mk3(0.f, 0.f, 0.f);
}

// params.rs:61-65
FH_HD float radius_step(float r, uint32_t n) {
    float nn = (float)(n > 1u ? n : 1u);
    float factor = powf(N_C / nn, RADIUS_EXP);
    return clampf_(r * factor, R_H, R_MAX);
}

// cell_of (cuda/flock.cu:77-79) with Rust `as i32` saturation, NaN -> 0.
FH_HD int32_t cell_coord(float p, float inv_cell) {
    float f = floorf(p * inv_cell);
#if defined(__HIP_DEVICE_COMPILE__)
    // v_cvt_i32_f32 clamps out-of-range values to INT32_MIN / INT32_MAX and
    // converts NaN to 0: the same saturation, in one instruction.
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(f));
    return r;
#else
    if (!(f == f)) return 0;
    if (f >= 2147483647.0f) return 2147483647;
    if (f <= -2147483648.0f) return (int32_t)(-2147483647 - 1);
    return (int32_t)f;
#endif
}
// bucket_of (cuda/flock.cu:80-83): wrapping u32 multiply-xor hash.
FH_HD uint32_t bucket_hash(int32_t cx, int32_t cy, int32_t cz) {
    return (uint32_t)cx * 73856093u ^ (uint32_t)cy * 19349663u ^ (uint32_t)cz * 83492791u;
}
// Locality-preserving bucket hash (FH_FLAG_SPATIAL; not the reference's): the
// Morton interleave of the low 10 bits of each cell coordinate, plus the
// reference hash of the 1024-cell block the cell lies in. Nearby cells get
// nearby buckets, so their ranges share lines of the offsets table and their
// birds neighbouring slots; after the bucket mask, cells 2^(log2(B)/3) apart
// along an axis share buckets as in any hash.
FH_HD uint32_t spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
FH_HD uint32_t spatial_hash(int32_t cx, int32_t cy, int32_t cz) {
    const uint32_t m = spread10((uint32_t)cx) | (spread10((uint32_t)cy) << 1) | (spread10((uint32_t)cz) << 2);
    return m + bucket_hash((int32_t)((uint32_t)cx >> 10), (int32_t)((uint32_t)cy >> 10),
                           (int32_t)((uint32_t)cz >> 10));
}
// cell_intersects_sphere (cuda/flock.cu:85-91), +1e-4 slack on r^2.
FH_HD bool cell_meets_sphere(int32_t cx, int32_t cy, int32_t cz, float cell, f3 p, float r2) {
    const float lo_x = (float)cx * cell, lo_y = (float)cy * cell, lo_z = (float)cz * cell;
    float dx = clampf_(p.x, lo_x, lo_x + cell) - p.x;
    float dy = clampf_(p.y, lo_y, lo_y + cell) - p.y;
    float dz = clampf_(p.z, lo_z, lo_z + cell) - p.z;
    return dx * dx + dy * dy + dz * dz <= r2 + 1e-4f;
}

// buckets_for (cuda/flock.cu:368-372): next_pow2(2n) clamped to [2^16, 2^22].
inline uint32_t buckets_for(uint32_t n) {
    uint64_t want = 1;
    while (want < (uint64_t)n * 2) want <<= 1;
    if (want < (1u << 16)) want = 1u << 16;
    if (want > (1u << 22)) want = 1u << 22;
    return (uint32_t)want;
}

}  // namespace fh
