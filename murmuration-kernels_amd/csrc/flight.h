// Banked flight (crates/flock-ref/src/model.rs:149-176, cuda/flock.cu:293-328)
// shared by the integrate passes and by decide's fused flight.
#pragma once

#include "layout.h"
#include "model.h"

namespace fh {

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

// Resident flight of slot s: banked flight of in[s] into out[s] (out may be
// in), its new bucket (by its flock's hash), and the mover accounting. A bird whose new cell falls
// in another bucket is a mover: it moves its count from its old bucket to its
// new one, takes an arrival rank there, and moves its count between the scan
// tiles' totals when the tile changes. Every lane of a wave calls this
// (active or not) with consecutive slots starting at a multiple of 64: the
// wave's mover bits are written with one ballot. Words past the flock's last
// slot are not written (nothing reads them).
__device__ __forceinline__ void resident_flight(const Geometry &g, const Records &in, const Records &out,
                                                const ResWork &w, uint32_t s, bool active) {
    bool mover = false;
    if (active) {
        const Seg &sg = g.segs[slot_segment(g, in, s)];
        float4 p4 = in.p4[s];
        float4 v4 = in.v4[s];
        float rate;
        if (sg.left)
            integrate_bird(p4, v4, in.s4[s], rate);
        else
            rate = in.spin[s];  // a frozen flock keeps its state
        out.p4[s] = p4;
        out.v4[s] = v4;
        out.spin[s] = rate;
        const uint32_t kn = bucket_of_cell(sg, cell_coord(p4.x, sg.inv_cell), cell_coord(p4.y, sg.inv_cell),
                                           cell_coord(p4.z, sg.inv_cell));
        const uint32_t ko = in.key[s];
        if (kn != ko) {
            mover = true;
            w.arank[s] = atomicAdd(&w.arr[kn], 1u);
            atomicAdd(&w.counts[kn], 1u);
            atomicSub(&w.counts[ko], 1u);
            if (kn / RES_TILE != ko / RES_TILE) {
                atomicAdd(&w.tsum[kn / RES_TILE], 1u);
                atomicSub(&w.tsum[ko / RES_TILE], 1u);
            }
        }
    }
    const uint64_t bits = __ballot(mover);
    const uint32_t lane = threadIdx.x & 63u;
    if ((lane == 0u || lane == 32u) && s < g.n_slots)
        w.mbits[s >> 5] = lane == 0u ? (uint32_t)bits : (uint32_t)(bits >> 32);
}

}  // namespace fh
