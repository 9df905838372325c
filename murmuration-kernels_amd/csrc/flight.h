// Banked flight (crates/flock-ref/src/model.rs:149-176, cuda/flock.cu:293-328)
// shared by the integrate passes and by decide's fused flight.
#pragma once

#include "layout.h"
#include "model.h"

namespace fh {

// Banked flight of one bird (model.rs:149-176): new (pos, next), (vel, roll)
// and the roll rate, from its (pos, next), (vel, roll) and (steering, radius).
// Divisions by the model constants take div_model_const (the same bits).
__device__ __forceinline__ void integrate_bird(float4 &p4, float4 &v4, const float4 s4, float &rate_out) {
    f3 vel = mk3(v4.x, v4.y, v4.z);
    const float roll = v4.w;
    const f3 steering = mk3(s4.x, s4.y, s4.z);

    float len;
    const f3 fwd = forward_and_length(vel, len);
    const f3 side = side_of(fwd);
    const float speed = fmaxf(len, 1e-6f);

    const float a_lat = div_model_const(dot3(steering, side), MASS, 1.0f / MASS);
    const float target = clampf_(atanf(div_model_const(a_lat, G, 1.0f / G)), -MAX_BANK, MAX_BANK);
    const float rate = clampf_(div_model_const(target - roll, TAU_ROLL, 1.0f / TAU_ROLL), -MAX_ROLL_RATE, MAX_ROLL_RATE);
    const float beta = clampf_(roll + rate * DT, -MAX_BANK, MAX_BANK);

    // body_up (model.rs:34-38): (side x f) cos b + side sin b
    // One range reduction for both (the device library's sincosf returns the
    // bits of its sinf and cosf).
    const f3 up0 = cross3(side, fwd);
    float sb, cb;
    sincosf(beta, &sb, &cb);
    const f3 up = up0 * cb + side * sb;
    const float cos_beta = fmaxf(cb, 0.2f);
    const float sr = div_model_const(speed, V0, 1.0f / V0);
    const float lift = sr * sr * L0 / cos_beta;
    const float drag = sr * sr * D0;
    const f3 flight = ((up * lift - fwd * drag) + fwd * D0) - (mk3(0.f, 0.f, 1.f) * MASS) * G;

    const f3 control = steering - side * dot3(steering, side);
    vel = vel + div_model_const(control + flight, MASS, 1.0f / MASS) * DT;
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

// A mover from bucket ko to bucket kn at slot s: its count moves between the
// buckets, it takes an arrival rank in kn, and its count moves between the
// scan tiles' totals when the tile changes.
__device__ __forceinline__ void count_mover(const ResWork &w, uint32_t s, uint32_t kn, uint32_t ko) {
    w.arank[s] = atomicAdd(&w.arr[kn], 1u);
    atomicAdd(&w.counts[kn], 1u);
    atomicSub(&w.counts[ko], 1u);
    if (kn / RES_TILE != ko / RES_TILE) {
        atomicAdd(&w.tsum[kn / RES_TILE], 1u);
        atomicSub(&w.tsum[ko / RES_TILE], 1u);
    }
}

// Resident flight of slot s from its loaded records (position and decision
// time p4, velocity and roll v4, steering and radius s4, old bucket ko):
// banked flight into out[s] (out may be in) unless its flock does not step
// (`live` false: frozen), its new bucket by its flock's hash, and the mover
// accounting. A bird whose new cell falls in another
// bucket is a mover: it moves its count from its old bucket to its new one,
// takes an arrival rank there, and moves its count between the scan tiles'
// totals when the tile changes. Returns whether the bird is a mover.
__device__ __forceinline__ bool fly_loaded(const Seg &sg, bool live, const Records &in, const Records &out,
                                           const ResWork &w, uint32_t s, float4 p4, float4 v4, const float4 s4,
                                           uint32_t ko) {
    float rate;
    if (live)
        integrate_bird(p4, v4, s4, rate);
    else
        rate = in.spin[s];  // a frozen flock keeps its state
    out.p4[s] = p4;
    out.v4[s] = v4;
    out.spin[s] = rate;
    const uint32_t kn = bucket_of_cell(sg, cell_coord(p4.x, sg.inv_cell), cell_coord(p4.y, sg.inv_cell),
                                       cell_coord(p4.z, sg.inv_cell));
    if (kn == ko) return false;
    count_mover(w, s, kn, ko);
    return true;
}

// The same flight between rebuilds (FH_FLAG_LAZY), up to the bucket test:
// banked flight into out[s]; returns the bird's new bucket through kn and
// whether it is a mover (its bucket now differs from its slot's, ko: the
// grid keeps the buckets of the last rebuild).
__device__ __forceinline__ bool fly_loaded_lazy(const Seg &sg, bool live, const Records &in, const Records &out,
                                                uint32_t s, float4 &p4, float4 &v4, const float4 s4, uint32_t ko,
                                                uint32_t &kn) {
    float rate;
    if (live)
        integrate_bird(p4, v4, s4, rate);
    else
        rate = in.spin[s];
    out.p4[s] = p4;
    out.v4[s] = v4;
    out.spin[s] = rate;
    kn = bucket_of_cell(sg, cell_coord(p4.x, sg.inv_cell), cell_coord(p4.y, sg.inv_cell),
                        cell_coord(p4.z, sg.inv_cell));
    return kn != ko;
}

// Appends the mover in slot s to the next step's list of its bucket
// kn & mmask: into the list's head record if the list is empty (claimed by a
// compare-and-swap of its slot word), else into the overflow chain, as entry
// s (entries are indexed by slot, so no counter is shared). Its velocity is
// entry s's either way.
__device__ __forceinline__ void append_mover(const ResWork &w, uint32_t s, uint32_t kn, float4 p4, float4 v4) {
    const uint32_t m = kn & w.mmask;
    float *head = reinterpret_cast<float *>(&w.mw_head4[m]);
    uint32_t *slot_word = reinterpret_cast<uint32_t *>(head + 3);
    if (atomicCAS(slot_word, MOVER_NIL, s) == MOVER_NIL) {
        head[0] = p4.x;
        head[1] = p4.y;
        head[2] = p4.z;
    } else {
        w.mw_p4[s] = make_float4(p4.x, p4.y, p4.z, __uint_as_float(s));
        w.mw_next[s] = atomicExch(&w.mw_ovf[m], s);
        atomicOr(slot_word, MOVER_OVF);
    }
    w.mw_v4[s] = v4;
}

// The mover bits of a wave whose lanes hold 64 consecutive slots from a
// multiple of 64 (every lane calls this): one ballot, written by lanes 0 and
// 32. Words past the flock's last slot are not written (nothing reads them).
__device__ __forceinline__ void write_mover_bits(const Geometry &g, const ResWork &w, uint32_t s, bool mover) {
    const uint64_t bits = __ballot(mover);
    const uint32_t lane = threadIdx.x & 63u;
    if ((lane == 0u || lane == 32u) && s < g.n_slots)
        w.mbits[s >> 5] = lane == 0u ? (uint32_t)bits : (uint32_t)(bits >> 32);
}

// Resident flight of slot s (every lane of a wave calls this, active or not,
// with consecutive slots from a multiple of 64).
__device__ __forceinline__ void resident_flight(const Geometry &g, const Records &in, const Records &out,
                                                const ResWork &w, uint32_t s, bool active) {
    bool mover = false;
    if (active) {
        const Seg &sg = g.segs[slot_segment(g, in, s)];
        mover = fly_loaded(sg, sg.left != 0u, in, out, w, s, in.p4[s], in.v4[s], in.s4[s], in.key[s]);
    }
    write_mover_bits(g, w, s, mover);
}

// Two slots per lane, s0 and s1 (each wave's lanes on 64 consecutive slots
// from a multiple of 64 for both): the records of both are loaded before
// either flies, so one memory latency covers two birds. Slots at or past
// `end` do not fly (`valid` is any slot below it, loaded in their place),
// and a wave's bits are written only for a run that starts below `end`.
template <bool MG>
__device__ __forceinline__ void resident_flight2(const Geometry &g, const Records &in, const Records &out,
                                                 const ResWork &w, uint32_t s0, uint32_t s1, uint32_t end,
                                                 uint32_t valid, bool lazy) {
    const bool a0 = s0 < end, a1 = s1 < end;
    const uint32_t l0 = a0 ? s0 : valid, l1 = a1 ? s1 : valid;
    const Seg &sg0 = g.segs[slot_segment(g, in, l0)];
    const Seg &sg1 = g.segs[slot_segment(g, in, l1)];
    const float4 p0 = in.p4[l0], v0 = in.v4[l0], q0 = in.s4[l0];
    const float4 p1 = in.p4[l1], v1 = in.v4[l1], q1 = in.s4[l1];
    const uint32_t k0 = in.key[l0], k1 = in.key[l1];
    // a batch's lazy steps: budgets as of mphase steps after the last commit
    float t_unused;
    const bool live0 = seg_step(sg0, MG ? w.mphase : 0u, t_unused), live1 = seg_step(sg1, MG ? w.mphase : 0u, t_unused);
    if (MG && lazy) {
        float4 f0 = p0, g0 = v0, f1 = p1, g1 = v1;
        uint32_t n0 = 0, n1 = 0;
        if (a0 && fly_loaded_lazy(sg0, live0, in, out, s0, f0, g0, q0, k0, n0)) append_mover(w, s0, n0, f0, g0);
        if (a1 && fly_loaded_lazy(sg1, live1, in, out, s1, f1, g1, q1, k1, n1)) append_mover(w, s1, n1, f1, g1);
        return;
    }
    const bool m0 = a0 && fly_loaded(sg0, live0, in, out, w, s0, p0, v0, q0, k0);
    if ((s0 & ~63u) < end) write_mover_bits(g, w, s0, m0);  // uniform over the wave
    const bool m1 = a1 && fly_loaded(sg1, live1, in, out, w, s1, p1, v1, q1, k1);
    if ((s1 & ~63u) < end) write_mover_bits(g, w, s1, m1);
}

}  // namespace fh
