/*
 * flockref: the CPU oracle for the flock step.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker, never the product: only the
 * tests, the smoke entry point and bench.py's CPU-baseline leg may load it.
 * The HIP library (libflock_hip.so) never links, calls or falls back to it.
 *
 * It restates, in C, the reference's Rust oracle `crates/flock-ref`:
 *   - constants and radius_step           crates/flock-ref/src/params.rs:6-65
 *   - forward / side_of / body_up          crates/flock-ref/src/model.rs:15-38
 *   - decide (brute force, index order)    crates/flock-ref/src/model.rs:53-145
 *   - step_bird                            crates/flock-ref/src/model.rs:149-176
 *   - Flock::step order                    crates/flock-ref/src/model.rs:182-200
 *   - SplitMix64 + scatter, .flk I/O       crates/flock-ref/src/state.rs:58-170
 *   - stats / diff / statdiff              crates/flock-ref/src/main.rs:83-208
 * plus the grid semantics of the reference GPU step
 *   - cell_of / bucket_of / sphere test    cuda/flock.cu:77-91
 *   - buckets_for                          cuda/flock.cu:368-372
 *   - the two grid walks of decide         cuda/flock.cu:205-273
 * and the grid validator check_grid        cutile/src/main.rs:304-362.
 *
 * Vector arithmetic follows glam 0.34.1's scalar Vec3 (the reference's only
 * maths dependency, Cargo.lock:219-221): dot = (x*x')+(y*y')+(z*z') left to
 * right, length = sqrtf(dot), normalize = v * (1/length),
 * normalize_or_zero keeps v*(1/length) only when that reciprocal is finite
 * and positive. Rust's f32 transcendental calls go to the platform libm, so
 * this file calls glibc's expf/powf/atanf/sinf/cosf/cbrtf. Build with
 * -ffp-contract=off (Rust never fuses a*b+c).
 *
 * State exchange uses the .flk record layout: 15 little-endian 4-byte words
 * per bird (pos xyz, vel xyz, roll, spin, radius, steering xyz,
 * next_decision, neighbours as u32 bits, centrality).
 */
#ifndef FLOCKREF_H
#define FLOCKREF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FR_WORDS 15

/* Word indices inside one 15-word record (state.rs:4-7). */
enum {
    FR_PX = 0, FR_PY, FR_PZ, FR_VX, FR_VY, FR_VZ, FR_ROLL, FR_SPIN, FR_RADIUS,
    FR_SX, FR_SY, FR_SZ, FR_NEXT, FR_NB, FR_CEN
};

/* ---- constants (params.rs:6-58) -------------------------------------- */
/* Writes the 28 constants in the order of params.rs (DT ... GRID_CELL) plus
 * the derived RADIUS_LERP/3 exponent; returns how many were written. */
int fr_constants(float *out, int cap);
float fr_radius_step(float r, uint32_t n);

/* ---- SplitMix64 and the reference scatter (state.rs:58-101) ---------- */
uint64_t fr_splitmix_next(uint64_t *state);
float fr_splitmix_unit(uint64_t *state);
void fr_scatter(uint32_t n, uint64_t seed, float *words);

/* ---- model (model.rs) ------------------------------------------------- */
void fr_forward(const float v[3], float out[3]);
void fr_side_of(const float f[3], float out[3]);
void fr_body_up(const float f[3], float beta, float out[3]);
/* Brute-force decision of bird i from the snapshot `words` (model.rs:53-145).
 * out: force xyz, radius, neighbours (as float bits), centrality. */
void fr_decide(const float *words, uint32_t n, uint32_t i, float z0, float out[6]);
/* Advances one 15-word record by dt under its held steering (model.rs:149-176). */
void fr_step_bird(float *rec, float dt);
/* `steps` x Flock::step (model.rs:182-200) in place; *time advances by DT per
 * step in f32. threads > 1 computes the (independent) decisions of one step
 * in parallel; the result is bit-identical to threads = 1. */
void fr_step(float *words, uint32_t n, float *time, uint32_t steps, int threads);
/* Number of birds due at time t (t >= next_decision). */
uint32_t fr_count_due(const float *words, uint32_t n, float t);

/* ---- checker CLI maths (main.rs:83-208) ------------------------------ */
/* mean speed, mean neighbours, mean radius, polarization (main.rs:83-90) */
void fr_stats(const float *words, uint32_t n, float out[4]);
/* Pointwise diff (main.rs:136-208). out[0..7): max |dpos|, |dvel|, |dradius|,
 * |dcentrality|, |dsteering|, |droll|, |dspin|; counts[0] = neighbour-count
 * mismatches, counts[1] = birds above pos/vel tolerance. Returns 1 on PASS. */
int fr_diff(const float *a, const float *b, uint32_t n, float pos_tol, float vel_tol,
            uint32_t allow_outliers, float outlier_max, float out[7], uint32_t counts[2]);
/* Aggregate diff with tolerances [0.02, 0.1, 0.01, 0.005] (main.rs:104-133).
 * out: sa[4], sb[4]. Returns 1 on PASS. */
int fr_statdiff(const float *a, const float *b, uint32_t n, float out[8]);

/* ---- .flk files (state.rs:104-170) ------------------------------------ */
/* Returns 0 on success, 1 cannot open, 2 not FLK1, 3 wrong words, 4 truncated.
 * With words == NULL only the header is read (n and time). */
int fr_read_flk(const char *path, uint32_t *n, float *time, float *words, uint32_t cap);
int fr_write_flk(const char *path, uint32_t n, float time, const float *words);

/* ---- grid semantics (cuda/flock.cu:77-91, 368-372) -------------------- */
uint32_t fr_buckets_for(uint32_t n);
/* floor(p * (1/cell)) with Rust `as i32` saturation (NaN -> 0). */
int32_t fr_cell_coord(float p, float inv_cell);
uint32_t fr_bucket_of(int32_t cx, int32_t cy, int32_t cz, uint32_t mask);
int fr_cell_intersects_sphere(int32_t cx, int32_t cy, int32_t cz, float cell,
                              const float p[3], float r2);
/* Stable counting sort of the birds by bucket: keys[n], counts[B],
 * offsets[B+1], sorted[n] (within a bucket, ascending bird index). */
void fr_grid(const float *words, uint32_t n, uint32_t buckets, float cell,
             uint32_t *keys, uint32_t *counts, uint32_t *offsets, uint32_t *sorted);
/* check_grid (cutile/src/main.rs:304-362). Returns 0 when valid, else a code
 * 1..6 naming the first violated invariant; msg receives the message. */
int fr_check_grid(const uint32_t *bucket, uint32_t n, const uint32_t *offsets,
                  const uint32_t *sorted, uint32_t buckets, char *msg, int msg_cap);

/* ---- neighbour sets --------------------------------------------------- */
/* Brute force (model.rs:67-75): indices j != i with |q_j - p_i|^2 <= (k r_i)^2,
 * k = 1 (walk 1) or 2 (walk 2), ascending. Returns the count (may exceed cap;
 * only cap are written). */
uint32_t fr_neighbours_brute(const float *words, uint32_t n, uint32_t i, int k,
                             uint32_t *out, uint32_t cap);

/* ---- grid-walk decide (cuda/flock.cu:190-290 semantics) -------------- */
/* The reference GPU decide restated on the CPU over a grid from fr_grid:
 * walk 1 over the r box with the r sphere test, walk 2 over the 2r box with
 * the 2r sphere test, exact-cell check, candidates in bucket order. For large
 * flocks where the brute force is too slow. Same out layout as fr_decide. */
void fr_decide_grid(const float *words, uint32_t n, uint32_t i, float z0, float cell,
                    uint32_t buckets, const uint32_t *offsets, const uint32_t *sorted,
                    float out[6]);
/* fr_decide_grid for `count` birds (threads > 1: in parallel); out is
 * count x 6 in the fr_decide layout. */
void fr_decide_grid_many(const float *words, uint32_t n, const uint32_t *birds, uint32_t count, float z0,
                         float cell, uint32_t buckets, const uint32_t *offsets, const uint32_t *sorted,
                         float *out, int threads);
/* Flock::step using fr_decide_grid for the decisions (threads as fr_step). */
void fr_step_grid(float *words, uint32_t n, float *time, uint32_t steps, float cell,
                  uint32_t buckets, int threads);

#ifdef __cplusplus
}
#endif
#endif
