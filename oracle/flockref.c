/*
 * flockref.c: the CPU oracle (test infrastructure only; see flockref.h).
 *
 * Every function cites the reference file:line it restates. Compile with
 * -O2 -ffp-contract=off and no fast-math so each float operation rounds as
 * the Rust oracle's does.
 */
#include "flockref.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- constants: crates/flock-ref/src/params.rs:6-58 -------------------- */
static const float DT = 0.005f;
static const float DU = 0.05f;
static const float V0 = 10.0f;
static const float MASS = 0.08f;
static const float CL_CD = 3.3f;
static const float G = 9.81f;
#define L0 (MASS * G)                /* params.rs:18, evaluated in f32 */
#define D0 (L0 / CL_CD)              /* params.rs:20 */
static const float TAU = 1.0f;
static const float TAU_ROLL = 0.15f;
static const float MAX_ROLL_RATE = 6.0f;
/* params.rs:28: 80 * std::f32::consts::PI / 180 in f32 */
#define PI_F32 3.14159265358979323846264338327950288f
#define MAX_BANK (80.0f * PI_F32 / 180.0f)
#define V_MIN (0.5f * V0)
#define V_MAX (2.0f * V0)
static const float N_C = 6.5f;
static const float R_MAX = 100.0f;
static const float RADIUS_LERP = 0.1f;
static const float R_H = 0.2f;
static const float SIGMA = 1.37f;
static const float W_SEP = 1.0f;
static const float W_ALIGN = 0.5f;
static const float W_COH = 0.65f;
/* params.rs:50: -FRAC_1_SQRT_2 */
#define BLIND_COS (-0.707106781186547524400844362104849039f)
static const float W_ROOST_V = 0.2f;
static const float H_ROOST_V = 40.0f;
static const float Z0 = 100.0f;
static const float GRID_CELL = 4.0f;
/* state.rs:52-54 */
static const float DENSITY = 0.5f;
static const float START_RADIUS = 1.5f;

int fr_constants(float *out, int cap) {
    const float c[] = {DT,    DU,    V0,        MASS,          CL_CD,    G,      L0,
                       D0,    TAU,   TAU_ROLL,  MAX_ROLL_RATE, MAX_BANK, V_MIN,  V_MAX,
                       N_C,   R_MAX, RADIUS_LERP, R_H,         SIGMA,    W_SEP,  W_ALIGN,
                       W_COH, BLIND_COS, W_ROOST_V, H_ROOST_V,  Z0,       GRID_CELL,
                       RADIUS_LERP / 3.0f};
    int k = (int)(sizeof c / sizeof c[0]);
    for (int i = 0; i < k && i < cap; ++i) out[i] = c[i];
    return k;
}

/* ---- glam 0.34.1 scalar Vec3 ------------------------------------------- */
typedef struct { float x, y, z; } v3;
static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 muls(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 divs(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline float dot(v3 a, v3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
static inline float length(v3 a) { return sqrtf(dot(a, a)); }
static inline v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline v3 normalize(v3 a) { return muls(a, 1.0f / length(a)); }
static inline v3 normalize_or_zero(v3 a) {
    float rcp = 1.0f / length(a);
    if (isfinite(rcp) && rcp > 0.0f) return muls(a, rcp);
    return mk(0.0f, 0.0f, 0.0f);
}
/* Rust f32::max / f32::clamp for non-NaN operands (IEEE maxNum). */
static inline float fmax_r(float a, float b) { return fmaxf(a, b); }
static inline float clamp_r(float v, float lo, float hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}

/* params.rs:61-65 */
float fr_radius_step(float r, uint32_t n) {
    float nn = (float)(n > 1u ? n : 1u);
    float factor = powf(N_C / nn, RADIUS_LERP / 3.0f);
    return clamp_r(r * factor, R_H, R_MAX);
}

/* ---- model.rs:15-38 ----------------------------------------------------- */
static inline v3 forward(v3 v) {
    if (dot(v, v) > 1e-12f) return normalize(v);
    return mk(1.0f, 0.0f, 0.0f);
}
static inline v3 side_of(v3 f) {
    v3 s = cross(f, mk(0.0f, 0.0f, 1.0f));
    if (dot(s, s) > 1e-8f) return normalize(s);
    return mk(1.0f, 0.0f, 0.0f);
}
static inline v3 body_up(v3 f, float beta) {
    v3 side = side_of(f);
    v3 up0 = cross(side, f);
    return add(muls(up0, cosf(beta)), muls(side, sinf(beta)));
}
void fr_forward(const float v[3], float out[3]) {
    v3 r = forward(mk(v[0], v[1], v[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void fr_side_of(const float f[3], float out[3]) {
    v3 r = side_of(mk(f[0], f[1], f[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void fr_body_up(const float f[3], float beta, float out[3]) {
    v3 r = body_up(mk(f[0], f[1], f[2]), beta);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

static inline v3 rec_pos(const float *r) { return mk(r[FR_PX], r[FR_PY], r[FR_PZ]); }
static inline v3 rec_vel(const float *r) { return mk(r[FR_VX], r[FR_VY], r[FR_VZ]); }

/* One neighbour of walk 1 (model.rs:73-92 / flock.cu:225-237). */
typedef struct {
    v3 sep, coh, align, csum;
    uint32_t n_all, n_star, n_align, n_g;
} acc_t;

static inline void walk1_visit(acc_t *a, v3 d, float d2, v3 f_i, const float *other) {
    float dist = sqrtf(d2);
    a->n_all += 1;
    v3 dhat = divs(d, fmax_r(dist, 1e-6f));
    float g;
    if (dist <= R_H) {
        g = 1.0f;
    } else {
        float t = dist - R_H;
        g = expf(-(t * t) / (2.0f * SIGMA * SIGMA));
    }
    a->sep = sub(a->sep, muls(dhat, g));
    if (dist > R_H) {
        a->coh = add(a->coh, dhat);
        a->n_star += 1;
    }
    if (dot(f_i, dhat) >= BLIND_COS) {
        a->align = add(a->align, forward(rec_vel(other)));
        a->n_align += 1;
    }
}

/* Combine the sums into a decision (model.rs:95-144 / flock.cu:242-289). */
static void finish_decision(const acc_t *a, const float *me, float z0, float out[6]) {
    v3 p = rec_pos(me);
    v3 v = rec_vel(me);
    float r = me[FR_RADIUS];
    v3 f_i = forward(v);
    v3 force = mk(0.0f, 0.0f, 0.0f);
    if (a->n_all > 0) force = add(force, muls(a->sep, W_SEP / (float)a->n_all));
    float centrality = a->n_g > 0 ? length(a->csum) / (float)a->n_g : 0.0f;
    if (a->n_star > 0) force = add(force, muls(normalize_or_zero(a->coh), W_COH * centrality));
    v3 al = a->n_align > 0 ? sub(divs(a->align, (float)a->n_align), f_i) : mk(0.0f, 0.0f, 0.0f);
    force = add(force, muls(al, W_ALIGN));
    force = add(force, muls(mk(0.0f, 0.0f, 1.0f), W_ROOST_V * clamp_r((z0 - p.z) / H_ROOST_V, -1.0f, 1.0f)));
    float speed = length(v);
    force = add(force, muls(f_i, (MASS / TAU) * (V0 - speed)));
    out[0] = force.x;
    out[1] = force.y;
    out[2] = force.z;
    out[3] = fr_radius_step(r, a->n_all);
    uint32_t nb = a->n_all;
    memcpy(&out[4], &nb, 4);
    out[5] = centrality;
}

/* model.rs:53-145 (brute force over j in index order). */
void fr_decide(const float *words, uint32_t n, uint32_t i, float z0, float out[6]) {
    const float *me = words + (size_t)i * FR_WORDS;
    v3 p = rec_pos(me);
    float r = me[FR_RADIUS];
    v3 f_i = forward(rec_vel(me));
    acc_t a;
    memset(&a, 0, sizeof a);
    float r2 = r * r;
    for (uint32_t j = 0; j < n; ++j) {
        if (j == i) continue;
        const float *o = words + (size_t)j * FR_WORDS;
        v3 d = sub(rec_pos(o), p);
        float d2 = dot(d, d);
        if (d2 > r2) continue;
        walk1_visit(&a, d, d2, f_i, o);
    }
    float c2 = (2.0f * r) * (2.0f * r);
    for (uint32_t j = 0; j < n; ++j) {
        if (j == i) continue;
        const float *o = words + (size_t)j * FR_WORDS;
        v3 d = sub(rec_pos(o), p);
        float d2 = dot(d, d);
        if (d2 > c2) continue;
        a.csum = add(a.csum, divs(d, fmax_r(sqrtf(d2), 1e-6f)));
        a.n_g += 1;
    }
    finish_decision(&a, me, z0, out);
}

/* model.rs:149-176 */
void fr_step_bird(float *b, float dt) {
    v3 vel = rec_vel(b);
    v3 steering = mk(b[FR_SX], b[FR_SY], b[FR_SZ]);
    v3 f = forward(vel);
    v3 side = side_of(f);
    float speed = fmax_r(length(vel), 1e-6f);

    float a_lat = dot(steering, side) / MASS;
    float target = clamp_r(atanf(a_lat / G), -MAX_BANK, MAX_BANK);
    float rate = clamp_r((target - b[FR_ROLL]) / TAU_ROLL, -MAX_ROLL_RATE, MAX_ROLL_RATE);
    b[FR_SPIN] = rate;
    float beta = clamp_r(b[FR_ROLL] + rate * dt, -MAX_BANK, MAX_BANK);
    b[FR_ROLL] = beta;

    v3 up = body_up(f, beta);
    float cos_beta = fmax_r(cosf(beta), 0.2f);
    float sr = speed / V0;
    float lift = sr * sr * L0 / cos_beta;
    float drag = sr * sr * D0;
    /* up*lift - f*drag + f*D0 - Z*MASS*G, left to right */
    v3 ff = sub(add(sub(muls(up, lift), muls(f, drag)), muls(f, D0)),
                muls(muls(mk(0.0f, 0.0f, 1.0f), MASS), G));

    v3 control = sub(steering, muls(side, dot(steering, side)));
    vel = add(vel, muls(divs(add(control, ff), MASS), dt));
    float sp = length(vel);
    if (sp > V_MAX) {
        vel = muls(vel, V_MAX / sp);
    } else if (sp < V_MIN && sp > 1e-6f) {
        vel = muls(vel, V_MIN / sp);
    }
    v3 pos = add(rec_pos(b), muls(vel, dt));
    b[FR_PX] = pos.x; b[FR_PY] = pos.y; b[FR_PZ] = pos.z;
    b[FR_VX] = vel.x; b[FR_VY] = vel.y; b[FR_VZ] = vel.z;
}

uint32_t fr_count_due(const float *words, uint32_t n, float t) {
    uint32_t c = 0;
    for (uint32_t i = 0; i < n; ++i) c += t >= words[(size_t)i * FR_WORDS + FR_NEXT];
    return c;
}

static void apply_decision(float *rec, const float d[6]) {
    rec[FR_SX] = d[0];
    rec[FR_SY] = d[1];
    rec[FR_SZ] = d[2];
    rec[FR_RADIUS] = d[3];
    memcpy(&rec[FR_NB], &d[4], 4);
    rec[FR_CEN] = d[5];
    rec[FR_NEXT] += DU;
}

/* model.rs:182-200: time += DT; decide (snapshot) for due birds; apply; step_bird all. */
void fr_step(float *words, uint32_t n, float *time, uint32_t steps, int threads) {
    float *dec = (float *)malloc((size_t)(n ? n : 1) * 6 * sizeof(float));
    unsigned char *due = (unsigned char *)malloc(n ? n : 1);
    for (uint32_t s = 0; s < steps; ++s) {
        *time += DT;
        const float t = *time;
        for (uint32_t i = 0; i < n; ++i) due[i] = t >= words[(size_t)i * FR_WORDS + FR_NEXT];
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
#endif
        for (int64_t i = 0; i < (int64_t)n; ++i)
            if (due[i]) fr_decide(words, n, (uint32_t)i, Z0, dec + (size_t)i * 6);
        for (uint32_t i = 0; i < n; ++i)
            if (due[i]) apply_decision(words + (size_t)i * FR_WORDS, dec + (size_t)i * 6);
        for (uint32_t i = 0; i < n; ++i) fr_step_bird(words + (size_t)i * FR_WORDS, DT);
    }
    (void)threads;
    free(dec);
    free(due);
}

/* ---- state.rs:58-101 ---------------------------------------------------- */
uint64_t fr_splitmix_next(uint64_t *s) {
    *s += 0x9E3779B97F4A7C15ull;
    uint64_t z = *s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
float fr_splitmix_unit(uint64_t *s) {
    return (float)(fr_splitmix_next(s) >> 40) / (float)(1ull << 24);
}

void fr_scatter(uint32_t n, uint64_t seed, float *words) {
    float side = cbrtf((float)n / DENSITY);
    v3 centre = mk(0.0f, 0.0f, Z0);
    uint64_t rng = seed;
    for (uint32_t i = 0; i < n; ++i) {
        float ox = (fr_splitmix_unit(&rng) - 0.5f) * side;
        float oy = (fr_splitmix_unit(&rng) - 0.5f) * side;
        float oz = (fr_splitmix_unit(&rng) - 0.5f) * side;
        float jx = fr_splitmix_unit(&rng) - 0.5f;
        float jy = fr_splitmix_unit(&rng) - 0.5f;
        float jz = fr_splitmix_unit(&rng) - 0.5f;
        v3 jitter = muls(mk(jx, jy, jz), 0.05f);
        v3 pos = add(centre, mk(ox, oy, oz));
        v3 vel = muls(normalize(add(mk(1.0f, 0.0f, 0.0f), jitter)), V0);
        float next = fr_splitmix_unit(&rng) * DU;
        float *r = words + (size_t)i * FR_WORDS;
        memset(r, 0, FR_WORDS * sizeof(float));
        r[FR_PX] = pos.x; r[FR_PY] = pos.y; r[FR_PZ] = pos.z;
        r[FR_VX] = vel.x; r[FR_VY] = vel.y; r[FR_VZ] = vel.z;
        r[FR_RADIUS] = START_RADIUS;
        r[FR_NEXT] = next;
    }
}

/* ---- main.rs:83-133 ----------------------------------------------------- */
void fr_stats(const float *w, uint32_t n, float out[4]) {
    float nf = (float)n;
    float speed = 0.0f, nb = 0.0f, radius = 0.0f;
    v3 pol = mk(0.0f, 0.0f, 0.0f);
    for (uint32_t i = 0; i < n; ++i) speed += length(rec_vel(w + (size_t)i * FR_WORDS));
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t k;
        memcpy(&k, &w[(size_t)i * FR_WORDS + FR_NB], 4);
        nb += (float)k;
    }
    for (uint32_t i = 0; i < n; ++i) radius += w[(size_t)i * FR_WORDS + FR_RADIUS];
    for (uint32_t i = 0; i < n; ++i) pol = add(pol, forward(rec_vel(w + (size_t)i * FR_WORDS)));
    out[0] = speed / nf;
    out[1] = nb / nf;
    out[2] = radius / nf;
    out[3] = length(divs(pol, nf));
}

int fr_statdiff(const float *a, const float *b, uint32_t n, float out[8]) {
    static const float tol[4] = {0.02f, 0.1f, 0.01f, 0.005f};
    fr_stats(a, n, out);
    fr_stats(b, n, out + 4);
    int ok = 1;
    for (int k = 0; k < 4; ++k) ok &= fabsf(out[k] - out[4 + k]) <= tol[k];
    return ok;
}

static inline float maxabs3(v3 d) {
    float x = fabsf(d.x), y = fabsf(d.y), z = fabsf(d.z);
    return fmax_r(x, fmax_r(y, z));
}

/* main.rs:136-208 */
int fr_diff(const float *a, const float *b, uint32_t n, float pos_tol, float vel_tol,
            uint32_t allow, float outlier_max, float out[7], uint32_t counts[2]) {
    float dp = 0, dv = 0, dr = 0, dc = 0, ds = 0, droll = 0, dspin = 0;
    uint32_t outliers = 0, nb_mismatch = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const float *x = a + (size_t)i * FR_WORDS, *y = b + (size_t)i * FR_WORDS;
        float bp = maxabs3(sub(rec_pos(x), rec_pos(y)));
        float bv = maxabs3(sub(rec_vel(x), rec_vel(y)));
        dp = fmax_r(dp, bp);
        dv = fmax_r(dv, bv);
        outliers += (bp > pos_tol || bv > vel_tol);
        dr = fmax_r(dr, fabsf(x[FR_RADIUS] - y[FR_RADIUS]));
        dc = fmax_r(dc, fabsf(x[FR_CEN] - y[FR_CEN]));
        ds = fmax_r(ds, maxabs3(sub(mk(x[FR_SX], x[FR_SY], x[FR_SZ]), mk(y[FR_SX], y[FR_SY], y[FR_SZ]))));
        droll = fmax_r(droll, fabsf(x[FR_ROLL] - y[FR_ROLL]));
        dspin = fmax_r(dspin, fabsf(x[FR_SPIN] - y[FR_SPIN]));
        uint32_t nx, ny;
        memcpy(&nx, &x[FR_NB], 4);
        memcpy(&ny, &y[FR_NB], 4);
        nb_mismatch += nx != ny;
    }
    out[0] = dp; out[1] = dv; out[2] = dr; out[3] = dc; out[4] = ds; out[5] = droll; out[6] = dspin;
    counts[0] = nb_mismatch;
    counts[1] = outliers;
    if (allow == 0) return dp <= pos_tol && dv <= vel_tol && nb_mismatch == 0;
    return outliers <= allow && dp <= outlier_max && dv <= outlier_max && nb_mismatch == 0;
}

/* ---- state.rs:104-170 ---------------------------------------------------- */
int fr_read_flk(const char *path, uint32_t *n, float *time, float *words, uint32_t cap) {
    FILE *fp = fopen(path, "rb");
    if (!fp) return 1;
    uint32_t head[4];
    if (fread(head, 4, 4, fp) != 4) { fclose(fp); return 4; }
    if (memcmp(head, "FLK1", 4) != 0) { fclose(fp); return 2; }
    if (head[2] != FR_WORDS) { fclose(fp); return 3; }
    *n = head[1];
    memcpy(time, &head[3], 4);
    if (words) {
        if (head[1] > cap) { fclose(fp); return 4; }
        size_t want = (size_t)head[1] * FR_WORDS;
        if (fread(words, 4, want, fp) != want) { fclose(fp); return 4; }
    }
    fclose(fp);
    return 0;
}

int fr_write_flk(const char *path, uint32_t n, float time, const float *words) {
    FILE *fp = fopen(path, "wb");
    if (!fp) return 1;
    uint32_t head[4] = {0x314B4C46u, n, FR_WORDS, 0};
    memcpy(&head[3], &time, 4);
    int ok = fwrite(head, 4, 4, fp) == 4 &&
             fwrite(words, 4, (size_t)n * FR_WORDS, fp) == (size_t)n * FR_WORDS;
    return fclose(fp) == 0 && ok ? 0 : 1;
}

/* ---- grid: cuda/flock.cu:77-91, 368-372 ---------------------------------- */
uint32_t fr_buckets_for(uint32_t n) {
    uint64_t want = 1;
    while (want < (uint64_t)n * 2) want <<= 1;
    if (want < (1u << 16)) want = 1u << 16;
    if (want > (1u << 22)) want = 1u << 22;
    return (uint32_t)want;
}

/* Rust `as i32` saturates (NaN -> 0); flock.cu:77-79 floors p * (1/cell). */
int32_t fr_cell_coord(float p, float inv_cell) {
    float f = floorf(p * inv_cell);
    if (f != f) return 0;
    if (f >= 2147483647.0f) return 2147483647;
    if (f <= -2147483648.0f) return (int32_t)(-2147483647 - 1);
    return (int32_t)f;
}

uint32_t fr_bucket_of(int32_t cx, int32_t cy, int32_t cz, uint32_t mask) {
    uint32_t h = (uint32_t)cx * 73856093u ^ (uint32_t)cy * 19349663u ^ (uint32_t)cz * 83492791u;
    return h & mask;
}

int fr_cell_intersects_sphere(int32_t cx, int32_t cy, int32_t cz, float cell, const float p[3],
                              float r2) {
    const float lo_x = (float)cx * cell, lo_y = (float)cy * cell, lo_z = (float)cz * cell;
    float dx = fminf(fmaxf(p[0], lo_x), lo_x + cell) - p[0];
    float dy = fminf(fmaxf(p[1], lo_y), lo_y + cell) - p[1];
    float dz = fminf(fmaxf(p[2], lo_z), lo_z + cell) - p[2];
    return dx * dx + dy * dy + dz * dz <= r2 + 1e-4f;
}

static inline uint32_t key_of(const float *rec, float inv_cell, uint32_t mask) {
    return fr_bucket_of(fr_cell_coord(rec[FR_PX], inv_cell), fr_cell_coord(rec[FR_PY], inv_cell),
                        fr_cell_coord(rec[FR_PZ], inv_cell), mask);
}

/* The counting sort of flock.cu:118-179 (histogram, exclusive scan, scatter),
 * made stable: within a bucket the birds keep ascending index order. */
void fr_grid(const float *words, uint32_t n, uint32_t buckets, float cell, uint32_t *keys,
             uint32_t *counts, uint32_t *offsets, uint32_t *sorted) {
    const float inv = 1.0f / cell;
    const uint32_t mask = buckets - 1;
    memset(counts, 0, (size_t)buckets * 4);
    for (uint32_t i = 0; i < n; ++i) {
        keys[i] = key_of(words + (size_t)i * FR_WORDS, inv, mask);
        counts[keys[i]] += 1;
    }
    uint32_t run = 0;
    for (uint32_t b = 0; b < buckets; ++b) {
        offsets[b] = run;
        run += counts[b];
    }
    offsets[buckets] = run;
    uint32_t *cursor = (uint32_t *)malloc((size_t)buckets * 4);
    memcpy(cursor, offsets, (size_t)buckets * 4);
    for (uint32_t i = 0; i < n; ++i) sorted[cursor[keys[i]]++] = i;
    free(cursor);
}

/* cutile/src/main.rs:304-362 */
int fr_check_grid(const uint32_t *bucket, uint32_t n, const uint32_t *offsets,
                  const uint32_t *sorted, uint32_t buckets, char *msg, int cap) {
    char local[160];
    if (!msg || cap <= 0) { msg = local; cap = (int)sizeof local; }
    msg[0] = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (bucket[i] >= buckets) {
            snprintf(msg, cap, "a bucket index is outside 0..buckets (grid_histogram)");
            return 1;
        }
    if (offsets[0] != 0) {
        snprintf(msg, cap, "offsets[0] = %u (grid_scan_blocks/totals)", offsets[0]);
        return 2;
    }
    if (offsets[buckets] != n) {
        snprintf(msg, cap, "offsets[buckets] = %u but n = %u (grid_scan_add)", offsets[buckets], n);
        return 3;
    }
    for (uint32_t b = 0; b < buckets; ++b)
        if (offsets[b + 1] < offsets[b]) {
            snprintf(msg, cap, "offsets decreases at bucket %u (grid_scan_*)", b);
            return 4;
        }
    uint32_t *count = (uint32_t *)calloc(buckets ? buckets : 1, 4);
    for (uint32_t i = 0; i < n; ++i) count[bucket[i]] += 1;
    for (uint32_t b = 0; b < buckets; ++b)
        if (offsets[b + 1] - offsets[b] != count[b]) {
            snprintf(msg, cap, "bucket %u holds %u birds but its range is %u (grid_histogram vs scan)",
                     b, count[b], offsets[b + 1] - offsets[b]);
            free(count);
            return 5;
        }
    free(count);
    unsigned char *seen = (unsigned char *)calloc(n ? n : 1, 1);
    for (uint32_t b = 0; b < buckets; ++b)
        for (uint32_t s = offsets[b]; s < offsets[b + 1]; ++s) {
            uint32_t j = sorted[s];
            if (j >= n || seen[j]) {
                snprintf(msg, cap, "sorted[%u] = %u is out of range or repeated (grid_scatter)", s, j);
                free(seen);
                return 6;
            }
            seen[j] = 1;
            if (bucket[j] != b) {
                snprintf(msg, cap, "bird %u sits in bucket %u's range but belongs to %u (grid_scatter)",
                         j, b, bucket[j]);
                free(seen);
                return 6;
            }
        }
    free(seen);
    return 0;
}

/* ---- neighbour sets ------------------------------------------------------ */
uint32_t fr_neighbours_brute(const float *words, uint32_t n, uint32_t i, int k, uint32_t *out,
                             uint32_t cap) {
    const float *me = words + (size_t)i * FR_WORDS;
    v3 p = rec_pos(me);
    float r = me[FR_RADIUS];
    float lim = k == 2 ? (2.0f * r) * (2.0f * r) : r * r;
    uint32_t c = 0;
    for (uint32_t j = 0; j < n; ++j) {
        if (j == i) continue;
        v3 d = sub(rec_pos(words + (size_t)j * FR_WORDS), p);
        if (dot(d, d) > lim) continue;
        if (c < cap) out[c] = j;
        ++c;
    }
    return c;
}

/* ---- grid-walk decide: cuda/flock.cu:190-290 ------------------------------ */
void fr_decide_grid(const float *words, uint32_t n, uint32_t i, float z0, float cell,
                    uint32_t buckets, const uint32_t *offsets, const uint32_t *sorted,
                    float out[6]) {
    (void)n;
    const float *me = words + (size_t)i * FR_WORDS;
    v3 p = rec_pos(me);
    const float pa[3] = {p.x, p.y, p.z};
    float r = me[FR_RADIUS];
    v3 f_i = forward(rec_vel(me));
    const float inv = 1.0f / cell;
    const uint32_t mask = buckets - 1;
    acc_t a;
    memset(&a, 0, sizeof a);
    for (int walk = 1; walk <= 2; ++walk) {
        /* walk 1: flock.cu:206-208 (r); walk 2: flock.cu:249-252 (4 r r, 2 r) */
        const float lim = walk == 1 ? r * r : 4.0f * r * r;
        const float rr = walk == 1 ? r : 2.0f * r;
        int32_t lo[3], hi[3];
        lo[0] = fr_cell_coord(p.x - rr, inv); hi[0] = fr_cell_coord(p.x + rr, inv);
        lo[1] = fr_cell_coord(p.y - rr, inv); hi[1] = fr_cell_coord(p.y + rr, inv);
        lo[2] = fr_cell_coord(p.z - rr, inv); hi[2] = fr_cell_coord(p.z + rr, inv);
        for (int64_t cz = lo[2]; cz <= hi[2]; ++cz)
            for (int64_t cy = lo[1]; cy <= hi[1]; ++cy)
                for (int64_t cx = lo[0]; cx <= hi[0]; ++cx) {
                    if (!fr_cell_intersects_sphere((int32_t)cx, (int32_t)cy, (int32_t)cz, cell, pa, lim))
                        continue;
                    uint32_t b = fr_bucket_of((int32_t)cx, (int32_t)cy, (int32_t)cz, mask);
                    for (uint32_t s = offsets[b], e = offsets[b + 1]; s < e; ++s) {
                        uint32_t j = sorted[s];
                        if (j == i) continue;
                        const float *o = words + (size_t)j * FR_WORDS;
                        if (fr_cell_coord(o[FR_PX], inv) != cx || fr_cell_coord(o[FR_PY], inv) != cy ||
                            fr_cell_coord(o[FR_PZ], inv) != cz)
                            continue;
                        v3 d = sub(rec_pos(o), p);
                        float d2 = dot(d, d);
                        if (d2 > lim) continue;
                        if (walk == 1) {
                            walk1_visit(&a, d, d2, f_i, o);
                        } else {
                            a.csum = add(a.csum, divs(d, fmax_r(sqrtf(d2), 1e-6f)));
                            a.n_g += 1;
                        }
                    }
                }
    }
    finish_decision(&a, me, z0, out);
}

/* fr_decide (model.rs:53-145, brute force) for count birds, in parallel over
 * the birds: the multi-core CPU baseline. out[k * 6 ...] as fr_decide's. */
void fr_decide_many(const float *words, uint32_t n, const uint32_t *birds, uint32_t count, float z0, float *out,
                    int threads) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads > 0 ? threads : 1)
#endif
    for (int64_t k = 0; k < (int64_t)count; ++k) fr_decide(words, n, birds[k], z0, out + (size_t)k * 6);
    (void)threads;
}

void fr_decide_grid_many(const float *words, uint32_t n, const uint32_t *birds, uint32_t count, float z0,
                         float cell, uint32_t buckets, const uint32_t *offsets, const uint32_t *sorted,
                         float *out, int threads) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
#endif
    for (int64_t k = 0; k < (int64_t)count; ++k)
        fr_decide_grid(words, n, birds[k], z0, cell, buckets, offsets, sorted, out + (size_t)k * 6);
    (void)threads;
}

void fr_step_grid(float *words, uint32_t n, float *time, uint32_t steps, float cell,
                  uint32_t buckets, int threads) {
    float *dec = (float *)malloc((size_t)(n ? n : 1) * 6 * sizeof(float));
    unsigned char *due = (unsigned char *)malloc(n ? n : 1);
    uint32_t *keys = (uint32_t *)malloc((size_t)(n ? n : 1) * 4);
    uint32_t *sorted = (uint32_t *)malloc((size_t)(n ? n : 1) * 4);
    uint32_t *counts = (uint32_t *)malloc((size_t)buckets * 4);
    uint32_t *offsets = (uint32_t *)malloc(((size_t)buckets + 1) * 4);
    for (uint32_t s = 0; s < steps; ++s) {
        *time += DT;
        const float t = *time;
        fr_grid(words, n, buckets, cell, keys, counts, offsets, sorted);
        for (uint32_t i = 0; i < n; ++i) due[i] = t >= words[(size_t)i * FR_WORDS + FR_NEXT];
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
#endif
        for (int64_t i = 0; i < (int64_t)n; ++i)
            if (due[i])
                fr_decide_grid(words, n, (uint32_t)i, Z0, cell, buckets, offsets, sorted,
                               dec + (size_t)i * 6);
        for (uint32_t i = 0; i < n; ++i)
            if (due[i]) apply_decision(words + (size_t)i * FR_WORDS, dec + (size_t)i * 6);
        for (uint32_t i = 0; i < n; ++i) fr_step_bird(words + (size_t)i * FR_WORDS, DT);
    }
    (void)threads;
    free(dec); free(due); free(keys); free(sorted); free(counts); free(offsets);
}
