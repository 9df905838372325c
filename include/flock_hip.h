/*
 * flock_hip.h: C ABI of libflock_hip.so, the MI355X (gfx950) flock step.
 *
 * The reference has no library interface for this path: its CUDA step is a
 * program (cuda/flock.cu:374-523) driven through a CLI and .flk files, with
 * the pass sequence of cuda/flock.cu:452-470 (memset, grid_histogram,
 * grid_scan_blocks/totals/add, grid_scatter, decide, integrate) and the
 * argument lists spelled out in oxide/src/main.rs:226-534. This header exposes
 * that pass-level API as plain C (opaque handles, pointers, sizes), so a
 * Rust/cgo/ctypes caller binds it without any C++ or HIP type.
 *
 * Conventions
 *   - Every call returns fh_status (0 = FH_OK). On failure fh_last_error()
 *     returns a thread-local message describing the last error of this thread.
 *   - One handle per host thread; calls on one handle are not thread-safe.
 *   - State crosses the boundary as .flk records: FH_WORDS 4-byte words per
 *     bird (pos xyz, vel xyz, roll, spin, radius, steering xyz,
 *     next_decision, neighbours as u32 bits, centrality), as in
 *     crates/flock-ref/src/state.rs:1-7. The caller owns host memory; the
 *     library copies it.
 *   - Device buffers are allocated once per handle (fh_create /
 *     fh_batch_create) and freed by the matching destroy call.
 *   - The library never falls back to a CPU path: without a usable gfx950
 *     device every device call fails with FH_ERR_NODEVICE.
 */
#ifndef FLOCK_HIP_H
#define FLOCK_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FH_WORDS 15

typedef int32_t fh_status;
enum {
    FH_OK = 0,
    FH_ERR_INVALID = 1,  /* bad argument (n = 0 where not allowed, bad size, NULL) */
    FH_ERR_FLK = 2,      /* malformed .flk: bad magic, words != 15, truncated */
    FH_ERR_NOMEM = 3,    /* device or host allocation failed */
    FH_ERR_HIP = 4,      /* a HIP runtime call or a kernel reported an error */
    FH_ERR_BUDGET = 5,   /* admission refused: over the HBM budget or slot cap */
    FH_ERR_STATE = 6,    /* call out of order (e.g. scatter before scan) */
    FH_ERR_IO = 7,       /* file cannot be opened or written */
    FH_ERR_NODEVICE = 8  /* no usable GPU */
};

/* Tunables of one flock (or one batch). Zero fields take the defaults. */
typedef struct fh_params {
    float cell;                /* grid cell edge [m]; 0 -> 4 (params.rs:58 GRID_CELL) */
    uint32_t buckets;          /* hash buckets; 0 -> buckets_for(n) (cuda/flock.cu:368-372); else a power of two */
    uint32_t decide_block;     /* decide workgroup size: 64, 128, 256 or 512; 0 -> 256 */
    uint32_t birds_per_thread; /* decide scans block * this sorted slots per workgroup; 0 -> 8 */
    uint32_t flags;            /* FH_FLAG_* */
    uint32_t decide_group;     /* lanes sharing one bird's walk: 1, 2, 4, 8, 16 or 64; 0 -> 1.
                                  With 1 the sums follow the walk order exactly; with 2..8 each
                                  lane sums every G-th cell, with 16 and 64 every lane walks
                                  every cell and sums every G-th candidate of each bucket (for
                                  dense cells); the G partial sums are then added */
    uint32_t bucket_shift;     /* when buckets == 0: buckets_for(n) >> bucket_shift (at least 64),
                                  0..8; a smaller table stays in L2 at the price of more hash
                                  collisions. Keys equal the reference's only at shift 0 */
    uint32_t reserved;
} fh_params;

/* Within a bucket, order birds by ascending index (a stable counting sort):
 * the step becomes bit-reproducible and independent of what else is batched. */
#define FH_FLAG_CANONICAL 1u
/* Run the histogram as its own pass instead of fusing it into the previous
 * step's integrate (the reference's pass layout, for per-pass timing). */
#define FH_FLAG_UNFUSED 2u
/* Replay each step from a captured hipGraph instead of launching kernels. */
#define FH_FLAG_GRAPH 4u
/* Wave-aggregated histogram atomics: lanes of a wave that share a bucket
 * issue one atomic (pays off when many birds share a cell). */
#define FH_FLAG_WAVE_AGG 8u
/* Never use decide's tiled path (cells shared by a workgroup's due birds
 * staged in LDS); its results are bit-identical, this is for checking. */
#define FH_FLAG_NO_TILE 16u
/* decide tests the bucket occupancy bitmap written by the scan before it
 * reads a bucket's range, so empty buckets cost a lookup in a small table
 * instead of one in the offsets table (pays off when most buckets are
 * empty and the offsets table is far larger than the cache); results are
 * bit-identical. */
#define FH_FLAG_OCC 32u
/* fh_step keeps the flock's state in bucket order between steps: only the
 * birds that change bucket are counted with atomics, and one near-sequential
 * pass moves every record to its new slot (instead of hashing every bird and
 * scattering sorted copies each step). Keys, counts and offsets stay the
 * reference's; within a bucket, birds that stay keep their order and birds
 * that arrive follow them. Single flocks (fh_step) and batches
 * (fh_batch_step); not with FH_FLAG_CANONICAL. Any other call on a single
 * flock first puts its state back in bird order; in a batch, flocks join and
 * leave the layout in place and a download flushes that flock's slots only. */
#define FH_FLAG_RESIDENT 64u
/* Bucket cells by a locality-preserving hash (the Morton order of the cell
 * coordinates within 1024-cell blocks) instead of the reference's hash: nearby
 * cells get nearby buckets and slots, so decide's walks share cache lines.
 * Neighbour sets are unchanged; keys differ from the reference's. */
#define FH_FLAG_SPATIAL 128u
/* decide runs as a persistent kernel: a grid of as many workgroups as the
 * GPU holds at once, each taking chunks of slots from a device counter until
 * none are left (instead of one workgroup per chunk). Results are
 * bit-identical; the autotuner decides whether it pays. */
#define FH_FLAG_PERSISTENT 256u
/* decide maps workgroups to chunks of slots XCD by XCD: the workgroups the
 * GPU dispatches to one XCD (one L2) take one contiguous run of chunks, so
 * neighbouring chunks share their halo cells in that L2 (no effect with
 * FH_FLAG_PERSISTENT). Results are bit-identical. */
#define FH_FLAG_XCD 512u
/* decide stages the positions of each chunk of slots, with a halo of half a
 * chunk on each side, in LDS with coalesced loads, and reads candidates that
 * fall in that window from LDS. Pays off with FH_FLAG_SPATIAL, whose bucket
 * order keeps most candidates near the chunk. Results are bit-identical. */
#define FH_FLAG_WINDOW 1024u
/* With FH_FLAG_RESIDENT: decide flies the birds of each chunk of slots right
 * after deciding them, in the same launch (the new positions go to a third
 * set of records, so other walks keep reading this step's positions). One
 * launch less per step, and the flight's streaming overlaps the walks' memory
 * latency; fh_pass_ms then counts the flight in `decide` (integrate 0). The
 * results equal the unfused resident step's. Ignored without
 * FH_FLAG_RESIDENT. */
#define FH_FLAG_FLIGHT 2048u
/* With FH_FLAG_RESIDENT and FH_FLAG_FLIGHT (not FH_FLAG_WINDOW), for single
 * flocks and batches: the layout is rebuilt every 4th step instead of every
 * step. In between, each step flies the records in place, and the birds whose
 * cell now hashes to another bucket than their slot's are also listed by
 * their current bucket, in lists that decide's walks visit too, so every bird
 * is found in the bucket of its current cell (in a batch, only by walks of
 * its own flock). Neighbour sets and counts are those of the rebuilt layout;
 * sums are taken in another order (rounding). In a batch, each flock's clock
 * and step budget between rebuilds follow from its values at the last
 * rebuild and the steps taken since, and the rebuild commits them. Any call
 * that reads the state or the grid, or changes a batch's flocks or budgets,
 * rebuilds the layout first. */
#define FH_FLAG_LAZY 4096u

/* Per-pass GPU time of fh_step, mean ms per timed step (hipEvents on the
 * handle's stream), in the rows of the reference report
 * (cuda/flock.cu:486-491). With the fused layout the histogram runs inside
 * the previous step's integrate and grid_histogram reads 0. */
typedef struct fh_pass_ms {
    double grid_histogram;
    double grid_scan;
    double grid_scatter;
    double decide;
    double integrate;
    double step_total;   /* sum of the rows above */
    double wall_ms;      /* host wall clock over the timed steps, synchronised */
    uint32_t timed_steps;
    uint32_t due_last;   /* birds that decided in the last step */
} fh_pass_ms;

typedef struct fh_flock fh_flock;

/* ---- library ------------------------------------------------------------ */
const char *fh_version(void);
const char *fh_last_error(void);
/* Writes the model constants in the order of crates/flock-ref/src/params.rs:6-58
 * (DT .. GRID_CELL) plus RADIUS_LERP/3; returns how many exist. Host only. */
int fh_constants(float *out, int cap);
/* buckets_for, cuda/flock.cu:368-372. Host only. */
uint32_t fh_buckets_for(uint32_t n);
fh_status fh_params_default(fh_params *p);
/* Device name, compute units and HBM bytes of device 0. */
fh_status fh_device_info(char *name, int name_cap, int *compute_units, size_t *hbm_bytes);
/* Resources of the kernel named `kernel` ("decide", "integrate", "scatter",
 * "hash_count", "scan"; "decide:BLOCK:G:K" for the decide instance of that
 * workgroup size and lane group over sorted copies (K = s), resident records
 * (r) or resident records with the fused flight (f)): VGPRs per lane, static
 * LDS bytes, scratch bytes. */
fh_status fh_kernel_info(const char *kernel, int *vgprs, int *lds_bytes, int *scratch_bytes);

/* ---- .flk files and generators (host only; state.rs:58-170) ------------- */
/* Replaces Flock::read (state.rs:135-170): words == NULL reads only the header. */
fh_status fh_read_flk(const char *path, uint32_t *n, float *time, float *words, uint32_t cap);
/* Replaces Flock::write (state.rs:104-132). */
fh_status fh_write_flk(const char *path, uint32_t n, float time, const float *words);
/* Flock::scatter (state.rs:79-101): the reference benchmark start (density
 * 0.5, +x heading, SplitMix64). */
fh_status fh_gen_reference(uint32_t n, uint64_t seed, float *words);
/* Uniform cube of edge `side` [m] centred on (0, 0, 100); directions uniform
 * on the sphere, speed 10 + U(-1, 1) m/s; radius 1.5; next ~ U[0, DU). */
fh_status fh_gen_cube(uint32_t n, float side, uint64_t seed, float *words);
/* `clumps` isotropic Gaussian clumps of std `sigma` [m] whose centres are
 * uniform in a cube of edge `cube` [m] about (0, 0, 100); each clump has a
 * heading uniform on the sphere, members fly it rotated by up to 10 degrees
 * (uniform in the cap), at 10 m/s. */
fh_status fh_gen_clumps(uint32_t n, uint32_t clumps, float sigma, float cube, uint64_t seed,
                        float *words);
/* The start state of a served request: fh_gen_cube with the cube edge
 * cell * cbrt(n / birds_per_cell), i.e. `birds_per_cell` birds per cell. */
fh_status fh_gen_request(uint32_t n, uint64_t seed, float birds_per_cell, float cell, float *words);

/* ---- one flock: the reference pass API ---------------------------------- */
/* Allocates the device state and grid for n birds (cuda/flock.cu:403-412). */
fh_status fh_create(uint32_t n, const fh_params *params, fh_flock **out);
fh_status fh_destroy(fh_flock *h);
fh_status fh_get_params(const fh_flock *h, fh_params *out, uint32_t *buckets);
/* Host -> device state (cuda/flock.cu:414-437). n must equal the handle's n. */
fh_status fh_upload_words(fh_flock *h, const float *words, uint32_t n, float time);
/* Device -> host state (cuda/flock.cu:494-521); *time gets the flock clock. */
fh_status fh_download_words(fh_flock *h, float *words, float *time);
/* grid_histogram (cuda/flock.cu:118-124), including the counts memset (:457). */
fh_status fh_hash_count(fh_flock *h);
/* grid_scan_blocks/totals/add (cuda/flock.cu:144-172): offsets[B+1]. */
fh_status fh_scan(fh_flock *h);
/* grid_scatter (cuda/flock.cu:175-179) into the sorted position/heading
 * records that decide walks. */
fh_status fh_scatter(fh_flock *h);
/* decide (cuda/flock.cu:190-290) for birds with time >= next_decision. */
fh_status fh_decide(fh_flock *h, float time);
/* integrate (cuda/flock.cu:293-328). */
fh_status fh_integrate(fh_flock *h);
/* integrate fused with the next step's grid_histogram. */
fh_status fh_integrate_hash(fh_flock *h);
/* `steps` whole steps in the order of cuda/flock.cu:452-470 (time += DT in
 * f32 first). The first `warmup` steps are not timed. out may be NULL. */
fh_status fh_step(fh_flock *h, uint32_t steps, uint32_t warmup, fh_pass_ms *out);
/* Waits for the handle's stream and reports any kernel-side error. */
fh_status fh_sync(fh_flock *h);
fh_status fh_get_time(const fh_flock *h, float *time);

/* ---- one flock without allocation --------------------------------------- */
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): steps the n birds whose .flk records (FH_WORDS floats
 * each) the caller holds in DEVICE memory at d_words, in place, `steps` times
 * from clock `time`, with the launches of fh_step's fused standard path in
 * canonical order; *time_out (host, may be NULL) gets the new clock. The
 * result equals fh_create + fh_upload_words + fh_step + fh_download_words with
 * the same tunables and FH_FLAG_CANONICAL, bit for bit. Launches go to
 * `stream` (a hipStream_t; NULL = the null stream) and are not waited for.
 * Every work array lives in one library-held device array per GPU (about
 * 8 MiB, held until the process ends): no hipMalloc, hipFree or memset is
 * made, and FH_ERR_INVALID is returned for n > FH_NOALLOC_MAX_BIRDS or more
 * than 2^17 buckets. Calls that overlap on one GPU share that array and race.
 * Flags other than FH_FLAG_CANONICAL, _WAVE_AGG, _NO_TILE, _OCC, _SPATIAL,
 * _XCD and _WINDOW are refused (FH_ERR_INVALID). Kernel-side error bits are
 * not read back. */
#define FH_NOALLOC_MAX_BIRDS 65536u
fh_status fh_step_device_noalloc(float *d_words, uint32_t n, float time, uint32_t steps, const fh_params *params,
                                 void *stream, float *time_out);

/* ---- debug read-back for parity ----------------------------------------- */
/* Grid of the last build (any pointer may be NULL): keys[n], counts[B],
 * offsets[B+1], sorted[n] (bird indices in slot order). After fh_scan the
 * counts are offsets[b+1] - offsets[b]. */
fh_status fh_read_grid(fh_flock *h, uint32_t *keys, uint32_t *counts, uint32_t *offsets,
                       uint32_t *sorted);
/* Neighbours of `bird` found by the device grid walk: walk 1 (d^2 <= r^2) or
 * walk 2 (d^2 <= 4 r^2), ascending bird index. *count may exceed cap. */
fh_status fh_neighbours(fh_flock *h, uint32_t bird, int walk, uint32_t *ids, uint32_t cap,
                        uint32_t *count);

/* ---- many flocks, one launch per stage ----------------------------------- */
/* No counterpart in the reference, which steps one flock per process
 * (cuda/flock.cu:374-523): independent flocks share one set of device pools
 * and every pass of a step runs once for all of them. Each flock keeps its own
 * clock, buckets (its own buckets_for(n)) and results; with
 * FH_FLAG_CANONICAL a flock's result is bit-identical to stepping it alone. */
typedef struct fh_batch fh_batch;
/* Capacities: total birds, total buckets, and flock slots. */
fh_status fh_batch_create(uint32_t bird_capacity, uint32_t bucket_capacity, uint32_t flock_capacity,
                          const fh_params *params, fh_batch **out);
fh_status fh_batch_destroy(fh_batch *b);
/* Adds a flock (n > 0 .flk records at clock `time`); FH_ERR_BUDGET when the
 * pools or slots are full. In a resident batch that has stepped, the flock
 * joins the layout in place: the records of the flocks after it move by n
 * slots. */
fh_status fh_batch_add(fh_batch *b, const float *words, uint32_t n, float time, uint32_t *flock_id);
/* Removes a flock; in a resident batch its records leave the layout in place. */
fh_status fh_batch_remove(fh_batch *b, uint32_t flock_id);
/* The flock's .flk records (in its bird order) and clock. */
fh_status fh_batch_download(fh_batch *b, uint32_t flock_id, float *words, float *time);
/* `steps` steps of every live flock (time += DT each), except the flocks of a
 * resident batch whose step budget is used up. */
fh_status fh_batch_step(fh_batch *b, uint32_t steps, uint32_t warmup, fh_pass_ms *out);
/* Resident batches (FH_FLAG_RESIDENT): the flock takes at most `steps` more
 * steps of the coming fh_batch_step calls; after them it is frozen (state and
 * clock kept) until its budget is set again or it is removed, so a serving
 * loop can step many flocks of different lengths together and collect them
 * once in a while. FH_STEPS_UNLIMITED lifts the budget (the default).
 * FH_ERR_INVALID on a batch without FH_FLAG_RESIDENT. No counterpart in the
 * reference, which steps one flock per process. */
#define FH_STEPS_UNLIMITED 0xffffffffu
fh_status fh_batch_limit(fh_batch *b, uint32_t flock_id, uint32_t steps);
fh_status fh_batch_info(const fh_batch *b, uint32_t *live_flocks, uint32_t *live_birds, uint32_t *bucket_range,
                        size_t *device_bytes);
/* Device bytes one flock of n birds occupies in a batch (state, sorted copies,
 * grid), used for admission against an HBM budget. Host only. */
size_t fh_flock_bytes(uint32_t n, const fh_params *params);

/* ---- the served loop: continuous batching of flock requests --------------- */
/* No counterpart in the reference. Requests (birds, steps, seed) arrive at any
 * time; a worker thread keeps admitted flocks resident in one fh_batch, admits
 * waiting requests in arrival order whenever a slot and HBM bytes (by
 * fh_flock_bytes) are free, steps every active flock together and retires
 * each one when its step count is reached. A request's start state is
 * fh_gen_cube(birds, cell * (birds / birds_per_cell)^(1/3), seed). */
typedef struct fh_server_config {
    double hbm_budget_bytes;   /* admission budget [bytes]; default 200e9 */
    uint32_t max_active;       /* flocks resident at once (slot cap); default 64 */
    uint32_t bird_capacity;    /* device pool size in birds; default 40 Mi */
    uint32_t bucket_capacity;  /* device pool size in buckets; default 64 Mi */
    uint32_t steps_per_iter;   /* batched steps between admissions; default 8 */
    uint32_t prep_threads;     /* host threads generating start states; default 4 */
    uint32_t keep_states;      /* keep each final state for fh_server_take_state */
    float birds_per_cell;      /* request density; default 0.64 */
    float cell_for_density;    /* cell edge the density refers to [m]; default 4 */
    fh_params params;          /* step tunables; default canonical order */
} fh_server_config;

typedef struct fh_result {
    uint64_t id, seed;
    uint32_t birds, steps;
    int32_t status;            /* FH_OK or the error that ended the request */
    float final_time;          /* flock clock at completion */
    double arrival_s, admit_s, done_s;  /* seconds since the server started */
    uint64_t checksum;         /* FNV-1a 64 over the 32-bit words of the final .flk records */
} fh_result;

typedef struct fh_server_stats {
    uint64_t submitted, iterations, steps;
    double bird_steps;         /* sum over iterations of active birds x steps */
    double busy_s;             /* time inside batched steps */
    double wall_s;             /* time since the server started */
    double peak_bytes;         /* largest admitted HBM bytes */
    uint32_t peak_active, queued, active;
    double admit_s;            /* GPU-idle time spent adding admitted flocks to the batch */
    double retire_s;           /* GPU-idle time spent downloading and removing finished flocks */
    double gap_s;              /* resident batches: all GPU-idle time between step batches (admit_s +
                                  retire_s + the loop's own work); with FH_FLAG_RESIDENT, start states
                                  are staged in pinned memory ahead of admission and copies run on a
                                  second stream while steps run, so admissions and downloads overlap
                                  the steps and only one splice and scan per iteration is idle time */
} fh_server_stats;

typedef struct fh_server fh_server;

fh_status fh_server_config_default(fh_server_config *c);
fh_status fh_server_create(const fh_server_config *cfg, fh_server **out);
/* Stops the worker (active and queued requests are dropped) and frees all. */
fh_status fh_server_destroy(fh_server *s);
/* FH_ERR_INVALID for birds == 0 or steps == 0; FH_ERR_BUDGET when one request
 * alone exceeds the budget or the pool. */
fh_status fh_server_submit(fh_server *s, uint32_t birds, uint32_t steps, uint64_t seed, uint64_t *id);
/* Moves up to cap completed results to out. */
fh_status fh_server_poll(fh_server *s, fh_result *out, uint32_t cap, uint32_t *count);
/* Waits until nothing is queued or active. */
fh_status fh_server_drain(fh_server *s, double timeout_s);
fh_status fh_server_stats_get(fh_server *s, fh_server_stats *out);
/* Final .flk records of a completed request (keep_states), once. */
fh_status fh_server_take_state(fh_server *s, uint64_t id, float *words, uint32_t cap, float *time);
/* Open-loop replay: request i is submitted arrival_s[i] seconds after start;
 * returns when all are done. results[i] belongs to request i. */
fh_status fh_server_replay(const fh_server_config *cfg, const uint32_t *birds, const uint32_t *steps,
                           const uint64_t *seeds, const double *arrival_s, uint32_t count,
                           fh_result *results, fh_server_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
