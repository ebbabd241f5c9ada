/*
 * physim_gpu: MI355X (gfx950) gravity and integration engine behind the
 * physim element API. Plain C types only; every call returns a pg_status
 * and pg_last_error() holds the message of the last failure on the calling
 * thread.
 *
 * Reference interfaces replaced (the reference's hot path is single-threaded
 * FP64 Rust on the CPU):
 *   - pg_transform_aos            AstroOctreeElement::transform /
 *                                 AstroElement::transform /
 *                                 SimpleAstroElement::transform
 *                                 (astro/src/transformers.rs:123-160,32-69,219-244),
 *                                 called through the C-ABI transform vtable
 *                                 (physim-core/src/plugin/transform.rs:85-104).
 *   - pg_sim_step                 one pipeline iteration of
 *                                 synths(none) -> Verlet::integrate(acc_fn =
 *                                 that transform) -> state = new_state
 *                                 (physim-core/src/pipeline.rs:143-182,
 *                                 integrators/src/verlet.rs:85-108), with the
 *                                 state resident on the device.
 *   - pg_sim_upload / _download   the Vec<Entity> hand-over of
 *                                 pipeline.rs:100-107,173,179.
 *   - pg_gen_cube / _star /       RandomCube / SingleStar / Plummer
 *     _plummer_ref                (astro/src/initialisers.rs:81-107,145-191,293-330).
 * Everything else (batching, serving, virial Plummer, energy) has no
 * reference counterpart and says so where declared.
 */
#ifndef PHYSIM_GPU_H
#define PHYSIM_GPU_H

#include <stddef.h>
#include <stdint.h>

#include "physim_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pg_status {
    PG_OK = 0,
    PG_ERR_HIP = 1,        /* a HIP runtime call failed */
    PG_ERR_ARG = 2,        /* invalid argument */
    PG_ERR_NOMEM = 3,      /* device allocation failed */
    PG_ERR_BUDGET = 4,     /* request exceeds the engine's HBM budget */
    PG_ERR_DEPTH = 5,      /* two bodies share every tree level the key holds */
    PG_ERR_COINCIDENT = 6, /* bodies within 1e-9 on every axis (reference merge rule) */
    PG_ERR_NONFINITE = 7,  /* NaN or infinite coordinate */
    PG_ERR_STATE = 8,      /* call out of order */
    PG_ERR_INTERNAL = 9,   /* a device-side capacity was exceeded */
} pg_status;

typedef enum pg_kind {
    PG_KIND_DIRECT = 0,   /* simple_astro: direct O(N^2) sum */
    PG_KIND_QUADTREE = 2, /* astro: Barnes-Hut, 4-way (x, y) tree */
    PG_KIND_OCTREE = 3,   /* astro2: Barnes-Hut, 8-way tree */
} pg_kind;

typedef enum pg_precision {
    PG_PREC_MIXED = 0, /* FP32 pair forces, compensated FP32 per-body sums,
                          FP64 positions/differences/tree/criterion/integrator */
    PG_PREC_FP64 = 1,  /* the reference's FP64 arithmetic throughout */
} pg_precision;

typedef struct pg_params {
    int32_t kind;      /* pg_kind */
    int32_t precision; /* pg_precision */
    double theta;      /* opening parameter (reference default 1.0) */
    double e;          /* softening added to r^2 (reference default 1.0; abs taken) */
    double dt;         /* Stormer time step */
} pg_params;

typedef struct pg_engine pg_engine;
typedef struct pg_sim pg_sim;

const char* pg_last_error(void);
const char* pg_version(void);

/* ---- engine ----------------------------------------------------------- */
/* One engine per device; owns one HIP stream and the shared scratch of the
 * batched step. hbm_budget_bytes = 0 means 90% of the device's memory. */
int pg_engine_create(int device, uint64_t hbm_budget_bytes, pg_engine** out);
void pg_engine_destroy(pg_engine* eng);
/* Bytes of device memory the engine holds / its budget. */
int pg_engine_memory(pg_engine* eng, uint64_t* used, uint64_t* budget);
/* Device memory a simulation of n bodies needs (state + its share of
 * scratch), used for admission against the budget. */
uint64_t pg_sim_bytes(uint64_t n, int32_t kind);
/* Capture repeated steps in hipGraphs (default on). */
int pg_engine_set_graphs(pg_engine* eng, int enable);
/* Key sorting of the tree build: a sample sort into at most max_buckets
 * buckets (<= 8192; a larger batch, or max_buckets = 0, uses the 8-bit LSD
 * radix sort), each sorted in LDS when it holds at most bucket_cap keys
 * (<= 4096) and in HBM otherwise. Both give the same order. Defaults 8192,
 * 4096. A tuning and testing control; no reference counterpart. */
int pg_engine_set_sort(pg_engine* eng, uint32_t max_buckets, uint32_t bucket_cap);
/* Launch choices that change the time of a step but not its results (the
 * sort's order, the tree and every sum are the same under all of them). No
 * reference counterpart. */
typedef struct pg_tuning {
    uint32_t max_buckets;   /* sample sort up to this many buckets (<= 8192; 0: radix sort) */
    uint32_t bucket_keys;   /* mean keys per sample-sort bucket (256..4096; default 1664) */
    uint32_t walk_run_log2; /* the walk's blocks go to each XCD in runs of 2^this (default 4) */
    uint32_t lds_buckets;   /* up to this many buckets every key tile holds the splitters in
                               LDS; more: in HBM behind a coarse table (<= 1024; 0: 1024) */
} pg_tuning;
int pg_engine_set_tuning(pg_engine* eng, const pg_tuning* tuning);
int pg_engine_get_tuning(pg_engine* eng, pg_tuning* tuning);
/* The autotuner: runs the batch of `sims` (verlet) under each candidate
 * tuning from the same saved state -- four warm-up steps, then trial_steps
 * timed steps -- keeps the fastest in the engine and restores the sims'
 * state (their results are unchanged). *best receives the choice; tried[k]
 * and ms_per_step[k] (up to cap entries; may be NULL) each candidate and its
 * time; *ntried the number of candidates. */
int pg_engine_autotune(pg_engine* eng, pg_sim* const* sims, uint32_t count, uint32_t trial_steps,
                       pg_tuning* best, pg_tuning* tried, double* ms_per_step, uint32_t cap,
                       uint32_t* ntried);
/* Per-kernel HIP-event timing of every launch (eager mode, no graphs). */
int pg_engine_set_profiling(pg_engine* eng, int enable);
/* Accumulated per-kernel times. names: '\n'-separated kernel names
 * written into `names` (cap bytes); ms[i], launches[i] per kernel. Returns
 * the kernel count in *count. */
int pg_engine_profile_read(pg_engine* eng, char* names, size_t cap, double* ms,
                           uint64_t* launches, uint32_t max_kernels, uint32_t* count);
int pg_engine_profile_reset(pg_engine* eng);
int pg_engine_synchronize(pg_engine* eng);
/* A plain streaming copy of `bytes` device bytes, `reps` times; *gbps
 * receives (read + write) bytes per second of the timed launches. A
 * calibration probe for memory counters and the achievable-bandwidth
 * figure; no reference counterpart. */
int pg_engine_copy_probe(pg_engine* eng, uint64_t bytes, uint32_t reps, double* gbps);

/* ---- simulations ------------------------------------------------------ */
int pg_sim_create(pg_engine* eng, uint64_t n, const pg_params* params, pg_sim** out);
void pg_sim_destroy(pg_sim* sim);
int pg_sim_set_params(pg_sim* sim, const pg_params* params);
/* Host AoS -> device SoA. Clears the Stormer history, so the next step is
 * the Taylor first step (verlet.rs:102-103). */
int pg_sim_upload(pg_sim* sim, const Entity* state, uint64_t n);
/* Host AoS -> device SoA with the Stormer history of the reference's
 * verlet: `previous` (may be NULL: as pg_sim_upload) is the state the
 * integrator was given on its previous call, whose positions the next step
 * takes as x_prev (verlet.rs:52-82,104-107: the history survives an edited
 * state of the same length). */
int pg_sim_upload_state(pg_sim* sim, const Entity* state, const Entity* previous, uint64_t n);
/* A checkpoint of a verlet simulation: its state (as pg_sim_download) and
 * its Stormer history as `previous` (the records with the previous
 * positions; may be NULL), *has_history = 0 before the first step (previous
 * then equals state). pg_sim_upload_state(sim, state, previous) resumes it:
 * the resumed run continues exactly as the checkpointed one (verlet.rs:
 * 52-82 needs x_prev, which a plain download drops). */
int pg_sim_checkpoint(pg_sim* sim, Entity* state, Entity* previous, uint64_t n,
                      int* has_history);
/* Device -> host AoS in upload order; radius, id and fixed are those
 * uploaded, as verlet copies them (verlet.rs:41,72). */
int pg_sim_download(pg_sim* sim, Entity* out, uint64_t n);
/* Advance `steps` pipeline iterations on the device. */
int pg_sim_step(pg_sim* sim, uint32_t steps);

typedef enum pg_integrator {
    PG_INTEG_VERLET = 0, /* Stormer verlet (integrators/src/verlet.rs:24-108), the default */
    PG_INTEG_EULER = 1,  /* Euler::integrate (integrators/src/euler.rs:17-49) */
    PG_INTEG_RK4 = 2,    /* Rk4::integrate (integrators/src/rk4.rs:24-182): four
                            acceleration passes per step */
} pg_integrator;
/* The integrator pg_sim_step applies: the reference's `verlet`, `euler` or
 * `rk4` element (IntegratorElement::integrate, physim-core/src/plugin/
 * integrator.rs:5-13) with acc_fn = the sim's transform. A change starts
 * the new integrator fresh from the current positions and velocities (a new
 * verlet begins with its Taylor step). pg_batch_step and the server take
 * verlet simulations. */
int pg_sim_set_integrator(pg_sim* sim, int32_t integrator);
/* Accelerations of the current device state (no integration), FP64, n*3
 * values in upload order. */
int pg_sim_accel(pg_sim* sim, double* acc_xyz);
/* FP64 kinetic, potential (force-consistent, O(N^2)), total energy and
 * momentum of the current device state: out = {K, V, E, Px, Py, Pz}. No
 * reference counterpart (the reference's energysink is not force-consistent). */
int pg_sim_energy(pg_sim* sim, double out[6]);

typedef struct pg_stats {
    uint64_t steps_done;
    uint64_t nodes;        /* tree nodes of the last build */
    uint64_t interactions; /* interaction-list entries of the last force pass */
    uint32_t max_level;    /* deepest leaf level of the last build */
    uint32_t key_levels;   /* tree levels the keys resolve: two bodies share at most key_levels - 1 */
    uint32_t error;        /* sticky device-side pg_status of this sim */
    uint32_t reserved;
} pg_stats;
int pg_sim_stats(pg_sim* sim, pg_stats* out);
/* Direct-sum near field (no reference counterpart; 0, the default, is the
 * reference's walk): a tree cell at the scale of the body's own leaf cell
 * (that level and the two below) is opened rather than taken whole when the
 * body lies within leaf_widths widths of its leaf of the cell's box, so the
 * bodies that close contribute their pair forces directly. A deviation from
 * astro2's interaction lists, for the colliding-spheres precision study (2
 * leaf widths). */
int pg_sim_set_near_field(pg_sim* sim, double leaf_widths);
/* Count interactions per body in the force pass (debug; off by default). */
int pg_sim_set_counting(pg_sim* sim, int enable);
/* Per-body interaction-list length of the last force pass (upload order). */
int pg_sim_interaction_counts(pg_sim* sim, int64_t* counts);
/* Tree of the last build in preorder, children in octant order:
 * rec[4k..] = {level, octant in parent (-1 root), leaf body (upload index) or
 * -1, first sorted position}; fl[8k..] = {mass, com x, y, z, centre x, y, z,
 * half-width}. *count receives the node count (arrays may be NULL). */
int pg_sim_tree(pg_sim* sim, int64_t* rec, double* fl, uint64_t cap, uint64_t* count);
/* Morton keys of the last build in sorted order and the body permutation
 * (upload index of each sorted position). */
int pg_sim_morton(pg_sim* sim, uint64_t* keys, uint32_t* perm, uint64_t n);

/* ---- batching (no reference counterpart) ------------------------------- */
/* Advance several simulations together: one launch sequence per step covers
 * all of them. All must share kind and precision; theta, e, dt and n may
 * differ. Each result is bit-identical to stepping that simulation alone. */
int pg_batch_step(pg_engine* eng, pg_sim* const* sims, uint32_t count, uint32_t steps);

/* ---- transform (plugin path) ------------------------------------------- */
/* acc[i] += acceleration of body i under params (kind, theta, e; dt unused).
 * Host AoS in, host AoS out; the device holds nothing between calls except
 * reusable scratch. */
int pg_transform_aos(pg_engine* eng, const pg_params* params, const Entity* state, uint64_t n,
                     Acceleration* acc);
/* pg_transform_aos without device allocation: no counterpart in the
 * reference. For callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): every device buffer
 * of the call lies in the library's static device array of 48 MiB per GPU,
 * so n is at most 65,536 bodies (PG_ERR_ARG beyond) and a tree at most 4
 * nodes per body (PG_ERR_INTERNAL beyond). Same arguments, results and error
 * behaviour as pg_transform_aos, which it equals bit for bit. Calls on one
 * GPU share the array and are serialised; it keeps the last call's contents. */
int pg_transform_aos_noalloc(pg_engine* eng, const pg_params* params, const Entity* state,
                             uint64_t n, Acceleration* acc);

/* ---- serving (no reference counterpart) -------------------------------- */
typedef struct pg_server pg_server;
typedef struct pg_request_info {
    uint64_t id;
    uint64_t n;
    uint32_t steps;
    uint32_t status;       /* pg_status of the request */
    double submit_s;       /* engine clock, seconds */
    double admit_s;
    double complete_s;
} pg_request_info;
/* max_active: simulations stepped together; 0 = unlimited (budget only). */
int pg_server_create(pg_engine* eng, uint32_t max_active, pg_server** out);
void pg_server_destroy(pg_server* srv);
/* Queue a whole simulation; copies `state`. */
int pg_server_submit(pg_server* srv, const Entity* state, uint64_t n, const pg_params* params,
                     uint32_t steps, uint64_t* request_id);
/* Admit queued requests that fit, advance every active simulation by up to
 * `max_steps` co-batched steps, retire finished ones. *active / *queued
 * receive the counts after the tick. */
int pg_server_tick(pg_server* srv, uint32_t max_steps, uint32_t* active, uint32_t* queued);
/* Result of a completed request (state after its steps); frees it. */
int pg_server_take(pg_server* srv, uint64_t request_id, Entity* out, uint64_t n,
                   pg_request_info* info);
int pg_server_info(pg_server* srv, uint64_t request_id, pg_request_info* info);
double pg_server_clock(pg_server* srv);

/* ---- the verlet element on the device ---------------------------------- */
/* The C side of a `verlet` integrator element that steps on the device:
 * Verlet::integrate (integrators/src/verlet.rs:85-108) behind the
 * IntegratorElement interface (physim-core/src/plugin/integrator.rs:5-13),
 * with acc_fn = the bound GPU gravity element's transform. A Rust shim keeps
 * the element name `verlet`; the GPU element announces itself on the bus
 * (topic "physim_gpu", message "transform", sender_id = the element), and
 * the shim binds it. Each integrate call steps the device state once and
 * fills new_state (every field). The device state is reused while the input
 * equals, field for field, what the last call returned; an input changed by
 * a synth or transmute is uploaded with the last call's input as the Stormer
 * history (a Taylor restart only when the length changes, as the reference).
 * Without a bound element integrate returns PG_ERR_STATE (the shim then runs
 * the stock verlet with the pipeline's acc_fn). */
typedef struct pg_verlet pg_verlet;
int pg_verlet_create(pg_verlet** out);
void pg_verlet_destroy(pg_verlet* v);
/* element: a transform element object of this library (its init's return). */
int pg_verlet_bind(pg_verlet* v, const void* element);
int pg_verlet_integrate(pg_verlet* v, const Entity* entities, Entity* new_state, uint64_t n,
                        double dt);

/* ---- device-resident pipeline runner ---------------------------------- */
/* Pipeline::run's simulation loop (physim-core/src/pipeline.rs:134-192) for
 * a pipeline of one GPU gravity transform (params) and an integrator
 * (pg_integrator), the state resident on the device, with csvsink renderers
 * (utilities/src/csvsink.rs:50-79: the first state and every print_n-th,
 * "x,y,z," per entity in Rust's shortest f64 Display, one line per frame).
 * A state crosses PCIe only when a sink prints it; the renderer formats on
 * its own thread behind a two-frame queue (pipeline.rs:127's
 * sync_channel(2)). Synths and transmutes are not run on the device. */
typedef struct pg_runner pg_runner;
int pg_runner_create(pg_engine* eng, const pg_params* params, int32_t integrator,
                     const Entity* initial, uint64_t n, pg_runner** out);
/* csvsink properties file (default "csvsink.csv", truncated now) and print_n
 * (default 1 when 0). Sinks are added before the first run. */
int pg_runner_add_csvsink(pg_runner* r, const char* file, uint32_t print_n);
/* `iterations` more pipeline iterations (the initial state is handed to the
 * sinks on the first call). */
int pg_runner_run(pg_runner* r, uint32_t iterations);
/* Waits until every handed-over frame is written. */
int pg_runner_flush(pg_runner* r);
int pg_runner_state(pg_runner* r, Entity* out, uint64_t n);
void pg_runner_destroy(pg_runner* r);
/* Rust's `{}` (Display) of an f64 into out (NUL-terminated, cap bytes);
 * returns its length. csvsink's number format. */
size_t pg_format_f64(double v, char* out, size_t cap);

/* ---- initial conditions ------------------------------------------------ */
/* RandomCube::create_entities (astro/src/initialisers.rs:81-107). */
int pg_gen_cube(uint64_t n, uint64_t seed, double spin, double size, const double centre[3],
                double mass, uint64_t id, Entity* out);
/* SingleStar (astro/src/initialisers.rs:145-191). */
int pg_gen_star(double x, double y, double z, double vx, double vy, double vz, double radius,
                double mass, uint64_t id, int fixed, Entity* out);
/* Plummer::create_entities (astro/src/initialisers.rs:293-330). */
int pg_gen_plummer_ref(uint64_t n, uint64_t seed, double mass, double a, const double centre[3],
                       const double v[3], double spin, uint64_t id, Entity* out);
/* Isotropic Plummer sphere in virial equilibrium (Plummer 1911 radii,
 * Aarseth-Henon-Wielen 1974 speeds), centre-of-mass frame shifted to
 * centre / v. No reference counterpart. */
int pg_gen_plummer_virial(uint64_t n, uint64_t seed, double mass, double a,
                          const double centre[3], const double v[3], uint64_t id, Entity* out);

#ifdef __cplusplus
}
#endif

#endif /* PHYSIM_GPU_H */
