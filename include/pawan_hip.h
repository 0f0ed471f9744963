/*
 * pawan_hip.h -- C ABI of the MI355X (gfx950) vortex-particle interaction and
 * RK4 integration path.
 *
 * The entry points replace the compute behind pawan's C++ classes
 * (namespace pawan):
 *   __interaction  src/interaction/interaction.h:21-235
 *   __parallel     src/interaction/parallel.h:20-145
 *   __integration  src/integration/integration.h:20-58
 *   __rk4          src/integration/rk4.h:14-41
 * A C++ subclass of __interaction / __rk4 (see the adapter header
 * pawan_gpu.hpp) forwards its virtuals here; nothing else in the driver
 * changes.
 *
 * Conventions
 *   - Plain pointers and sizes; no GSL or torch types.  Particle arrays are
 *     the reference's storage: positions/vorticities N x 3 row-major doubles
 *     (gsl_matrix data with tda = 3), radius/volume N doubles.
 *   - State and rate vectors use the reference layout: for each wake in the
 *     order added, [positions 3n | vorticities 3n] (wake.cpp:87-120).
 *   - Every function returns PW_OK (0) or a PW_E* code; pw_last_error()
 *     gives the message.  No C++ exception crosses the ABI.
 *   - Host arrays are borrowed for the duration of the call only.
 *   - A context owns one HIP stream and all of its device memory; use it
 *     from one host thread at a time.  Distinct contexts are independent.
 *   - Results are deterministic: each particle's sums are taken in an order
 *     that depends only on its own simulation, never on other work sharing
 *     the GPU.
 */
#ifndef PAWAN_HIP_H_
#define PAWAN_HIP_H_

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define PW_OK 0
#define PW_EINVAL 1     /* bad argument or call order                        */
#define PW_ENOMEM 2     /* device or host allocation failed / over budget    */
#define PW_EHIP 3       /* HIP runtime error                                 */
#define PW_ENONFINITE 4 /* a rate or state became inf/NaN (coincident pairs) */
#define PW_ENOTFOUND 5  /* unknown job id                                    */

/* smoothing kernel: interaction_utils.h:13-24 selects one at compile time */
#define PW_KERNEL_HIGHORDER 0 /* highorder.h:20-60 (reference default)  */
#define PW_KERNEL_GAUSSIAN 1  /* gaussian.h:20-58                      */

/* arithmetic of the pair evaluation */
#define PW_FP64 0  /* binary64 throughout (the reference method)              */
#define PW_FP32C 1 /* binary32 pair terms, compensated per-target sums,
                      binary64 cross-tile reduction, state and diagnostics    */

/* how the pair pass visits pairs */
#define PW_SCHEME_DEFAULT 0   /* binary32 symmetric, binary64 one-sided (the
                                 plain method: a thread per target)            */
#define PW_SCHEME_ONESIDED 1  /* every ordered pair, per target (no source-side
                                 sums), both precisions                        */
#define PW_SCHEME_SYMMETRIC 2 /* every unordered pair once, both particles
                                 updated, as the reference's INTERACT does
                                 (interaction_utils.h:429-481); deterministic
                                 per-block source-side slots; both precisions  */

typedef struct pw_config {
    int device;       /* HIP device ordinal                                      */
    int kernel;       /* PW_KERNEL_*                                             */
    int precision;    /* PW_FP64 / PW_FP32C                                      */
    double nu;        /* kinematic viscosity (reference: 2.5e-3, interaction.cpp:13) */
    double uinf[3];   /* uniform freestream added to every velocity (0 = none)     */
    int block;        /* threads per workgroup 64/128/256; 0 = tuned or default;
                         512 for the binary32 symmetric HIGHORDER pass at one
                         target per thread (half the source-side slot rows; the
                         default for such a context past 524,288 particles)    */
    int chunk;        /* sources per work item (multiple of block); 0 = policy   */
    int targets_per_thread; /* 1 or 2; 0 = tuned or default                      */
    int unroll;       /* source-loop unroll 1/2/4; 0 = tuned or default          */
    int scheme;       /* PW_SCHEME_*: how the pair pass visits pairs             */
    int chunks;       /* source chunks per job when chunk is 0 (0 = policy or the
                         tuning cache's); symmetric pass: also the strip count */
} pw_config;

/* A compiled kernel variant of the pair pass. */
typedef struct pw_variant {
    int block;              /* 64, 128 or 256 threads per workgroup; 512: binary32
                               symmetric HIGHORDER jobs at one target per thread
                               (binary64 jobs of that engine run at 256)       */
    int targets_per_thread; /* 1 or 2                                             */
    int unroll;             /* 1, 2 or 4; applies to the precision being tuned    */
    int scheme;             /* PW_SCHEME_* (0: default)                            */
    int chunks;             /* source chunks per job (0: the chunk policy, at most
                               16 one-sided, max(128, blocks / 8) symmetric)    */
} pw_variant;

typedef struct pw_ctx pw_ctx;

/* library identification: "pawan_hip <version> gfx950" */
const char *pw_version(void);
/* fill cfg with defaults: device 0, HIGHORDER, FP64, nu 2.5e-3, no freestream */
int pw_config_default(pw_config *cfg);

/* __interaction() (interaction.cpp:11-18) */
int pw_create(pw_ctx **out, const pw_config *cfg);
/* ~__interaction() (interaction.cpp:29-33) */
int pw_destroy(pw_ctx *ctx);
/* message for the last failing call on ctx (or on ctx == NULL: last global) */
const char *pw_last_error(const pw_ctx *ctx);

/* __interaction::_nu (interaction.h:24); the reference has no setter */
int pw_set_nu(pw_ctx *ctx, double nu);

/* __interaction::addWake (interaction.cpp:35-39).  Copies the particles of
 * one wake (__wake::_position/_vorticity/_radius/_volume, wake.h:35-40) to
 * the device.  Wakes interact with themselves and with each other, as
 * interact() does (interaction.cpp:95-108). */
int pw_add_wake(pw_ctx *ctx, size_t n, const double *pos, const double *vort,
                const double *radius, const double *volume);
/* __system::_size: length of the state vector, 6 * total particles */
size_t pw_size(const pw_ctx *ctx);
size_t pw_num_wakes(const pw_ctx *ctx);

/* __interaction::setStates / getStates (interaction.cpp:204-229) */
int pw_set_state(pw_ctx *ctx, const double *state);
int pw_get_state(pw_ctx *ctx, double *state);
/* __interaction::solve -> interact() (interaction.cpp:41-43,95-108):
 * velocity and vorticity rate of every particle, kept on the device */
int pw_solve(pw_ctx *ctx);
/* __interaction::getRates (interaction.cpp:213-220): [vel 3n | dvort 3n] per wake */
int pw_get_rates(pw_ctx *ctx, double *rates);
/* one wake's __wake::_position/_vorticity and _velocity/_retvorcity (N x 3) */
int pw_get_wake(pw_ctx *ctx, size_t wake, double *pos, double *vort, double *vel,
                double *dvort);

/* __rk4::step (rk4.cpp:12-77) repeated nsteps times on device-resident state */
int pw_rk4_steps(pw_ctx *ctx, double dt, size_t nsteps);
/* __integration::step (integration.cpp:37-45), forward Euler */
int pw_euler_steps(pw_ctx *ctx, double dt, size_t nsteps);

/* calculateTotalVorticity / calculateLinearImpulse (interaction.cpp:240-259)
 * and calculateKineticEnergy (interaction.cpp:351-366) over all particles;
 * any pointer may be NULL to skip that quantity (ke == NULL skips the O(N^2)
 * pass) */
int pw_diagnostics(pw_ctx *ctx, double omega[3], double impulse[3], double *ke);

/* Everything __interaction::diagnose() reports (interaction.cpp:49-93), plus
 * the kinetic energy it leaves out: total vorticity, linear and angular
 * impulse (:240-275), helicity (:317-349), enstrophy (:277-315), kinetic
 * energy (:351-366), over all particles. */
typedef struct pw_diag {
    double total_vorticity[3];
    double linear_impulse[3];
    double angular_impulse[3];
    double helicity;
    double enstrophy;
    double kinetic_energy;
} pw_diag;
int pw_diagnose(pw_ctx *ctx, pw_diag *out);

/* __interaction::resolve -> influence() (interaction.cpp:45-47,110-155):
 * the vorticity field at every particle, N x 3 with wakes in order
 * (each wake's __wake::_vorticityfield) */
int pw_resolve(pw_ctx *ctx, double *field);

/* __interaction::write (interaction.cpp:197-202): nWake (size_t), then per
 * wake N (size_t), positions N x 3, vorticities N x 3, radii N, all binary64
 * (wake.cpp:68-73); integrate() writes the time before each frame */
int pw_write(pw_ctx *ctx, FILE *f);

/* Asynchronous frames.  No reference function of its own: it replaces the
 * per-step "fwrite(&t); S->write(f)" of __integration::integrate
 * (integration.cpp:18-29), which stalls the stepping on the copy and the file
 * write.  pw_write_async snapshots the current state on the device, after
 * every step already queued, and returns; a copy stream moves the snapshot
 * into one of two pinned host buffers and a writer thread appends to f the
 * time t (when with_time != 0, as integrate writes it) followed by exactly
 * the bytes pw_write would write for that state.  Frames reach f in call
 * order; with two frames in flight a third call waits for the older one.
 * The caller must not touch f between pw_write_async and the pw_write_wait
 * that follows it.  A failed frame (fwrite short) is reported by the next
 * pw_write_async (which then queues nothing) or pw_write_wait. */
int pw_write_async(pw_ctx *ctx, FILE *f, double t, int with_time);
/* wait until every frame queued by pw_write_async is in its file (no fflush);
 * returns the first failure since the last wait */
int pw_write_wait(pw_ctx *ctx);

/* Particle emission: append n particles to wake `wake` (its _numParticles
 * grows; the state vector grows by 6n).  Device state of the particles
 * already present is kept.  No reference counterpart (__wake has a fixed
 * particle count); the growing rotor-wake workload needs it. */
int pw_append_particles(pw_ctx *ctx, size_t wake, size_t n, const double *pos, const double *vort,
                        const double *radius, const double *volume);

/* The pair-pass variant the context runs for its precision (builds the
 * device job if needed): threads per workgroup, targets per thread, unroll
 * (0 on the symmetric pass) and PW_SCHEME_ONESIDED / PW_SCHEME_SYMMETRIC.
 * No reference counterpart: reports what pw_config's zero fields resolved to
 * (defaults, the tuning cache). */
int pw_get_variant(pw_ctx *ctx, pw_variant *out);

/* device time of the last pw_solve / pw_rk4_steps / pw_euler_steps call,
 * measured with HIP events on the context's stream: total ms and the summed
 * ms of the pair-interaction kernel alone (0 after pw_rk4_steps on a context
 * of at most 2^30 pair interactions per RHS, about 32,768 particles, whose
 * steps replay as one captured graph each), plus the pairs it evaluated */
int pw_last_timing(const pw_ctx *ctx, double *total_ms, double *pair_ms, double *pairs);

/* ------------------------------------------------------ no allocation
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the rates of one wake, as pw_solve + pw_get_rates give
 * them for a context holding that wake (binary64, one-sided pass), with
 * every intermediate (particle records, partial sums, work tables, counters)
 * in a fixed array of this library in each device's memory
 * (pw_noalloc_scratch_bytes(), under 32 MiB, held until the process ends).
 *   pos, vort  N x 3 row-major, radius, volume  N: DEVICE arrays (inputs)
 *   vel, dvort N x 3 row-major: DEVICE arrays (outputs)
 *   uinf       host array of 3 (NULL: no freestream); stream: a hipStream_t
 *              (NULL: the default stream) on the current device
 * Asynchronous: returns once the kernels are queued on `stream`.  At most
 * PW_NOALLOC_MAX_N particles (PW_EINVAL beyond, with a message).  A
 * non-finite rate is not reported by the return code (no host
 * synchronisation); it shows in the outputs.  Every call on a device shares
 * that device's array, so calls that overlap on one device (on different
 * streams or host threads) race; calls in order on one stream do not.  The
 * first call on each device sets up (looks up the array's address there). */
#define PW_NOALLOC_MAX_N 16384
int pw_rates_noalloc(int kernel, double nu, const double *uinf, size_t n, const double *pos,
                     const double *vort, const double *radius, const double *volume, double *vel,
                     double *dvort, void *stream);
/* bytes of the per-device scratch array of pw_rates_noalloc */
size_t pw_noalloc_scratch_bytes(void);

/* ------------------------------------------------------------------ batch
 * Many independent simulations stepped together: one persistent launch over
 * a table of (job, target tile, source chunk) work items per RHS, and the RK4
 * stage combinations fused with the per-target reduction.  The step is
 * captured once as a HIP graph and replayed. */
typedef struct pw_batch pw_batch;

typedef struct pw_job_spec {
    size_t n;             /* particles                                     */
    const double *pos;    /* N x 3                                         */
    const double *vort;   /* N x 3                                         */
    const double *radius; /* N                                             */
    const double *volume; /* N                                             */
    double nu;            /* viscosity                                     */
    double dt;            /* RK4 time step                                 */
    double uinf[3];       /* freestream                                    */
    int precision;        /* PW_FP64 / PW_FP32C                            */
} pw_job_spec;

/* variant may be NULL; zero fields take the defaults: block 256, two targets
 * per thread for binary64 jobs (one for GAUSSIAN) and one for binary32 jobs,
 * default unroll, symmetric binary32 pass; a nonzero targets_per_thread
 * applies to both precisions */
int pw_batch_create(pw_batch **out, int device, int kernel, const pw_variant *variant);
int pw_batch_destroy(pw_batch *b);
const char *pw_batch_last_error(const pw_batch *b);
/* returns the job id in *id */
int pw_batch_add(pw_batch *b, const pw_job_spec *spec, int64_t *id);
int pw_batch_remove(pw_batch *b, int64_t id);
/* RK4-advance every job by nsteps (use_graph: replay a captured step) */
int pw_batch_step(pw_batch *b, size_t nsteps, int use_graph);
/* evaluate rates only, for all jobs */
int pw_batch_solve(pw_batch *b);
/* one job's particle positions/vorticities and last rates (N x 3 each; NULL skips) */
int pw_batch_get(pw_batch *b, int64_t id, double *pos, double *vort, double *vel, double *dvort);
/* per-job status: PW_OK or PW_ENONFINITE after the last step */
int pw_batch_status(pw_batch *b, int64_t id);
int pw_batch_last_timing(const pw_batch *b, double *total_ms, double *pair_ms, double *pairs);
/* source chunk length of the one-sided pass for a job with n particles: 16
 * chunks, or up to one block of sources per chunk below 256 blocks (a
 * function of n and block only; jobs on the symmetric pass use max(128,
 * blocks / 8) chunks, capped at the block count) */
size_t pw_chunk_for(size_t n, int block);

/* ---------------------------------------------------------------- serving
 * A continuous-batching request loop: jobs are queued with an arrival time,
 * admitted while their HBM footprint fits the budget, advanced together with
 * every other resident job, rebatched every pass as jobs join and finish, and
 * retired with their final state kept until fetched.
 *
 * A pass advances resident jobs by 1, 2 or 4 RK4 stages (never past the end
 * of a job's current step), or by a slice of one stage.  The reference cost
 * is the largest resident job's stage, held within [2e9, 1.6e10] pair
 * interactions (a binary32 pair priced at half).  A job whose stage costs
 * at most a quarter of it ("small") runs 4 stages in every pass; the others
 * ("large") run 2 if their stage costs at most half of it, else 1, and a
 * stage costing more than the reference runs in slices of about the
 * reference's work (its pair work items cut into consecutive parts, the last
 * part with the stage's reduction and update).  Large jobs take turns: a
 * pass takes them in id order from a rotating cursor while their work stays
 * within the reference (at least one).  So small jobs take whole steps per
 * pass, no pass holds much more than the reference besides them, and every
 * job advances.  Each stage is the same computation whichever passes it
 * falls in, so a job's result does not depend on the mix.
 * Jobs that arrive during a pass are uploaded beside it (a second stream)
 * and join the next one; the non-finite flags of all jobs come back in one
 * copy per pass. */
typedef struct pw_server pw_server;

typedef struct pw_server_stats {
    int64_t submitted, admitted, finished, failed, cancelled, queued, running;
    int64_t steps;             /* passes executed                              */
    double pairs;              /* ordered pair-interactions evaluated         */
    double busy_ms;            /* device time of the passes                   */
    double step_p50_ms, step_p99_ms; /* pass latency (admit, run, retire)     */
    double job_p50_ms, job_p99_ms;   /* arrival -> final state ready          */
    size_t hbm_in_use, hbm_budget;   /* bytes                                 */
    double job_step_p50_ms, job_step_p99_ms; /* per finished job: admission ->
                                       retirement / RK4 steps, the wall time
                                       per step that job saw                  */
} pw_server_stats;

int pw_server_create(pw_server **out, int device, int kernel, size_t hbm_budget_bytes,
                     const pw_variant *variant);
int pw_server_destroy(pw_server *s);
const char *pw_server_last_error(const pw_server *s);
/* the server's clock: milliseconds since pw_server_create (steady clock) */
double pw_server_now_ms(const pw_server *s);
/* queue a job that runs nsteps RK4 steps; arrival_ms on the server clock
 * (negative: now).  Job latency is measured from arrival to retirement. */
int pw_server_submit(pw_server *s, const pw_job_spec *spec, size_t nsteps, double arrival_ms,
                     int64_t *id);
int pw_server_cancel(pw_server *s, int64_t id);
/* one pass of the request loop: queue a pass (the small jobs and a turn of
 * the large ones, 1, 2 or 4 RK4 stages or a stage slice each), admit queued
 * jobs that fit the budget while it runs, then retire finished and failed
 * jobs.  *running
 * receives the number of resident jobs after the pass. */
int pw_server_step(pw_server *s, int64_t *running);
/* 0 queued, 1 running, 2 done, 3 failed (non-finite), 4 cancelled, -1 unknown */
int pw_server_job_state(pw_server *s, int64_t id, int64_t *steps_done);
/* final positions/vorticities of a finished job (N x 3 each); frees it */
int pw_server_fetch(pw_server *s, int64_t id, double *pos, double *vort);
int pw_server_get_stats(pw_server *s, pw_server_stats *st);
/* device bytes one job of n particles occupies while resident, for the
 * default variant of a context of n particles (a binary32 job includes the
 * symmetric pass's slot rows, 80 bytes per particle per strip with
 * max(128, blocks / 8) strips: 2.8 GB at n = 262,144, 21.9 GB at 1,048,576) */
size_t pw_job_footprint(size_t n, int precision);

/* -------------------------------------------------------------- autotuner
 * Times every compiled variant (block x targets per thread x unroll, and
 * both pair schemes) on the given particles at the chunk policy, then the
 * fastest variant of each scheme at 1/2, 3/4, 3/2 and 2x the policy's source
 * chunk count (reps solves each, HIP events), and returns them in `all` (up
 * to max_all entries; variant.chunks 0 = the policy) and the fastest in
 * *best.  One-sided variants give bitwise identical results at one chunk
 * count; symmetric ones group their sums by block and strip and agree to
 * binary32 rounding. */
typedef struct pw_tune_entry {
    pw_variant variant;
    double pair_ms;      /* mean device time per solve: pair pass and
                            reduction (HIP events)                  */
    double pairs_per_s;  /* ordered pair-interactions per second   */
} pw_tune_entry;

int pw_autotune(int device, int kernel, int precision, size_t n, const double *pos,
                const double *vort, const double *radius, const double *volume, int reps,
                pw_tune_entry *all, size_t max_all, size_t *n_all, pw_tune_entry *best);
/* A tuning cache file: one line per (kernel, precision, log2 N bucket):
 * "kernel precision log2N block targets_per_thread unroll scheme ms chunks"
 * (a line without the chunks field reads as chunks 0).  pw_create consults the
 * file named by the PW_TUNE_CACHE environment variable when block,
 * targets_per_thread, unroll and scheme are all 0; its chunk count applies
 * when the config's chunk and chunks are 0. */
int pw_tune_save(const char *path, int kernel, int precision, size_t n, const pw_tune_entry *e);
int pw_tune_lookup(const char *path, int kernel, int precision, size_t n, pw_variant *out);

#ifdef __cplusplus
}
#endif
#endif
