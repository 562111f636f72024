/*
 * mq_engine.h -- C ABI of the MI355X batched ion-trajectory integrator.
 *
 * Drop-in boundary for chsim's batch/scan API (simulation/chsim/src/channel.rs,
 * simulation/chsim/src/scan.rs). Plain pointers and sizes only; every call is
 * synchronous on the engine's own HIP stream; an engine handle is not
 * thread-safe (the reference Engine is !Sync, channel.rs:309-318).
 *
 * Return convention: 0 = ok, < 0 = error; mq_last_error() gives the message of
 * the last failing call on the calling thread. Where the reference panics
 * (channel.rs:328,331,444,446) these entry points return an error instead.
 */
#ifndef MQ_ENGINE_H
#define MQ_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Ion status codes: ChIon p.w in channel.metal:32. */
enum {
    MQ_ALIVE = 0,
    MQ_CRASHED = 1,
    MQ_RETURNED = 2,
    MQ_TRANSMITTED = 3,
    MQ_LEFT = 4,
    MQ_TIMEOUT = 5
};

/* Execution path: Path { Tiled, Direct } in channel.rs:304.
 * BATCHED = FP32 fields/interpolation/push with FP64 position, time and phase,
 *           persistent wavefront workers with compaction (tiled.rs:159-218).
 * PLAIN   = FP64 everything, one thread per ion, relaunched until every ion
 *           has finished (Engine::run_direct, channel.rs:478-493). */
enum { MQ_PATH_BATCHED = 0, MQ_PATH_PLAIN = 1 };

/* Injection model. REFERENCE follows Physics.legacy_injection exactly as
 * Engine::inject (channel.rs:342-374): legacy 10-degree cone on axis, or a
 * thermal disc with flux-weighted axial speed. GAUSSIAN: Gaussian spot of
 * sigma inj_sigma_mm about the axis, Maxwell-Boltzmann velocity at temp_k. */
enum { MQ_INJ_REFERENCE = 0, MQ_INJ_GAUSSIAN = 1 };

/* Random streams. PHILOX: Philox4x32-10 keyed by the job seed, counter
 * (ion, step, ...): results depend on (seed, ion, step) only. COMPAT: the
 * reference's own streams: one xorshift64 stream per job drives injection
 * exactly as Engine::inject (channel.rs:342-374, on the host, with its
 * fixed-point positions and f32 velocities), and a per-ion xoshiro128**
 * stream seeded from it drives collisions in the draw order of
 * channel.metal:38-53, 93-113. */
enum { MQ_RNG_PHILOX = 0, MQ_RNG_COMPAT = 1 };

/* Species (channel.rs:236): mass [Da], charge number, N2 CCS [A^2]. */
typedef struct { double mass, z, ccs_n2; } mq_species;

/* RF + AC drive of one job (channel.rs:247). Volts and Hz. */
typedef struct { double v_rf, f_rf, v_ac, f_ac; } mq_rfac;

/* One scan point (channel.rs:252). Extensions for setting sweeps: dc_scale
 * multiplies the batch's combined DC field for this job only (1.0 reproduces
 * the reference, which shares DC per batch); pressure_pa > 0 overrides the
 * batch pressure for this job; per-job v_rf/f_rf are allowed (the reference
 * asserts a shared RF, channel.rs:446). */
typedef struct {
    mq_species sp;
    mq_rfac drive;
    uint64_t n;
    uint64_t seed;
    double dc_scale;
    double pressure_pa;
} mq_job;

/* Physics switches (channel.rs:255-259) plus the injection model and an
 * explicit step cap. max_steps = 0 means round(t_max / dt). */
typedef struct {
    int32_t legacy_force;       /* 1: half kick 0.5*q/m (their 1/2 a dt) */
    int32_t legacy_collisions;  /* 1: stationary-gas countdown collisions  */
    int32_t legacy_fields;      /* 1: field of the lower-corner node, no interpolation */
    int32_t legacy_injection;   /* 1: legacy cone injection (MQ_INJ_REFERENCE only) */
    int32_t legacy_criteria;    /* 1: legacy return/exit planes and rod-only crash window */
    int32_t injection;          /* MQ_INJ_* */
    double pressure_pa, temp_k, dt, t_max, r_inject_mm;
    double inj_sigma_mm;
    uint64_t max_steps;
    int32_t rng;                /* MQ_RNG_* */
    int32_t pad_;
} mq_physics;

/* Per-job outcome counts (channel.rs:297). mean_time in s. */
typedef struct {
    uint64_t n, crashed, returned, transmitted, left, timeout;
    double mean_coll, mean_time;
} mq_outcome;

/* Basis E-field grids, E per volt, node-major index i + nx*(j + ny*k), node
 * (i,j,k) at (i*h, j*h, k*h) [m]. Fields (channel.rs:122-129): RF, AC, four DC
 * bases (q1, plate, downstream, box) of which any may be NULL (zero), and the
 * conductor mask (RF .w in the reference). Plane and window positions are the
 * values channel.rs:397-410 derives from its geometry. */
typedef struct {
    int32_t nx, ny, nz;
    double h;
    double axis_x, axis_y, z_inject;
    double z_return, z_exit;                 /* correct criteria */
    double z_return_legacy, z_exit_legacy;   /* legacy criteria */
    double crash_z0_legacy, crash_z1_legacy; /* legacy crash window (rods) */
    const double* rf;      /* 3*N */
    const double* ac;      /* 3*N */
    const double* dc[4];   /* 3*N each, or NULL */
    const uint8_t* mask;   /* N, nonzero = conductor node */
} mq_fields;

/* Final ion states, structure of arrays, caller-owned, n entries each. */
typedef struct {
    double *x, *y, *z;
    double *vx, *vy, *vz;
    uint32_t *steps;   /* steps executed (time = steps * dt) */
    uint32_t *ncoll;   /* collisions */
    uint8_t *status;   /* MQ_* status */
    /* Optional (NULL = not returned): */
    double *phi_ac0;   /* AC phase at injection [cycles, 0..1); the phase at the
                          end is frac(phi_ac0 + f_ac * dt * steps) */
    uint32_t *cd;      /* legacy collision countdown (steps to the next collision) */
} mq_ion_soa;

/* Statistics of the last run on an engine. */
typedef struct {
    uint64_t ion_steps;      /* sum over ions of steps executed */
    uint64_t n_ions;
    uint32_t launches;       /* worker / direct kernel launches */
    double kernel_ms;        /* device time of the worker kernels (HIP events) */
    double wall_ms;          /* host wall time of the call */
    uint32_t quad_launches;  /* of the launches, those of the four-lane tail worker (mq_tune.tail_quads) */
    uint32_t plain_recs;     /* plain path: 1 if its launches read the FP64 cell records (mq_tune.plain_recs) */
} mq_run_stats;

/* Tuning of the batched path (TileCfg, tiled.rs:12-29), and the plain path's
 * field access. mq_set_tune: zero in block, waves_per_cu, refill_period,
 * chunk, chunk_cap, sync_every and coll_batch keeps the current value; resort,
 * graph, tail_lanes, tail_quads and plain_recs are always taken. Results per
 * ion do not depend on any of these. */
typedef struct {
    int32_t block;           /* threads per workgroup (64..256, multiple of 64) */
    int32_t waves_per_cu;    /* persistent waves launched per CU */
    int32_t refill_period;   /* steps between idle-lane refills (compaction period) */
    int32_t chunk;           /* max steps per ion per launch (the re-sort interval) */
    int32_t chunk_cap;       /* without re-sort the chunk doubles per launch up to this */
    int32_t resort;          /* 1: sort the live list by grid cell before every launch,
                                one queue per XCD over the sorted list; 2: as 1, and
                                each 256-thread workgroup takes 256 consecutive entries
                                of the sorted list for a whole chunk with the records of
                                the 6x6x6 cells around their bin staged in LDS (the
                                reference's tiled rounds, channel_tiled.metal:237-334;
                                Philox runs only, COMPAT runs sort as with 1) */
    int32_t sync_every;      /* launches enqueued per host check of the live count */
    int32_t graph;           /* 1: capture those launches as one hipGraph and replay it
                                (sync_every must be even) */
    int32_t coll_batch;      /* lanes of a wavefront whose collision is pending before the
                                collision code runs for all of them (1: at once) */
    int32_t tail_lanes;      /* once the live ions fit the launch's lanes: ions per wavefront
                                at most (1..64; 0: spread them evenly over every wavefront) */
    int32_t tail_quads;      /* 1: once no more ions are live than 16 per wavefront, step
                                each on four lanes (the step's field components, drive sines
                                and candidate draws split over the lanes; Philox runs only);
                                0: one lane per ion throughout */
    int32_t plain_recs;      /* plain path: 1 (default): per-cell FP64 records of the 8 corner
                                values of the three bases (576 B per cell, built once per
                                combined DC when device memory allows), kept per lane while
                                its ion stays in the cell; 0: gather the node tables at
                                every step. The same arithmetic either way */
} mq_tune;

typedef struct mq_engine mq_engine;

/* Message of the last failing call on this thread ("" if none). */
const char* mq_last_error(void);

/* Replaces Engine::new (channel.rs:327-340): selects the GPU, uploads the RF,
 * AC and DC basis grids and the conductor mask once. */
int mq_engine_create(const mq_fields* fields, mq_engine** out);

/* Drops the engine and all device memory it owns. NULL is a no-op. */
void mq_engine_destroy(mq_engine* eng);

/* Replaces Engine::run_batch (channel.rs:437-440): runs every job of the batch
 * to completion and writes one outcome per job into out[njobs]. dc[4] are the
 * voltages of the four DC bases, shared by the batch. */
int mq_run_batch(mq_engine* eng, const mq_job* jobs, size_t njobs, const double dc[4],
                 const mq_physics* ph, int32_t path, mq_outcome* out);

/* Replaces Engine::run_batch_ions (channel.rs:458-475): as mq_run_batch but
 * returns the final state of every ion (concatenated in job order, sum of
 * jobs[k].n entries) and the job index of each ion (out_job may be NULL). */
int mq_run_batch_ions(mq_engine* eng, const mq_job* jobs, size_t njobs, const double dc[4],
                      const mq_physics* ph, int32_t path, const mq_ion_soa* out, uint32_t* out_job);

/* Largest batch of mq_run_batch_ions_noalloc. */
#define MQ_NOALLOC_MAX_IONS 65536
#define MQ_NOALLOC_MAX_JOBS 4096

/* mq_run_batch_ions for callers that cannot allocate device memory (code that
 * runs under HIP graph capture elsewhere, tight serving loops, harnesses that
 * forbid hipMalloc). No counterpart in the reference. Same arguments, results
 * and errors as mq_run_batch_ions, bit for bit, with these limits:
 *  - at most MQ_NOALLOC_MAX_IONS ions and MQ_NOALLOC_MAX_JOBS jobs (larger
 *    batches are rejected with an error);
 *  - ph->rng must be MQ_RNG_PHILOX (the reference's streams are injected on
 *    the host into per-call buffers);
 *  - the engine's field tables for this path must exist: one earlier
 *    mq_run_batch / mq_run_batch_ions call on the same path builds them (a new
 *    DC configuration is then recomputed into them in place).
 * The per-call device buffers (ion state, live lists, re-sort bins, counters,
 * job table, accumulators, about 12 MB) are a static array of the library that
 * every GPU holds once, for the life of the process; the call calls no
 * hipMalloc or hipFree. All engines on one GPU share that array: calls that
 * overlap on one GPU, from several threads or engines, race. The batched path
 * runs without hipGraph capture here. */
int mq_run_batch_ions_noalloc(mq_engine* eng, const mq_job* jobs, size_t njobs, const double dc[4],
                              const mq_physics* ph, int32_t path, const mq_ion_soa* out, uint32_t* out_job);

/* Engine::inject (channel.rs:342-374) on the host with the reference's
 * xorshift64 stream (the injection of MQ_RNG_COMPAT): the initial state of
 * the n ions of one job in s (positions after the reference's fixed-point
 * truncation, velocities rounded to f32, phi_ac0 the AC phase in cycles
 * rounded to f32, cd the legacy countdown) and the four xoshiro128** seed
 * words of each ion in rng[4n]. Host only, no GPU. */
int mq_inject_reference(const mq_fields* f, const mq_job* job, const mq_physics* ph, const mq_ion_soa* s,
                        uint32_t* rng);

/* Replaces summarize (channel.rs:496-506): per-job counts and means from final
 * states. dt converts steps to time. */
int mq_summarize(const uint8_t* status, const uint32_t* ncoll, const uint32_t* steps,
                 const uint32_t* job, size_t n, size_t njobs, double dt, mq_outcome* out);

/* Statistics of the last run_batch / run_batch_ions call on this engine. */
int mq_last_stats(const mq_engine* eng, mq_run_stats* out);

/* Achievable gather bandwidth of the engine's FP32 cell table (no reference
 * counterpart: the measurement that a roofline of the batched path needs).
 * Random whole-record reads (304 B, the worker's access unit) by persistent
 * wavefronts at the worker's occupancy, over the first window_cells records
 * of the table built by the last batched run (0 = the whole table); iters
 * records per lane. *bytes_per_s receives the bandwidth. */
int mq_gather_probe(mq_engine* eng, uint64_t window_cells, uint32_t iters, double* bytes_per_s);

/* The batched path's FP32 field at n positions [m] (no reference counterpart:
 * it exposes the worker's own field evaluation -- cell location, FP32 cell
 * record of the combined DC of dc[4], polynomial, drive amplitudes -- so it can
 * be checked against the FP64 trilinear interpolation of the node tables).
 * E = a_rf E_rf + s_dc E_dc + a_ac E_ac, trilinear, or the lower-corner node
 * when nearest = 1; crash[i] = conductor bit of the node nearest position i. */
int mq_field_probe(mq_engine* eng, const double dc[4], int32_t nearest, double a_rf, double s_dc, double a_ac,
                   size_t n, const double* x, const double* y, const double* z, float* ex, float* ey, float* ez,
                   uint8_t* crash);

/* Tuning of the batched path; replaces TileCfg::from_env (tiled.rs:24-28). */
int mq_set_tune(mq_engine* eng, const mq_tune* t);
int mq_get_tune(const mq_engine* eng, mq_tune* out);

/* Autotuner of the batched path (the reference tunes by environment
 * variables, tiled.rs:24-28): a coordinate search over re-sort on/off, chunk
 * (re-sort interval), waves per CU, block size, refill (compaction) period,
 * collision batch and hipGraph replay, timed on an evenly spaced subset of the jobs of
 * about sample_ions ions (0: 262144), sets the fastest on the engine and
 * appends it to the TSV file cache_path (NULL: no file) keyed by grid and
 * physics switches; a later call with the same key reads it back instead of
 * measuring unless force is 1. *out and *ion_steps_per_s (0 when read from
 * the file) are optional. Per-ion results do not depend on the tuning. */
int mq_autotune(mq_engine* eng, const mq_job* jobs, size_t njobs, const double dc[4], const mq_physics* ph,
                uint64_t sample_ions, const char* cache_path, int32_t force, mq_tune* out,
                double* ion_steps_per_s);

/* ------------------------------------------------- differentiable form */

/* Objective of the differentiable form: every ion of one job is integrated
 * for exactly nsteps FP64 steps (no termination), then
 *   J = mean_i [ sigmoid((z_i - z_exit) / scale)
 *               - w_leak * sigmoid((r_i - r_leak) / scale) ],
 * r_i = distance from the injection axis. Checkpoints: ckpt_every = 1 keeps
 * every step's position (24 B per ion-step, 48 B with gas) so that mq_backward
 * needs no recompute; 0 does so when that fits a quarter of the free device
 * memory and otherwise checkpoints every ceil(sqrt(nsteps)) steps, recomputing
 * each segment in mq_backward; K > 1 checkpoints every K steps. The results are
 * the same bit for bit.
 * With no gas (pressure 0 or ccs_n2 0) the steps are collision-free. With gas
 * the thermal collisions of the plain path apply with common random numbers:
 * each ion's Philox draws depend on (seed, ion, step) only, the candidate and
 * acceptance decisions are those of the run, and the gradient is the pathwise
 * derivative through each collision's velocity map with the decisions held
 * fixed (it omits the jumps where a decision changes with theta). Legacy
 * collisions are rejected. */
typedef struct {
    uint32_t nsteps, ckpt_every;
    double z_exit, scale, r_leak, w_leak;
} mq_objective;

/* Forward op: injects job->n ions (as mq_run_batch would) and returns the
 * objective in *J, keeping the checkpoints on the engine for mq_backward.
 * Parameters theta = (drive.v_rf, drive.f_rf, drive.v_ac, drive.f_ac,
 * dc_scale). No counterpart in the reference, which has no gradients: its
 * Optimize search is a grid of sweeps (src-tauri/src/research.rs:133-163). */
int mq_forward(mq_engine* eng, const mq_job* job, const double dc[4], const mq_physics* ph,
               const mq_objective* obj, double* J);

/* Backward op of the last mq_forward: grad[5] = dJ_up * dJ/dtheta, and, if
 * grad_state is not NULL, dJ_up * dJ/d(initial x, y, z, vx, vy, vz) of every
 * ion as 6 arrays of n (x for all ions, then y, ...). */
int mq_backward(mq_engine* eng, double dJ_up, double grad[5], double* grad_state);

/* Per-ion objective J_i of the last mq_forward (n values, J = mean J_i) and
 * per-ion gradient contributions of the last mq_backward (5 arrays of n, in
 * the order of grad; grad = sum over ions). For conditioning diagnostics. */
int mq_forward_per_ion(mq_engine* eng, double* J_ion);
int mq_backward_per_ion(mq_engine* eng, double* grad_ion);

/* Final position of each ion after the last mq_forward (n values each, metres).
 * No counterpart in the reference: the differentiable form is new. Lets a
 * caller tell whether a finite-difference step moved an ion by more than the
 * objective's length scale, i.e. left the range where differences are linear. */
int mq_forward_final(mq_engine* eng, double* x, double* y, double* z);

/* ------------------------------------------------------------ scan API */

/* One point of an adaptive scan: AC frequency [kHz], transmitted and crashed
 * fractions, ions simulated (scan.rs:51, the (f, T, C, n) tuple). */
typedef struct { double f_khz, transmitted, crashed; uint64_t n; } mq_scan_point;

/* Replaces scan::cached_batch (scan.rs:31-47): outcomes of jobs whose key is
 * already in the TSV file `cache_path` are read from it, the rest run as one
 * batch and are appended. geom_key is the caller's geometry key (the Rust
 * Debug string of Geom in the reference). *n_computed (optional) receives the
 * number of jobs run. */
int mq_cached_batch(mq_engine* eng, const char* geom_key, const mq_job* jobs, size_t njobs, const double dc[4],
                    const mq_physics* ph, int32_t path, const char* cache_path, mq_outcome* out,
                    size_t* n_computed);

/* Replaces scan::adaptive_scan (scan.rs:52-78): coarse AC-frequency scan
 * f_lo..f_hi [kHz] step f_step with n_lo ions (seeds seed0 + i), then a fine
 * scan at f_step/4 with n_hi ions (seeds seed0 + 100000 + i) around points with
 * T >= 0.35 T_max and T - base > 4 sigma; both through mq_cached_batch.
 * Writes up to max_out points sorted by f; *n_out = number of points. Returns
 * 0, < 0 on error, or the required size if max_out is too small. */
int mq_adaptive_scan(mq_engine* eng, const char* geom_key, const mq_species* sp, const mq_rfac* rf,
                     const double dc[4], const mq_physics* ph, double f_lo, double f_hi, double f_step,
                     uint64_t n_lo, uint64_t n_hi, double base, uint64_t seed0, int32_t path,
                     const char* cache_path, mq_scan_point* out, size_t max_out, size_t* n_out);

/* The refinement rule of adaptive_scan (scan.rs:59-67) on given coarse
 * points; same return convention as mq_adaptive_scan. */
int mq_adaptive_fine_points(const double* f_khz, const double* t, size_t nf, uint64_t n_lo, double base,
                            double f_step, double* out, size_t max_out, size_t* n_out);

/* The cache key of one job (scan.rs:9-11) into buf; returns 0, < 0 on error,
 * or the required length if len is too small. path is the execution path the
 * job runs on (the reference's key() reads it from the Engine). For Philox
 * streams on the batched path the key is the reference's; ph->rng other than
 * MQ_RNG_PHILOX appends ", rng: <n>" and a path other than MQ_PATH_BATCHED
 * appends ", path: <n>" inside the Physics braces, so a cached result is only
 * returned to a request for the same streams and the same path. */
int mq_scan_key(const char* geom_key, const mq_job* job, const double dc[4], const mq_physics* ph, int32_t path,
                char* buf, size_t len);

/* An f64 in Rust Debug ({:?}, display = 0) or Display ({}, display = 1)
 * notation, as used in the cache file; same return convention as mq_scan_key. */
int mq_format_f64(double x, int32_t display, char* buf, size_t len);

/* ---------------------------------------------------------- record compat */

/* The reference's 64-byte per-ion record, ChIon (channel.rs:292-294,
 * channel.metal:31-36): p = fixed-point position at MQ_POS_SCALE per metre
 * (62.5 pm) with the status in p[3]; v = velocity [m/s] with the ion's own
 * time [s] in v[3]; aux = {AC phase [cycles], collisions, legacy countdown,
 * RF phase [cycles]}; rng = the random-stream position of this build,
 * {key lo, key hi, ion index in its job, steps} (the reference holds its
 * xoshiro128** state there). Little-endian, no padding. */
#define MQ_POS_SCALE 1.6e10
typedef struct {
    int32_t p[4];
    float v[4];
    float aux[4];
    uint32_t rng[4];
} mq_chion;

/* Writes n ChIon records from final states (ions in job order, as returned by
 * mq_run_batch_ions). s->phi_ac0 and s->cd are read when non-NULL (else the
 * AC phase and countdown are written as 0). Positions are rounded to the
 * nearest quantum, as the reference's drift does (channel.metal:88-90). */
int mq_export_chion(const mq_ion_soa* s, const uint32_t* job, size_t n, const mq_job* jobs, size_t njobs,
                    const mq_physics* ph, mq_chion* out);

/* summarize (channel.rs:496-506) over ChIon records: status from p[3],
 * collisions from aux[1], time from v[3], summed in f64. */
int mq_summarize_chion(const mq_chion* ions, const uint32_t* job, size_t n, size_t njobs, mq_outcome* out);

/* FNV-1a 64-bit hash of a byte range, as check.rs:100-104 hashes the raw
 * ChIon bytes of a run (offset basis 14695981039346656037, prime
 * 1099511628211). */
uint64_t mq_fnv1a64(const void* data, size_t nbytes);

#ifdef __cplusplus
}
#endif

#endif /* MQ_ENGINE_H */
