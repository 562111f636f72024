/*
 * mq_live.h -- a live ion population on the engine: the block loop of the
 * reference's interactive model (simulation/mqviz/src/engine.rs:448-538,
 * Sim::block), with its transport on the HIP batched path.
 *
 * The population shares one running clock: an ion born at global step s sees
 * the RF phase f_rf dt s and the AC phase of the clock, whose frequency may
 * change between blocks with a continuous phase (Clock::fill, world.rs:442-465).
 * Each block advances every live ion block_steps steps (fewer if it finishes),
 * then, as Sim::block does: tallies the finished ions by tag and species;
 * schedules each returned ion's re-entry after an exponential delay when
 * recirculation is on; spawns the re-entries that are due; and spawns a Poisson
 * number of fresh ions with species drawn by weight. The block's host draws use
 * the reference's xorshift64 stream and its Poisson sampler (engine.rs:318-328);
 * finished ions are processed in the order of their birth. Transport uses each
 * species' Philox stream (key = seed + species index, stream number = the
 * ion's birth order), not the reference's per-ion seeds.
 *
 * Conventions as mq_engine.h: 0 = success, < 0 = error (mq_last_error()).
 */
#ifndef MQ_LIVE_H
#define MQ_LIVE_H

#include <stddef.h>
#include <stdint.h>

#include "mq_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t block_steps; /* steps per block (a multiple of 4; the reference uses 128) */
    uint32_t capacity;    /* live ions at most (MAXION); spawns beyond it are dropped and counted */
    double inj_rate;      /* fresh ions per simulated second: Poisson(inj_rate block_steps dt) per block */
    int32_t recirculate;  /* 1: a returned ion re-enters after an exponential delay */
    int32_t max_tags;     /* outcome tallies are kept for tags 0 .. max_tags - 1 */
    double residence_s;   /* mean re-entry delay [s of simulated time] */
    uint64_t seed;        /* the host stream (odd-ised) and the species' Philox keys (seed + k) */
} mq_live_opts;

/* Counters of a population since its creation. */
typedef struct {
    uint64_t step;      /* global steps run */
    uint64_t alive;     /* live ions now */
    uint64_t injected;  /* fresh spawns (Poisson and mq_live_inject) */
    uint64_t reentered; /* recirculated spawns */
    uint64_t dropped;   /* spawns refused at capacity */
    uint64_t finished;  /* ions that ended (any class) */
    uint64_t ion_steps; /* steps executed by all ions */
} mq_live_state;

typedef struct mq_live mq_live;

/* A population of nsp species (weights: relative spawn probabilities, NULL =
 * equal) with the drive rf (v_rf, f_rf, v_ac, f_ac) and DC voltages dc on the
 * engine's grid, on the batched path with the engine's tuning. Physics as
 * mq_run_batch (Philox streams, thermal or legacy collisions); every ion's
 * step cap is its own age. */
int mq_live_create(mq_engine* eng, const mq_species* sp, const double* weights, size_t nsp, const mq_rfac* rf,
                   const double dc[4], const mq_physics* ph, const mq_live_opts* o, mq_live** out);
void mq_live_destroy(mq_live* l);

/* Spawns n fresh ions of one species at the current step (outside the block's
 * Poisson draw, which it does not touch). */
int mq_live_inject(mq_live* l, uint32_t species, uint32_t n);

/* Runs nblocks blocks; *out (optional) receives the counters after them. */
int mq_live_run(mq_live* l, uint32_t nblocks, mq_live_state* out);

/* A new AC frequency [Hz] of the live tone from the current step on, phase
 * continuous (the Params update of Sim::apply_params, engine.rs:399), and the
 * tag of the ions spawned from now on. While a sweep runs the live tone stays
 * with the sweep and takes the new frequency when the sweep stops. */
int mq_live_set_ac(mq_live* l, double f_ac, int32_t tag);

/* A frequency sweep of the population (Sim::start_sweep, stop_sweep and the
 * sweep bookkeeping of Sim::block, engine.rs:420-432, 512-535). The points are
 * f_lo_khz + i step_khz for i = 0 .. round((f_hi_khz - f_lo_khz) / step_khz),
 * 101 at most (one when f_hi_khz < f_lo_khz). Each point has its own AC tone,
 * a clock of its frequency whose phase runs on from the previous sweep's tone
 * of the same index, or starts at 0 where that sweep had fewer points (the
 * clock's resize, engine.rs:426); every ion spawned (fresh, injected or
 * re-entered) while the sweep runs and does not drain carries the current
 * point's tone for its whole life, and its outcome is counted for that point.
 * The tone of untagged ions (the live tone) follows the current point while
 * the sweep runs. Each fresh ion of the block's Poisson draw (mq_live_inject's
 * do not count) advances the point after ions_per_point of them; after
 * `sweeps` passes over the points the sweep drains: fresh injection stops
 * until the last tagged ion has ended, then the sweep stops, the live tone
 * returns to the AC frequency last set (mq_live_create or mq_live_set_ac) and
 * injection resumes. The reference replaces a running sweep on a new start;
 * here mq_live_sweep_start fails while a sweep runs or drains. */
typedef struct {
    double f_lo_khz, f_hi_khz, step_khz;
    uint32_t ions_per_point; /* fresh ions per point (>= 1) */
    uint32_t sweeps;         /* passes over the points (>= 1) */
} mq_live_sweep_cfg;

/* The state of the last sweep (sweep_view, engine.rs:434-446). */
typedef struct {
    uint32_t npoints;     /* 0: no sweep was started */
    uint32_t point;       /* current point (the last one once the sweep has stopped) */
    uint32_t sweeps_done; /* completed passes */
    int32_t running;      /* 1 while the sweep runs or drains */
    int32_t draining;     /* 1 while fresh injection is off and the tagged ions finish */
    int32_t pad_;
    double progress;      /* (sweeps_done npoints + point) / (sweeps npoints), at most 1 */
} mq_live_sweep_state;

#define MQ_LIVE_MAX_POINTS 101

int mq_live_sweep_start(mq_live* l, const mq_live_sweep_cfg* c);

/* Stops fresh injection; the sweep ends once its tagged ions have ended
 * (stop_sweep, engine.rs:431-432). As there, a sweep that has already ended is
 * marked running and draining again, so the next block draws no fresh ions and
 * ends it. No-op before the first sweep. */
int mq_live_sweep_stop(mq_live* l);

/* The last sweep's state, its frequencies [kHz] (npoints) and its counts
 * acc[point][species][4] of ended tagged ions: transmitted, crashed,
 * returned, other (left the grid or timed out). Any output pointer may be
 * NULL; f_khz and acc hold MQ_LIVE_MAX_POINTS points at most. */
int mq_live_sweep(const mq_live* l, mq_live_sweep_state* st, double* f_khz, uint64_t* acc);

/* Outcomes of the ions of one tag that have ended, one mq_outcome per species
 * (n, classes, mean collisions, mean time). */
int mq_live_tally(const mq_live* l, int32_t tag, mq_outcome* per_species);

/* The current contents of every slot (capacity entries): final or current
 * state (s), random-stream number, species, birth step and status 255 for a
 * slot never used. Any pointer may be NULL. */
int mq_live_slots(mq_live* l, const mq_ion_soa* s, uint32_t* sid, uint32_t* species, uint64_t* birth);

#ifdef __cplusplus
}
#endif

#endif /* MQ_LIVE_H */
