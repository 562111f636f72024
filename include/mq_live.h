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

/* A new AC frequency [Hz] from the current step on, phase continuous, and the
 * tag of the ions spawned from now on (the sweep point of Sim::block,
 * engine.rs:518-533). */
int mq_live_set_ac(mq_live* l, double f_ac, int32_t tag);

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
