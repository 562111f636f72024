/*
 * mq_tandem.h -- tandem multi-domain batch on one MI355X.
 *
 * Ions share one clock while they move between domains: a splitter, nine
 * channel ports and a combiner. The transport and bookkeeping follow
 * trapsim's Model::advance, enter and leave
 * (simulation/trapsim/src/lib.rs:176-385):
 *   - the splitter and the combiner sample one "cube" grid with two DC
 *     configurations and no excitation (cube_field, lib.rs:70-90);
 *   - each port samples one shared channel grid with its own excitation tone;
 *     a port whose tone is 0 is not driven and gets a blocking DC increment
 *     (lib.rs:196-203);
 *   - in the cube an ion on a conductor node is lost (classify, lib.rs:208-212);
 *   - above the seam, an ion inside the aperture of the nearest mouth enters
 *     that port (lib.rs:213-225, enter 326-356); outside it, the ion keeps
 *     moving and is lost only once it passes the top of the cube (lib.rs:226-229);
 *   - past the cube's y = (ny - 1) h face an ion outside the entrance aperture
 *     is lost (lib.rs:231-236); one through it that gets l_ext beyond the face
 *     is back (from the splitter) or detected (from the combiner)
 *     (lib.rs:237-243); otherwise an ion past any other face is lost;
 *   - an ion that leaves a port through its return plane re-enters the
 *     splitter with its state intact, and one that leaves through the exit
 *     plane passes to the combiner (leave, lib.rs:357-377).
 * Differences from trapsim: the cube is a grid the caller supplies (trapsim
 * builds its 486-post cube in geometry.rs / fields.rs; its posts and board
 * are conductor nodes of the caller's mask here); both regions share one
 * bath with chsim's thermal collisions and Philox streams (mq_engine.h).
 *
 * Conventions as mq_engine.h: 0 = success, < 0 = error (mq_last_error()).
 */
#ifndef MQ_TANDEM_H
#define MQ_TANDEM_H

#include <stddef.h>
#include <stdint.h>

#include "mq_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MQ_TANDEM_PORTS 9
#define MQ_TANDEM_MAX_JOBS 8

/* Region of an ion (Region, lib.rs:15-22): splitter, port i (1 + i), combiner. */
enum { MQ_REG_SPLIT = 0, MQ_REG_PORT0 = 1, MQ_REG_COMBINE = 10 };

/* Final status of a tandem ion: 1 lost (electrode, wall, board or box),
 * 2 back out of the splitter's entrance, 3 detected past the combiner,
 * 5 still travelling at the step cap (region tells where). */
enum { MQ_T_ALIVE = 0, MQ_T_LOST = 1, MQ_T_BACK = 2, MQ_T_DETECTED = 3, MQ_T_TIMEOUT = 5 };

/* "No previous port" in the per-ion last-port record. */
#define MQ_TANDEM_NO_PORT 255

/* Coupling geometry, metres, in the cube grid's frame. The port's own planes
 * (z_return, z_exit) and axis come from the port grid's mq_fields. */
typedef struct {
    double seam_z;                   /* z of the port board (seam, lib.rs:126) */
    double mouth[MQ_TANDEM_PORTS][2];/* mouth centres (x, y) (port_axis_xy, lib.rs:107-110) */
    double ap_r;                     /* aperture radius of the mouths and of the entrance (geom.ap_r) */
    double entrance[2];              /* (x, z) of the entrance aperture's centre on the cube's
                                        y = (ny - 1) h face (port_axis_xy(entrance), lib.rs:232) */
    double l_ext;                    /* entrance tube length (geom.l_ext, lib.rs:237) */
} mq_tandem_geom;

/* Drive of one tandem run (Model::advance, lib.rs:176-183): one RF on every
 * grid with phase 0 at step 0, one excitation amplitude, a tone per port
 * (0 = not driven), DC voltages on the cube's four DC bases for the splitter
 * and the combiner, on the port grid's bases for a driven port, and the
 * increment added for a port that is not driven. */
typedef struct {
    double v_rf, f_rf;
    double v_ac;
    double tones[MQ_TANDEM_PORTS];
    double dc_split[4], dc_combine[4];
    double dc_port[4], bias_idle[4];
} mq_tandem_drive;

/* Per-port counters (Counts, lib.rs:28-39); by_job[k] counts forwarded ions
 * of job k (by_charge). retry_from[j]: entries into this port from the
 * splitter by ions whose previous port was j. */
typedef struct {
    uint64_t entered, returned, forwarded, revisits, from_other, reverse_entries, lost;
    uint64_t retry_from[MQ_TANDEM_PORTS];
    uint64_t by_job[MQ_TANDEM_MAX_JOBS];
} mq_port_counts;

/* Where the ions ended (Frame, lib.rs:41-53, at the end of the run):
 * travelling ions by region at the step cap, then the terminal classes. */
typedef struct {
    uint64_t n, split, channel, combine, lost, back, detected;
    uint64_t ion_steps;
} mq_tandem_totals;

typedef struct mq_tandem mq_tandem;

/* Uploads the cube grid (splitter and combiner; its ac may be NULL) and the
 * port grid (shared by the nine ports; needs rf, ac and the DC bases the
 * drives use) once. The port grid's z_return / z_exit are the port's planes;
 * the cube grid's axis and z_inject are where ions are injected. */
int mq_tandem_create(const mq_fields* cube, const mq_fields* port, const mq_tandem_geom* geom, mq_tandem** out);
void mq_tandem_destroy(mq_tandem* t);

/* Runs every ion of every job (species, n, seed and an optional pressure;
 * the jobs' drives are ignored, the tandem drive applies; dc_scale must be 1;
 * at most MQ_TANDEM_MAX_JOBS jobs) until it is lost, back, detected or at the
 * step cap, on MQ_PATH_PLAIN (FP64) or MQ_PATH_BATCHED (FP32 fields and push,
 * FP64 position and clock). Physics: full or half kick, trilinear or nearest
 * fields, thermal collisions; legacy collisions and criteria are rejected.
 * Writes counts[MQ_TANDEM_PORTS] and *totals (either may be NULL). */
int mq_tandem_run(mq_tandem* t, const mq_job* jobs, size_t njobs, const mq_tandem_drive* d, const mq_physics* ph,
                  int32_t path, mq_port_counts* counts, mq_tandem_totals* totals);

/* As mq_tandem_run, and returns every ion's final state in job order: s
 * (positions in the frame of its final region), region, last port
 * (MQ_TANDEM_NO_PORT if none) and job. Any output pointer may be NULL. */
int mq_tandem_run_ions(mq_tandem* t, const mq_job* jobs, size_t njobs, const mq_tandem_drive* d,
                       const mq_physics* ph, int32_t path, mq_port_counts* counts, mq_tandem_totals* totals,
                       const mq_ion_soa* s, uint8_t* region, uint8_t* last, uint32_t* job);

/* Statistics of the last run (ion-steps, launches, device and wall time). */
int mq_tandem_last_stats(const mq_tandem* t, mq_run_stats* out);

/* Field access of the FP64 plain path (no counterpart in trapsim, whose
 * transport is a CPU loop): 1 (default) reads per-cell FP64 records of the four
 * field sets (576 B per cell each, built once per drive's DC sets when device
 * memory allows), 0 gathers the node tables at every step. Results are the
 * same; mq_run_stats.plain_recs reports which one the last run used. */
int mq_tandem_set_plain_recs(mq_tandem* t, int32_t on);

#ifdef __cplusplus
}
#endif

#endif /* MQ_TANDEM_H */
