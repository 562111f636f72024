/*
 * mq_legacy.h -- the legacy Java potentials, regenerated on the GPU (MI355X).
 *
 * The reference's legacy Fig. 4 columns (docs/mass-scan.csv, java_original
 * and rust_legacy) were computed on potentials that the Java sketch
 * MQ_MZ_ejection_singleQ exported and that Fields::from_theirs loads
 * (simulation/chsim/src/channel.rs:187-202) as eject_{RF,DC,AC}.f32 and
 * eject_{RF,AC}_electrode.u8. The snapshot does not hold those files. This
 * API regenerates them from the sketch's own algorithm:
 *   - float32 electrode rasterization of the ejection geometry
 *     (reference/MQ_MZ_ejection_singleQ/GEOMETRY.pde:11-47 calculateElectrodeGeometry,
 *     54-127 define_RF, 129-204 define_DC, 206-288 define_AC, 291-318 mapLayer,
 *     362-393 withinQuadElectrode, 407-414 SIGN, 463-467 withinHole,
 *     512-550 withinSqueezedQuadElectrode; dimensions
 *     MQ_MZ_ejection_singleQ.pde:27-73, 118, 152-162);
 *   - float32 lexicographic Gauss-Seidel sweeps (z outer, x, y inner), run
 *     while fewer than max_iter sweeps are done and the last sweep changed
 *     the potential at the check point by more than precision_pct percent
 *     (POTENTIAL.pde:31-84, called with 0.001 % and 6000 at IO.pde:180-183;
 *     check point CONVERSIONS.pde:12-15, MQ_MZ_ejection_singleQ.pde:71-73).
 * The GPU runs the sweeps as hyperplane wavefronts: node (i, j, k) takes its
 * sweep t at step i + j + k + 2t, when its three lower neighbours already
 * hold sweep t and its three upper neighbours still hold sweep t - 1, so
 * every update reads the operands the sequential sweep reads and the result
 * is bit-identical to it.
 *
 * Layout: node-major i + nx (j + ny k), i along x (fastest), k along z.
 * Conventions: 0 = success, < 0 = error (message in mq_last_error()).
 */
#ifndef MQ_LEGACY_H
#define MQ_LEGACY_H

#include <stddef.h>
#include <stdint.h>

#include "mq_fieldsolve.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The three exported bases (define_RF, define_DC, define_AC). */
enum { MQ_LEGACY_RF = 0, MQ_LEGACY_DC = 1, MQ_LEGACY_AC = 2 };

typedef struct {
    float scl;             /* nodes per mm (the sketch: 10, a 160 x 160 x 242 grid) */
    float precision_pct;   /* stop rule threshold in percent (the sketch: 0.001) */
    uint32_t max_iter;     /* sweep cap (the sketch: 6000) */
    int32_t pad_;
} mq_legacy_opts;

typedef struct {
    uint32_t sweeps;       /* sweeps run (the sketch's Nit) */
    float dif_pct;         /* change at the check point in the last sweep, percent */
    int32_t check[3];      /* check point node (i, j, k) */
    int32_t pad_;
    uint64_t node_updates; /* free-node updates over all sweeps */
    double ms;             /* device time of the sweeps (both passes) */
} mq_legacy_stats;

/* The sketch's settings: scl 10, 0.001 %, 6000 sweeps. */
void mq_legacy_default(mq_legacy_opts* o);

/* Grid size (int)(16 scl) x (int)(16 scl) x (int)(zDevice scl) and the check
 * point ((int)(9 scl), (int)(8.5 scl), (int)(14 scl)); check may be NULL. */
int mq_legacy_dims(const mq_legacy_opts* o, int32_t* nx, int32_t* ny, int32_t* nz, int32_t check[3]);

/* Initial state of one basis on the GPU: electrode nodes at their potential
 * with electrode = 1, every other node 0 (calculateElectrodeGeometry). */
int mq_legacy_rasterize(const mq_legacy_opts* o, int32_t basis, float* phi, uint8_t* electrode);

/* The sweeps of calculatePotential on any grid: phi (in/out) holds the start
 * state, nodes with electrode != 0 and the box faces are never updated.
 * check is the check point node; the stop rule is the sketch's (float32:
 * dif = |100 ((prev + 1e-10) - now) / (prev + 1e-10)|, prev starting at 1,
 * dif at 100). */
int mq_legacy_sweeps(int32_t nx, int32_t ny, int32_t nz, float* phi, const uint8_t* electrode,
                     const int32_t check[3], float precision_pct, uint32_t max_iter, mq_legacy_stats* stats);

/* Rasterize + sweeps for one basis; phi[N], electrode[N] (may be NULL). */
int mq_legacy_potential(const mq_legacy_opts* o, int32_t basis, float* phi, uint8_t* electrode,
                        mq_legacy_stats* stats);

/* All three bases written to dir as the files Fields::from_theirs reads:
 * eject_RF.f32, eject_DC.f32, eject_AC.f32 (float32 little-endian,
 * node-major) and eject_RF_electrode.u8, eject_AC_electrode.u8 (0/1).
 * stats[3] (may be NULL) in RF, DC, AC order. */
int mq_legacy_export(const mq_legacy_opts* o, const char* dir, mq_legacy_stats stats[3]);

/* Replaces Fields::from_theirs (channel.rs:187-202): reads the five files
 * from dir (a 160 x 160 x 242 grid at 0.1 mm), E by central differences on
 * the GPU (mq_efield), the RF potential's E into rf, AC into ac, the DC file
 * into the plate slot dc[1] (the drive then puts the DC voltage on dc[1]),
 * dc[0], dc[2], dc[3] zero, mask = RF electrode | AC electrode. phi[] gets
 * RF, AC and the DC file in slot 3 (DcPlate), the other slots zero. Any
 * output pointer may be NULL. Geometry scalars come from mq_geom_default. */
int mq_fields_from_theirs(const char* dir, mq_channel_fields* out);

#ifdef __cplusplus
}
#endif

#endif /* MQ_LEGACY_H */
