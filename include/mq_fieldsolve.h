/*
 * mq_fieldsolve.h -- basis-field grids of the channel geometry (MI355X).
 *
 * The step before the integrator: electrode rasterization of the parametric
 * channel, one Laplace solve per basis on the GPU, E = -grad(phi) by central
 * differences and the conductor mask, as the reference's Fields::build does
 * (simulation/chsim/src/channel.rs:18-217, fieldsolve/src/lib.rs:104-166).
 * The output feeds mq_engine_create (mq_fields) unchanged.
 *
 * Conventions: 0 = success, < 0 = error (message in mq_last_error()).
 */
#ifndef MQ_FIELDSOLVE_H
#define MQ_FIELDSOLVE_H

#include <stddef.h>
#include <stdint.h>

#include "mq_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces Geom (channel.rs:18-36); lengths in mm. Booleans as 0/1. */
typedef struct {
    double de;            /* rod diameter */
    double coef;          /* gap / de */
    double q1_start, q1_len;
    double alpha;         /* squeeze per mm of z */
    double gap, plate, hole_d, q2_len, stem_d;
    int32_t q2_flare_bug; /* Q2 reuses the Q1 squeeze formula */
    double grow;          /* convex electrodes grown by grow * h */
    int32_t tilt_mode;    /* 0 all four Q1 rods shifted in y, 1 upper pair converges in x, 2 radial */
    int32_t ac_mode;      /* 0 left rods -1 / right +1, 1 AC on the lower pair only */
    double mis[4][4];     /* per Q1 rod [LL,RL,LU,RU]: dx, dy [mm], dx/dz, dy/dz [rad] */
    int32_t n_xy;         /* nodes across (x and y) */
    double h_mm;          /* node spacing */
} mq_geom;

/* Electrode codes of mq_geom_classify (Elec, channel.rs:45-46): quadrant q
 * = 0 LL, 1 RL, 2 LU, 3 RU is added to the rod families. */
enum {
    MQ_ELEC_NONE = 0,
    MQ_ELEC_Q1 = 1,      /* 1..4 */
    MQ_ELEC_STEM3 = 5,   /* 5..8 */
    MQ_ELEC_PLATE = 9,
    MQ_ELEC_STEM5 = 10,  /* 10..13 */
    MQ_ELEC_Q2 = 14,     /* 14..17 */
    MQ_ELEC_STEM7 = 18,  /* 18..21 */
    MQ_ELEC_SLAB = 22
};

/* Basis identifiers (Basis, channel.rs:105-107). */
enum { MQ_BASIS_RF = 0, MQ_BASIS_AC, MQ_BASIS_DC_Q1, MQ_BASIS_DC_PLATE, MQ_BASIS_DC_DOWN, MQ_BASIS_DC_BOX };

/* Geom::default (channel.rs:38-43). */
void mq_geom_default(mq_geom* g);

/* Grid size (n_xy, n_xy, nz(); channel.rs:55) and the eight layer planes
 * (layers(), channel.rs:49-53) [mm]; layers may be NULL. */
int mq_geom_dims(const mq_geom* g, int32_t* nx, int32_t* ny, int32_t* nz, double layers[8]);

/* Fills the geometry-derived scalars of an mq_fields (grid size, spacing in
 * metres, axis, injection plane, return/exit planes and the legacy crash
 * window, as Engine::batch_params and inject derive them, channel.rs:342-410).
 * The field pointers are left untouched. */
int mq_geom_fields_params(const mq_geom* g, mq_fields* f);

/* Electrode rasterization on the GPU: one MQ_ELEC_* code per node,
 * node-major i + nx (j + ny k) (Geom::classify, channel.rs:59-103). */
int mq_geom_classify(const mq_geom* g, uint8_t* codes);

/* Solver statistics (SolveStats, fieldsolve/src/mg.rs). */
typedef struct {
    uint32_t iterations;
    double residual;      /* max over free nodes |sum of 6 neighbours - 6 phi| / max(|fixed|, |outer|) */
    int32_t converged;
    double ms;            /* device time */
} mq_solve_stats;

/* Replaces fieldsolve::solve (fieldsolve/src/lib.rs:150-166): the 7-point
 * Laplace system on an nx x ny x nz grid, nodes with mask != 0 fixed at
 * fixed[idx], unmasked box-face nodes fixed at outer_bc, solved on the GPU
 * (conjugate gradients in FP64) until the true max-norm residual above is
 * <= tol or max_iter iterations (0 = 20000). phi[N] receives the solution.
 * Returns 0 when converged, -5 when not (phi holds the last iterate). */
int mq_laplace_solve(int32_t nx, int32_t ny, int32_t nz, const uint8_t* mask, const double* fixed, double outer_bc,
                     double tol, uint32_t max_iter, double* phi, mq_solve_stats* stats);

/* E = -grad(phi) per unit potential by central differences, one-sided at the
 * box faces (efield, channel.rs:168-184), on the GPU; e[3N] node-major. */
int mq_efield(const double* phi, int32_t nx, int32_t ny, int32_t nz, double h, double* e);

/* Output of mq_fields_build, caller-allocated (N = nx ny nz from
 * mq_geom_dims). Any pointer may be NULL to skip that output. */
typedef struct {
    double* rf;           /* 3N: E of the RF basis per volt */
    double* ac;           /* 3N */
    double* dc[4];        /* 3N each: DcQ1, DcPlate, DcDown, DcBox */
    double* phi[6];       /* N each: potentials in basis order */
    uint8_t* mask;        /* N: 1 = electrode node */
    mq_solve_stats stats[6];
} mq_channel_fields;

/* Replaces Fields::build (channel.rs:204-217) with solve_basis
 * (channel.rs:140-166): classify, the six basis solves at tolerance tol,
 * E and mask. With cache_dir non-NULL, each potential is read from / written
 * to cache_dir/ch_<fnv>.bin in the reference's cache format (raw f64
 * little-endian, channel.rs:130-139, 148-165). */
int mq_fields_build(const mq_geom* g, double tol, const char* cache_dir, mq_channel_fields* out);

/* Cache file name of one basis, "ch_%016x.bin" of the FNV-1a hash of the
 * Rust Debug strings of the geometry and the basis (cache_path,
 * channel.rs:130-135); same return convention as mq_scan_key. */
int mq_field_cache_name(const mq_geom* g, int32_t basis, char* buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* MQ_FIELDSOLVE_H */
