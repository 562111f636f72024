/*
 * oracle.h -- CPU FP64 oracle of the batched ion-trajectory integrator.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline. The product (multiq-workbench_amd/) never links,
 * loads or calls it.
 *
 * What it restates (reference file:line):
 *   step arithmetic and order ........ simulation/chsim/src/channel.metal:82-121
 *   trilinear/nearest sampling ....... channel.metal:55-68
 *   thermal null-collision bath ...... channel.metal:93-99 (form of world.rs:473-488)
 *   legacy stationary-gas collisions . channel.metal:100-113
 *   classification order ............. channel.metal:114-120
 *   job and batch parameters ......... simulation/chsim/src/channel.rs:397-426
 *   injection models ................. channel.rs:342-374 (legacy cone, thermal disc)
 *   summarize ........................ channel.rs:496-506
 *   DC combination ................... channel.rs:220-226 (in FP64 here)
 *
 * Deliberate, specified differences from the reference (the new design):
 *   - Random numbers: Philox4x32-10 keyed by the job seed with counter
 *     (ion index in job, step, sub-block, domain) instead of xorshift64 +
 *     xoshiro128** (channel.rs:343-373, channel.metal:38-44); with
 *     rng = MQ_RNG_COMPAT the reference's own streams and draw order.
 *   - FP64 position and velocity (the reference stores 62.5 pm fixed-point
 *     positions and FP32 velocities, channel.metal:31-36), RF/AC phases
 *     computed from the step counter in FP64 (the reference accumulates them
 *     in FP32, channel.metal:91-92), step cap in steps.
 *   - sin/cos/log are the self-written routines below, identical on the GPU,
 *     so the FP64 GPU path can be compared bit for bit.
 */
#ifndef MQ_ORACLE_H
#define MQ_ORACLE_H

#include <stdint.h>
#include <stddef.h>
#include "../include/mq_engine.h"
#include "../include/mq_fieldsolve.h"
#include "../include/mq_tandem.h"
#include "../include/mq_legacy.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 block (Salmon et al., SC'11). */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* sin(2 pi p), cos(2 pi p); log(x) for positive normal x. */
void or_sincos2pi(double p, double* s, double* c);
double or_log(double x);

/* Initial state of ion `ion` of job `job` (FP64). */
typedef struct {
    double x[3], v[3];
    double phi_ac;
    uint32_t steps, ncoll, cd;
    uint8_t status;
    uint32_t rng[4]; /* xoshiro128** state (MQ_RNG_COMPAT) */
} or_ion;

void or_inject(const mq_fields* f, const mq_job* job, const mq_physics* ph, uint32_t ion, or_ion* out);

/* Engine::inject (channel.rs:342-374) with the reference's xorshift64 stream:
 * all job->n ions of one job, in order (the stream is sequential). */
void or_inject_reference(const mq_fields* f, const mq_job* job, const mq_physics* ph, or_ion* out);

/* The generators of MQ_RNG_COMPAT: one xorshift64 step (13, 7, 17) and one
 * xoshiro128** output with its state update. */
uint64_t or_xorshift64(uint64_t* s);
uint32_t or_xoshiro128ss(uint32_t s[4]);

/* Runs every ion of every job to completion with `nthreads` CPU threads.
 * out_* are arrays of sum(n) entries, job order. Returns 0 on success. */
int or_run_batch(const mq_fields* f, const mq_job* jobs, size_t njobs, const double dc[4],
                 const mq_physics* ph, int nthreads, or_ion* out, uint32_t* out_job);

/* Runs jobs[0] ions [i0, i1) only (bounded CPU-baseline sample). */
int or_run_range(const mq_fields* f, const mq_job* job, const double dc[4], const mq_physics* ph,
                 uint32_t i0, uint32_t i1, int nthreads, or_ion* out);

void or_summarize(const or_ion* ions, const uint32_t* job, size_t n, size_t njobs, double dt, mq_outcome* out);

/* Differentiable form, forward-mode reference: integrates job->n ions for
 * obj->nsteps FP64 steps without termination, carrying 11 tangent directions
 * (dual numbers): the parameters (V_rf, f_rf, V_ac, f_ac, dc_scale) and the
 * initial state (x, y, z, vx, vy, vz). With gas (pressure > 0) the thermal
 * collisions of the Philox streams are applied with frozen draws: decisions on
 * the primal velocity, tangents through the collision map (pathwise
 * derivative). Returns J (mean objective), grad[5] and, if grad_state is not
 * NULL, dJ/d(initial state) as 6 arrays of n; -3 for legacy collisions or the
 * COMPAT streams with gas. */
int or_grad_reference(const mq_fields* f, const mq_job* job, const double dc[4], const mq_physics* ph,
                      const mq_objective* obj, int nthreads, double* J, double grad[5], double* grad_state);

/* The same for the listed ions of the job only: Ji[s] and gi[k * nions + s] =
 * objective of ion ions[s] and its derivative by parameter k (per ion, not
 * divided by job->n). Returns -2 if an index is out of range. */
int or_grad_ions(const mq_fields* f, const mq_job* job, const double dc[4], const mq_physics* ph,
                 const mq_objective* obj, int nthreads, const uint32_t* ions, uint32_t nions, double* Ji, double* gi);

/* Trilinear sample of one node-major FP64 basis at grid coordinates g. */
void or_sample(const double* F, int nx, int ny, int nz, const double g[3], int nearest, double out[3]);

/* Tandem multi-domain batch (trapsim Model::advance / enter / leave,
 * simulation/trapsim/src/lib.rs:176-385) in the FP64 operation order of the
 * GPU plain path. out/out_job/region/last: sum(n) entries in job order;
 * counts: 9 x 24 u64 (entered, returned, forwarded, revisits, from_other,
 * reverse entries, lost, retry_from[9], by_job[8]); totals: n, split,
 * channel, combine, lost, back, detected, ion-steps. */
int or_tandem_run(const mq_fields* cube, const mq_fields* port, const mq_tandem_geom* g, const mq_job* jobs,
                  size_t njobs, const mq_tandem_drive* d, const mq_physics* ph, int nthreads, or_ion* out,
                  uint32_t* out_job, uint8_t* region, uint8_t* last, uint64_t* counts, uint64_t* totals);

/* ---- basis-field pipeline (oracle_fields.c) ---- */

/* Geom::default, layers(), nz() (channel.rs:38-57). */
void or_geom_default(mq_geom* g);
void or_geom_layers(const mq_geom* g, double l[8]);
int or_geom_nz(const mq_geom* g);

/* Geom::classify (channel.rs:59-103): MQ_ELEC_* code of node (i, j, k), and
 * of every node of the n_xy x n_xy x nz grid. */
int or_geom_classify_node(const mq_geom* g, int i, int j, int k);
void or_geom_classify(const mq_geom* g, uint8_t* codes);

/* basis_value (channel.rs:110-119) of an MQ_ELEC_* code. */
double or_basis_value(int basis, int code, int ac_mode);

/* The 7-point Laplace problem of fieldsolve::solve by red-black SOR (small
 * grids): returns 0 when the scaled max-norm residual is <= tol. */
int or_laplace_sor(int nx, int ny, int nz, const uint8_t* mask, const double* fixed, double outer_bc, double tol,
                   uint32_t max_iter, double* phi, uint32_t* iters, double* residual);

/* efield (channel.rs:168-184): e[3N] = -grad(phi). */
void or_efield(const double* phi, int nx, int ny, int nz, double h, double* e);

/* ---- the Java sketch's potential export (oracle_legacy.c) ---- */

/* Grid size and check point node of the ejection sketch at scl nodes/mm. */
int or_legacy_dims(float scl, int* nx, int* ny, int* nz, int check[3]);

/* define_RF / define_DC / define_AC (basis MQ_LEGACY_*) of every node:
 * phi[N] and el[N] (1 = electrode), node-major i + nx (j + ny k). */
void or_legacy_rasterize(float scl, int basis, float* phi, uint8_t* el);

/* One sequential sweep of calculatePotential (z outer, x, y inner). */
void or_legacy_sweep(int nx, int ny, int nz, float* phi, const uint8_t* el);

/* The stop rule's change |100 ((prev + 1e-10) - now) / (prev + 1e-10)|. */
float or_legacy_dif(float prev, float now);

/* calculatePotential's loop: sweeps while nit < max_iter and dif > precision;
 * returns the sweep count, the last dif, and (trace != NULL) the check point
 * value after every sweep. */
uint32_t or_legacy_sweeps(int nx, int ny, int nz, float* phi, const uint8_t* el, const int check[3], float precision,
                          uint32_t max_iter, float* dif_out, float* trace);

#ifdef __cplusplus
}
#endif
#endif
