// mq_types.h -- device-side parameter blocks and the ion-state layout.
#pragma once

#include <stdint.h>

namespace mq {

// Per-job parameters, derived on the host in FP64 (ChJob, channel.rs:285-290 and
// Engine::job_params, channel.rs:412-426). kq = qm_kick * dt, w_* = f * dt
// (cycles per step), s_dc = per-job DC scale.
struct JobP {
    double kq, v_rf, w_rf, v_ac, w_ac, s_dc;
    double nsig_dt, s_gas, v_bound, mr, one_m_mr, floor_v, mfp_mean, m, zc;
    uint32_t key0, key1;
    uint64_t off;  // global index of the job's first ion
};

// Batch-wide parameters (ChParams, channel.rs:276-282 and batch_params 397-410).
struct BatchP {
    double dt, h, inv_h, lim0, lim1, lim2;
    double z_ret, z_exit, cz0, cz1;
    int32_t nx, ny, nz, nearest, legacy_coll, compat;  // compat: MQ_RNG_COMPAT streams
    uint32_t max_steps;
    uint32_t njobs;
};

// Injection parameters (Engine::inject, channel.rs:342-374).
struct InjP {
    double axis_x, axis_y, z_inject, temp_k, r_inject_mm, inj_sigma_mm, dt;
    int32_t model;  // 0 thermal disc, 1 legacy cone, 2 gaussian spot
    int32_t pad;
};

// Ion state, structure of arrays in HBM. FP64 position, AC phase offset and
// (plain path) velocity; FP32 velocity for the batched path.
struct IonState {
    double *x, *y, *z, *phi;
    double *vx, *vy, *vz;
    float *fvx, *fvy, *fvz;
    uint32_t *steps, *ncoll, *cd, *job;
    uint8_t *status;
    uint8_t *region = nullptr, *last = nullptr;  // tandem batch only: domain and previous port
    uint32_t* rng = nullptr;  // MQ_RNG_COMPAT only: xoshiro128** state, 4 words per ion [4 i + k]
    // live population only (batched path): RF phase at the ion's birth [cycles] and its random-stream
    // number; without them the RF phase starts at 0 and the stream number is the ion's index in its job
    double* phr = nullptr;
    uint32_t* sid = nullptr;
};

// Per-job accumulators (u64): [0] n, [1..5] crashed, returned, transmitted,
// left, timeout, [6] sum of steps, [7] sum of collisions, [8] ions whose status
// is not a final one (never finished: the engine reports these as an error).
constexpr int ACC = 9;
constexpr int ACC_UNFINISHED = 8;

// FP32 cell record of the batched path: trilinear polynomial coefficients
// c0..c7 (1, tx, ty, tz, tx ty, ty tz, tx tz, tx ty tz) x 3 components for the
// RF, combined-DC and AC bases, then the conductor mask of the 8 corners
// (bit dx + 2 dy + 4 dz): 19 float4 = 304 B loaded per cell change. Records
// sit at a stride of 20 float4 = 320 B, so each starts on a 64-byte boundary
// and spans exactly three 128-byte lines (a 304-byte stride leaves 1 record in 4
// straddling a fourth line).
#ifndef MQ_CELL_STRIDE_F4
#define MQ_CELL_STRIDE_F4 20
#endif
constexpr int CELL_F4 = 19;
constexpr int CELL_STRIDE_F4 = MQ_CELL_STRIDE_F4;
static_assert(CELL_STRIDE_F4 >= CELL_F4, "cell record stride below its size");

// ---- tandem batch (kernels_tandem.hip)
// One grid of the tandem: FP64 bases and mask (plain path), geometry.
struct TGrid {
    const double *rf, *ac;
    const uint8_t* mask;
    double inv_h, lim0, lim1, lim2;
    int32_t nx, ny, nz, pad;
};

// Coupling geometry and drive, FP64; w_* = f * dt (cycles per step).
struct TandemP {
    double dt, seam, ap_r2, ent_x, ent_z, l_ext, z_ret, z_exit, ax_p, ay_p;
    double mx[9], my[9];
    double v_rf, w_rf, v_ac;
    double w_ac[9];
    uint32_t max_steps, nearest, pad0, pad1;
};

// Per-port u64 counters: entered, returned, forwarded, revisits, from_other,
// reverse entries, lost, retry_from[9], by_job[8].
constexpr int TC_PORT = 24;
constexpr int TC_ENTERED = 0, TC_RETURNED = 1, TC_FORWARDED = 2, TC_REVISITS = 3, TC_FROM_OTHER = 4,
              TC_REVERSE = 5, TC_LOST = 6, TC_RETRY = 7, TC_BYJOB = 16;
constexpr int TC_MAX_JOBS = 8;
// Totals accumulated at the end: n, split, channel, combine, lost, back, detected, steps.
constexpr int TT_N = 8;
constexpr uint32_t REG_SPLIT = 0, REG_COMBINE = 10, NO_PORT = 255;

}  // namespace mq
