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
    int32_t nx, ny, nz, nearest, legacy_coll, pad;
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
};

// Per-job accumulators (u64): [0] n, [1..5] crashed, returned, transmitted,
// left, timeout, [6] sum of steps, [7] sum of collisions.
constexpr int ACC = 8;

// FP32 cell record of the batched path: trilinear polynomial coefficients
// c0..c7 (1, tx, ty, tz, tx ty, ty tz, tx tz, tx ty tz) x 3 components for the
// RF, combined-DC and AC bases, then the conductor mask of the 8 corners
// (bit dx + 2 dy + 4 dz), padded to 19 float4 = 304 B.
constexpr int CELL_F4 = 19;

}  // namespace mq
