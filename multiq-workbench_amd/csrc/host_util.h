// host_util.h -- host helpers shared by the engine and the tandem batch:
// error type, HIP status check, pooled device buffers, and the per-job
// parameter derivation (Engine::job_params, channel.rs:412-426).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

#include "../../include/mq_engine.h"
#include "mq_types.h"

namespace mq {
namespace host {

constexpr double AMU = 1.6605390666e-27;
constexpr double QE = 1.602176634e-19;
constexpr double KB = 1.380649e-23;

struct Err : std::runtime_error {
    int code;
    Err(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                                     \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            throw Err(-3, std::string("HIP error ") + hipGetErrorString(e_) + " at " #expr);             \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t b)
    {
        if (b <= bytes) return;
        release();
        size_t nb = std::max<size_t>(b, 256);
        HIPCHK(hipMalloc(&p, nb));
        bytes = nb;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// channel.rs:376-381
inline double legacy_mfp_mean(const mq_species& sp, double pressure_pa)
{
    double p_torr = pressure_pa / 133.322 + 1e-10;
    double np = 2.68e5 * (p_torr / 760.0);
    double mm = sp.mass;
    double ccs_he = 2.81e-9 * mm * mm * mm - 3.55e-5 * mm * mm + 0.232 * mm + 41.91;
    double ccs = ccs_he * 10.0 * std::sqrt(sp.z);
    return 1.0 / (np * ccs);
}

// Engine::job_params (channel.rs:412-426), FP64.
inline JobP job_params(const mq_job& j, const mq_physics& ph, uint64_t off)
{
    JobP J{};
    double m = j.sp.mass * AMU, mg = 28.0 * AMU;
    double qm = j.sp.z * QE / m;
    double p = j.pressure_pa > 0.0 ? j.pressure_pa : ph.pressure_pa;
    double n_gas = p / (KB * ph.temp_k);
    double s_gas = std::sqrt(KB * ph.temp_k / mg);
    J.m = m;
    J.zc = j.sp.z;
    J.kq = (ph.legacy_force ? 0.5 : 1.0) * qm * ph.dt;
    J.v_rf = j.drive.v_rf;
    J.w_rf = j.drive.f_rf * ph.dt;
    J.v_ac = j.drive.v_ac;
    J.w_ac = j.drive.f_ac * ph.dt;
    J.s_dc = j.dc_scale;
    J.nsig_dt = n_gas * j.sp.ccs_n2 * 1e-20 * ph.dt;
    J.s_gas = s_gas;
    J.v_bound = 6.0 * s_gas;
    J.mr = mg / (m + mg);
    J.one_m_mr = 1.0 - J.mr;
    J.floor_v = ph.legacy_collisions ? std::sqrt(2.0 * 0.035 * 1.6e-19 / (j.sp.mass * 1.66e-27)) : -1.0;
    J.mfp_mean = legacy_mfp_mean(j.sp, p);
    J.key0 = (uint32_t)j.seed;
    J.key1 = (uint32_t)(j.seed >> 32);
    J.off = off;
    return J;
}

inline uint32_t max_steps_of(const mq_physics& ph)
{
    uint64_t m = ph.max_steps ? ph.max_steps : (uint64_t)std::llround(ph.t_max / ph.dt);
    if (m > 0xFFFFFFF0ull) m = 0xFFFFFFF0ull;
    if (m < 1) m = 1;
    return (uint32_t)m;
}


// Device tables and settings of an engine for a DC configuration (engine.cpp).
struct EngineView {
    hipStream_t stream = nullptr;
    int cus = 0;
    mq_tune tune{};
    mq_fields meta{};
    const float4* cells = nullptr;
    BatchP B{};
};
EngineView engine_view(mq_engine* e, const double dc[4], const mq_physics& ph, uint32_t njobs);

// Engine::inject with the reference's xorshift64 stream (inject_ref.cpp): the
// n ions of one job, as the reference stores them, plus 4 xoshiro128** seed
// words per ion.
void inject_reference(const mq_fields& f, const mq_job& job, const mq_physics& ph, uint64_t n, double* x, double* y,
                      double* z, double* vx, double* vy, double* vz, double* phi, uint32_t* cd, uint32_t* rng);

}  // namespace host
}  // namespace mq
