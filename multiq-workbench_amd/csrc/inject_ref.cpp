// inject_ref.cpp -- the reference's injection with its own random stream
// (the injection of MQ_RNG_COMPAT), on the host as the reference does it.
//
// Engine::inject (simulation/chsim/src/channel.rs:342-374): one xorshift64
// stream per job, state = seed * 0x9E3779B97F4A7C15 ^ 0xD1B54A32D192ED03,
// x ^= x << 13, x ^= x >> 7, x ^= x << 17, u = (x >> 11) / 2^53; per ion, in
// order, the draws of the injection model, the AC phase, the first legacy
// free path, then four words that seed the ion's xoshiro128** stream (the
// first with its low bit set). The stored state is the reference's ChIon:
// positions truncated to the 62.5 pm fixed-point grid, velocities and the AC
// phase rounded to f32, the countdown as f32.
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/mq_engine.h"
#include "host_util.h"
#include "mq_types.h"

void mq_set_last_error(const std::string& msg);  // engine.cpp

namespace mq {
namespace host {

namespace {

struct XorShift64 {
    uint64_t s;
    explicit XorShift64(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull ^ 0xD1B54A32D192ED03ull) {}
    uint64_t next()
    {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return s;
    }
    double u() { return (double)(next() >> 11) / 9007199254740992.0; }  // 2^53
};

constexpr double PI = 3.14159265358979323846;

}  // namespace

void inject_reference(const mq_fields& f, const mq_job& job, const mq_physics& ph, uint64_t n, double* x, double* y,
                      double* z, double* vx, double* vy, double* vz, double* phi, uint32_t* cd, uint32_t* rng)
{
    XorShift64 g(job.seed);
    const double m = job.sp.mass * AMU;
    const double mfp_mean = legacy_mfp_mean(job.sp, job.pressure_pa > 0.0 ? job.pressure_pa : ph.pressure_pa);
    for (uint64_t i = 0; i < n; ++i) {
        double pos[3], vel[3];
        if (ph.injection == MQ_INJ_GAUSSIAN) {
            // Gaussian spot and Maxwell-Boltzmann start (no reference counterpart): pairs (r, theta)
            const double sig = ph.inj_sigma_mm * 1e-3, s = std::sqrt(KB * ph.temp_k / m);
            double r = std::sqrt(-2.0 * std::log(std::fmax(g.u(), 1e-300))), th = 2.0 * PI * g.u();
            pos[0] = f.axis_x + sig * (r * std::cos(th));
            pos[1] = f.axis_y + sig * (r * std::sin(th));
            pos[2] = f.z_inject;
            r = std::sqrt(-2.0 * std::log(std::fmax(g.u(), 1e-300)));
            th = 2.0 * PI * g.u();
            vel[0] = s * (r * std::cos(th));
            vel[1] = s * (r * std::sin(th));
            r = std::sqrt(-2.0 * std::log(std::fmax(g.u(), 1e-300)));
            th = 2.0 * PI * g.u();
            vel[2] = s * (r * std::cos(th));
        } else if (ph.legacy_injection) {
            const double kin = 0.1 * (0.75 + 0.75 * g.u()) * job.sp.z;  // eV
            const double vin = std::sqrt(2.0 * kin * QE / m);
            const double ang = (-10.0 + 20.0 * g.u()) * (PI / 180.0);
            const double vxy = 0.707 * vin * std::sin(ang);
            pos[0] = f.axis_x; pos[1] = f.axis_y; pos[2] = f.z_inject;
            vel[0] = vxy; vel[1] = vxy; vel[2] = vin * std::cos(ang);
        } else {
            const double s = std::sqrt(KB * ph.temp_k / m);
            const double r = ph.r_inject_mm * 1e-3 * std::sqrt(g.u());
            const double th = 2.0 * PI * g.u();
            auto gs = [&](double u1, double u2) {
                return std::sqrt(-2.0 * std::log(std::fmax(u1, 1e-300))) * std::cos(2.0 * PI * u2);
            };
            const double vzz = s * std::sqrt(-2.0 * std::log(std::fmax(g.u(), 1e-300)));
            pos[0] = f.axis_x + r * std::cos(th);
            pos[1] = f.axis_y + r * std::sin(th);
            pos[2] = f.z_inject;
            const double u1 = g.u(), u2 = g.u();
            vel[0] = s * gs(u1, u2);
            const double u3 = g.u(), u4 = g.u();
            vel[1] = s * gs(u3, u4);
            vel[2] = vzz;
        }
        const double phase = 2.0 * PI * g.u();
        const double vmag = std::sqrt(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
        const double first_mfp = -std::log(std::fmax(g.u(), 1e-7)) * mfp_mean / 10.0;
        const double steps0 = std::floor(first_mfp / vmag / ph.dt) + 1.0;
        const double* outp[3] = {x, y, z};
        for (int d = 0; d < 3; ++d) {
            const double q = pos[d] * MQ_POS_SCALE;
            // `as i32`: truncation toward zero, saturating
            const int32_t p = !(q == q) ? 0 : q >= 2147483647.0 ? 2147483647 : q <= -2147483648.0 ? INT32_MIN : (int32_t)q;
            const_cast<double*>(outp[d])[i] = (double)p / MQ_POS_SCALE;
        }
        vx[i] = (double)(float)vel[0];
        vy[i] = (double)(float)vel[1];
        vz[i] = (double)(float)vel[2];
        phi[i] = (double)(float)(phase / (2.0 * PI));
        const float s0 = (float)steps0;
        cd[i] = !(s0 >= 1.0f) ? 1u : (s0 >= 4294967040.0f ? 4294967040u : (uint32_t)s0);
        rng[4 * i + 0] = (uint32_t)g.next() | 1u;
        rng[4 * i + 1] = (uint32_t)g.next();
        rng[4 * i + 2] = (uint32_t)g.next();
        rng[4 * i + 3] = (uint32_t)g.next();
    }
}

}  // namespace host
}  // namespace mq

extern "C" int mq_inject_reference(const mq_fields* f, const mq_job* job, const mq_physics* ph, const mq_ion_soa* s,
                                   uint32_t* rng)
{
    try {
        if (!f || !job || !ph || !s || !rng) throw std::runtime_error("NULL argument");
        if (!s->x || !s->y || !s->z || !s->vx || !s->vy || !s->vz || !s->phi_ac0 || !s->cd)
            throw std::runtime_error("mq_ion_soa: x, y, z, vx, vy, vz, phi_ac0 and cd are required");
        if (!(ph->dt > 0.0)) throw std::runtime_error("dt must be positive");
        mq::host::inject_reference(*f, *job, *ph, job->n, s->x, s->y, s->z, s->vx, s->vy, s->vz, s->phi_ac0, s->cd,
                                   rng);
        if (s->steps)
            for (uint64_t i = 0; i < job->n; ++i) s->steps[i] = 0;
        if (s->ncoll)
            for (uint64_t i = 0; i < job->n; ++i) s->ncoll[i] = 0;
        if (s->status)
            for (uint64_t i = 0; i < job->n; ++i) s->status[i] = 0;
        return 0;
    } catch (const std::exception& e) {
        mq_set_last_error(e.what());
        return -1;
    }
}
