// Record compatibility with the reference: the 64-byte ChIon export, summarize
// over ChIon records and the FNV-1a hash that check.rs prints for a run.
//
// Host-side format code: it converts final states that mq_run_batch_ions has
// already brought back; no integration happens here.
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mq_engine.h"

void mq_set_last_error(const std::string& msg);  // engine.cpp

static_assert(sizeof(mq_chion) == 64, "ChIon is 64 bytes (channel.rs:292-294)");

namespace {

// Phase in cycles, [0, 1): frac(phi0 + w * steps) with the product and sum
// formed as the kernels form them (one fma, then the floor).
double phase_at(double w, uint32_t steps, double phi0)
{
    const double p = std::fma(w, (double)steps, phi0);
    return p - std::floor(p);
}

int32_t fixed_point(double x)
{
    const double q = std::nearbyint(x * MQ_POS_SCALE);
    if (!(q >= -2147483648.0 && q <= 2147483647.0)) throw std::runtime_error("position outside the fixed-point range");
    return (int32_t)q;
}

}  // namespace

extern "C" {

int mq_export_chion(const mq_ion_soa* s, const uint32_t* job, size_t n, const mq_job* jobs, size_t njobs,
                    const mq_physics* ph, mq_chion* out)
{
    try {
        if (!s || !ph || (n && (!job || !jobs || !out))) throw std::runtime_error("NULL argument");
        if (n && (!s->x || !s->y || !s->z || !s->vx || !s->vy || !s->vz || !s->steps || !s->ncoll || !s->status))
            throw std::runtime_error("mq_ion_soa: a required array is NULL");
        // First ion of each job: ions are concatenated in job order.
        std::vector<size_t> off(njobs + 1, 0);
        for (size_t k = 0; k < njobs; ++k) off[k + 1] = off[k] + (size_t)jobs[k].n;
        for (size_t i = 0; i < n; ++i) {
            const uint32_t k = job[i];
            if (k >= njobs) throw std::runtime_error("job index out of range");
            if (i < off[k] || i >= off[k + 1]) throw std::runtime_error("ions are not in job order");
            const mq_job& J = jobs[k];
            mq_chion r;
            r.p[0] = fixed_point(s->x[i]);
            r.p[1] = fixed_point(s->y[i]);
            r.p[2] = fixed_point(s->z[i]);
            r.p[3] = (int32_t)s->status[i];
            r.v[0] = (float)s->vx[i];
            r.v[1] = (float)s->vy[i];
            r.v[2] = (float)s->vz[i];
            r.v[3] = (float)((double)s->steps[i] * ph->dt);
            r.aux[0] = s->phi_ac0 ? (float)phase_at(J.drive.f_ac * ph->dt, s->steps[i], s->phi_ac0[i]) : 0.0f;
            r.aux[1] = (float)s->ncoll[i];
            r.aux[2] = s->cd ? (float)s->cd[i] : 0.0f;
            r.aux[3] = (float)phase_at(J.drive.f_rf * ph->dt, s->steps[i], 0.0);  // RF phase 0 at injection
            r.rng[0] = (uint32_t)J.seed;
            r.rng[1] = (uint32_t)(J.seed >> 32);
            r.rng[2] = (uint32_t)(i - off[k]);
            r.rng[3] = s->steps[i];
            out[i] = r;
        }
        return 0;
    } catch (const std::exception& e) {
        mq_set_last_error(e.what());
        return -1;
    }
}

int mq_summarize_chion(const mq_chion* ions, const uint32_t* job, size_t n, size_t njobs, mq_outcome* out)
{
    try {
        if (!out || (n && (!ions || !job))) throw std::runtime_error("NULL argument");
        for (size_t k = 0; k < njobs; ++k) out[k] = mq_outcome{};
        for (size_t i = 0; i < n; ++i) {
            if (job[i] >= njobs) throw std::runtime_error("job index out of range");
            mq_outcome& o = out[job[i]];
            o.n += 1;
            switch (ions[i].p[3]) {
                case 1: o.crashed += 1; break;
                case 2: o.returned += 1; break;
                case 3: o.transmitted += 1; break;
                case 4: o.left += 1; break;
                default: o.timeout += 1; break;
            }
            o.mean_coll += (double)ions[i].aux[1];
            o.mean_time += (double)ions[i].v[3];
        }
        for (size_t k = 0; k < njobs; ++k)
            if (out[k].n) {
                out[k].mean_coll /= (double)out[k].n;
                out[k].mean_time /= (double)out[k].n;
            }
        return 0;
    } catch (const std::exception& e) {
        mq_set_last_error(e.what());
        return -1;
    }
}

uint64_t mq_fnv1a64(const void* data, size_t nbytes)
{
    const unsigned char* b = (const unsigned char*)data;
    uint64_t h = 14695981039346656037ull;
    for (size_t i = 0; i < nbytes; ++i) {
        h ^= (uint64_t)b[i];
        h *= 1099511628211ull;
    }
    return h;
}

}  // extern "C"
