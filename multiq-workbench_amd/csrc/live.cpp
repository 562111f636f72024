// live.cpp -- a live ion population on the engine (include/mq_live.h): the
// block loop of mqviz's Sim::block (simulation/mqviz/src/engine.rs:448-538)
// with the transport on the batched FP32 worker (kernels_batched.hip).
//
// Device: a pool of `capacity` ion slots (structure of arrays, as the engine's
// batches), two ping-pong live lists of slot ids with their counter blocks,
// the finished-ion records of the last block and the job table: one job per
// (AC tone, species), tone 0 the live tone and tone p + 1 the clock of sweep
// point p, job = tone * nsp + species. Host: the free slots, each slot's
// species, tone, tag and birth step, the tones' phases, the xorshift64 stream
// of the block's draws, the re-entry queue, the tallies and the sweep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <queue>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mq_engine.h"
#include "../../include/mq_live.h"
#include "host_util.h"
#include "kernels.h"
#include "mq_types.h"

using namespace mq;
using namespace mq::host;

void mq_set_last_error(const std::string& msg);  // engine.cpp

namespace {

// The reference's xorshift64 step and uniform (engine.rs:319, 367-368).
double xs_u01(uint64_t& r)
{
    r ^= r << 13;
    r ^= r >> 7;
    r ^= r << 17;
    return (double)(r >> 11) / (double)(1ull << 53);
}

// poisson (engine.rs:318-328): product of uniforms below 30, else a rounded normal.
uint64_t poisson(uint64_t& r, double lam)
{
    if (lam < 30.0) {
        const double l = std::exp(-lam);
        uint64_t k = 0;
        double p = 1.0;
        for (;;) {
            p *= xs_u01(r);
            if (p <= l) return k;
            ++k;
        }
    }
    const double u1 = xs_u01(r);
    const double u2 = xs_u01(r);
    const double g = std::sqrt(-2.0 * std::log(std::max(u1, 1e-300))) * std::cos(2.0 * M_PI * u2);
    return (uint64_t)std::max(0.0, std::round(lam + std::sqrt(lam) * g));
}

double frac(double x) { return x - std::floor(x); }

}  // namespace

struct mq_live {
    mq_engine* eng = nullptr;
    EngineView V;
    mq_physics ph{};
    mq_rfac rf{};
    double dc[4] = {0, 0, 0, 0};
    mq_live_opts o{};
    uint32_t nsp = 0, cap = 0;
    std::vector<double> cdf;
    std::vector<JobP> jp;  // [tone][species]
    InjP ip{};
    // device
    DevBuf x, y, z, phi, phr, vx, vy, vz, fvx, fvy, fvz, steps, ncoll, cd, job, status, sid;
    DevBuf jobs, list[2], ctr, fin, nfin, rec;
    IonState S{};
    int cur = 0;
    // host
    uint64_t step = 0;
    std::vector<double> tone_ph;  // phase of each AC tone at `step` [cycles]
    double f_live = 0.0;          // the live tone's frequency outside a sweep [Hz]
    uint64_t rng = 1;
    uint32_t next_id = 0;
    int32_t tag = 0;
    uint32_t fresh = 0;  // slots [fresh, cap) were never used; they are taken before recycled ones
    std::vector<uint32_t> free_slots, slot_sp, slot_sid;
    std::vector<uint8_t> slot_tone;
    std::vector<int32_t> slot_tag;
    std::vector<uint64_t> slot_birth;
    std::priority_queue<std::pair<uint64_t, uint32_t>, std::vector<std::pair<uint64_t, uint32_t>>,
                        std::greater<std::pair<uint64_t, uint32_t>>>
        returns;  // (due step, species)
    std::vector<uint64_t> tally;  // [tag][species][8]: n, classes 1..5, steps, collisions
    mq_live_state st{};
    // the sweep (SweepRun, engine.rs:420-432): points [kHz], counts [point][species][4]
    struct Sweep {
        mq_live_sweep_cfg cfg{};
        std::vector<double> f;
        std::vector<uint64_t> acc;
        uint32_t point = 0, inj_in_point = 0, sweeps_done = 0;
        bool running = false, draining = false;
    } sw;
    uint64_t tagged_alive = 0;  // live ions on a sweep tone
    std::vector<LiveNew> pending;
    std::vector<uint32_t> fin_h;

    uint32_t* ctr_of(int k) { return ctr.as<uint32_t>() + (size_t)k * CTR_BLOCK; }

    void create(mq_engine* e, const mq_species* sp, const double* w, size_t n, const mq_rfac& drive, const double d[4],
                const mq_physics& p, const mq_live_opts& opt)
    {
        if (!e || !sp || n == 0) throw Err(-1, "a live population needs an engine and at least one species");
        if (n > 64) throw Err(-1, "at most 64 species");
        if (opt.block_steps == 0 || opt.block_steps % 4) throw Err(-1, "block_steps must be a positive multiple of 4");
        if (opt.capacity == 0 || opt.capacity > (1u << 26)) throw Err(-1, "capacity must be in [1, 2^26]");
        if (opt.max_tags < 1 || opt.max_tags > 4096) throw Err(-1, "max_tags must be in [1, 4096]");
        if (!(opt.inj_rate >= 0.0) || !(opt.residence_s >= 0.0)) throw Err(-1, "inj_rate and residence_s must be >= 0");
        if (!(p.dt > 0.0)) throw Err(-1, "dt must be positive");
        if (p.rng != MQ_RNG_PHILOX) throw Err(-1, "a live population uses the Philox streams (rng = 0)");
        eng = e;
        ph = p;
        rf = drive;
        o = opt;
        nsp = (uint32_t)n;
        cap = opt.capacity;
        for (int b = 0; b < 4; ++b) dc[b] = d[b];
        double tot = 0.0;
        for (size_t k = 0; k < n; ++k) tot += w ? w[k] : 1.0;
        if (!(tot > 0.0)) throw Err(-1, "species weights must sum to a positive number");
        double acc = 0.0;
        for (size_t k = 0; k < n; ++k) {
            acc += (w ? w[k] : 1.0) / tot;
            cdf.push_back(acc);
        }
        V = engine_view(e, dc, ph, nsp);
        jp.resize(nsp);
        for (uint32_t k = 0; k < nsp; ++k) {
            mq_job j{};
            j.sp = sp[k];
            j.drive = rf;
            j.n = 0;
            j.seed = opt.seed + k;
            j.dc_scale = 1.0;
            jp[k] = job_params(j, ph, 0);
        }
        ip.axis_x = V.meta.axis_x; ip.axis_y = V.meta.axis_y; ip.z_inject = V.meta.z_inject;
        ip.temp_k = ph.temp_k; ip.r_inject_mm = ph.r_inject_mm; ip.inj_sigma_mm = ph.inj_sigma_mm; ip.dt = ph.dt;
        ip.model = ph.injection == MQ_INJ_GAUSSIAN ? 2 : (ph.legacy_injection ? 1 : 0);
        const size_t c = cap;
        for (DevBuf* b : {&x, &y, &z, &phi, &phr, &vx, &vy, &vz}) b->ensure(c * 8);
        for (DevBuf* b : {&fvx, &fvy, &fvz, &steps, &ncoll, &cd, &job, &sid, &list[0], &list[1]}) b->ensure(c * 4);
        status.ensure(c);
        jobs.ensure((size_t)nsp * (1 + MQ_LIVE_MAX_POINTS) * sizeof(JobP));
        ctr.ensure(2 * CTR_BLOCK * 4);
        fin.ensure(c * 16);
        nfin.ensure(4);
        rec.ensure(c * sizeof(LiveNew));
        S.x = x.as<double>(); S.y = y.as<double>(); S.z = z.as<double>(); S.phi = phi.as<double>();
        S.vx = vx.as<double>(); S.vy = vy.as<double>(); S.vz = vz.as<double>();
        S.fvx = fvx.as<float>(); S.fvy = fvy.as<float>(); S.fvz = fvz.as<float>();
        S.steps = steps.as<uint32_t>(); S.ncoll = ncoll.as<uint32_t>(); S.cd = cd.as<uint32_t>();
        S.job = job.as<uint32_t>(); S.status = status.as<uint8_t>();
        S.phr = phr.as<double>(); S.sid = sid.as<uint32_t>();
        hipStream_t s = V.stream;
        HIPCHK(hipMemsetAsync(status.p, 255, c, s));
        HIPCHK(hipMemsetAsync(ctr.p, 0, 2 * CTR_BLOCK * 4, s));
        HIPCHK(hipMemcpyAsync(jobs.p, jp.data(), nsp * sizeof(JobP), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        rng = opt.seed | 1ull;
        fresh = 0;
        free_slots.clear();
        slot_sp.assign(cap, 0);
        slot_sid.assign(cap, 0);
        slot_tone.assign(cap, 0);
        slot_tag.assign(cap, 0);
        slot_birth.assign(cap, 0);
        tally.assign((size_t)o.max_tags * nsp * 8, 0);
        tone_ph.assign(1, 0.0);
        f_live = rf.f_ac;
    }

    void destroy()
    {
        for (DevBuf* b : {&x, &y, &z, &phi, &phr, &vx, &vy, &vz, &fvx, &fvy, &fvz, &steps, &ncoll, &cd, &job, &status,
                          &sid, &jobs, &list[0], &list[1], &ctr, &fin, &nfin, &rec})
            b->release();
    }

    // spawn_ion (engine.rs:372-382): the next free slot, or a drop at capacity;
    // while a sweep runs and does not drain, the ion is on the current point's tone.
    void spawn(uint32_t k, bool reentry)
    {
        uint32_t slot;
        if (fresh < cap) {
            slot = fresh++;
        } else if (!free_slots.empty()) {
            slot = free_slots.back();
            free_slots.pop_back();
        } else {
            st.dropped += 1;
            return;
        }
        const uint32_t s = next_id++;
        const uint32_t tone = (sw.running && !sw.draining) ? sw.point + 1 : 0;
        pending.push_back(LiveNew{slot, tone * nsp + k, s, 0, tone_ph[tone]});
        slot_sp[slot] = k;
        slot_sid[slot] = s;
        slot_tone[slot] = (uint8_t)tone;
        if (tone) tagged_alive += 1;
        slot_tag[slot] = tag;
        slot_birth[slot] = step;
        if (reentry) st.reentered += 1;
        else st.injected += 1;
    }

    // The pending spawns enter the live list with the clock's RF phase and their tones' phases.
    void flush()
    {
        const uint32_t k = (uint32_t)pending.size();
        if (!k) return;
        hipStream_t s = V.stream;
        HIPCHK(hipMemcpyAsync(rec.p, pending.data(), k * sizeof(LiveNew), hipMemcpyHostToDevice, s));
        const double phr_birth = frac(jp[0].w_rf * (double)step);
        launch_live_inject(S, jobs.as<JobP>(), ip, rec.as<LiveNew>(), k, phr_birth, list[cur].as<uint32_t>(),
                           ctr_of(cur) + CTR_COUNT, s);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));  // `pending` is reused
        st.alive += k;
        pending.clear();
    }

    // Sim::block (engine.rs:448-538).
    void block()
    {
        // the engine may have run other DC configurations since: rebuild its tables for ours if so
        V = engine_view(eng, dc, ph, (uint32_t)jp.size());
        const uint32_t B = o.block_steps;
        const uint64_t alive0 = st.alive;
        hipStream_t s = V.stream;
        uint32_t nf = 0;
        if (alive0) {
            WorkerArgs a{};
            a.S = S;
            a.jobs = jobs.as<JobP>();
            a.B = V.B;
            a.cells = V.cells;
            a.list_in = list[cur].as<uint32_t>();
            a.list_out = list[cur ^ 1].as<uint32_t>();
            a.ctr_in = ctr_of(cur);
            a.ctr_out = ctr_of(cur ^ 1);
            a.chunk = B;
            a.period = (uint32_t)std::max(4, (V.tune.refill_period + 3) & ~3);
            a.nqueues = 1;
            a.coll_batch = (uint32_t)std::max(1, V.tune.coll_batch);
            a.bin_out = nullptr;
            a.lanes = 64;
            const uint32_t blk = (uint32_t)V.tune.block;
            const uint64_t want = ((uint64_t)alive0 + blk - 1) / blk;
            const uint32_t grid = (uint32_t)std::max<uint64_t>(
                1, std::min<uint64_t>(want, (uint64_t)V.cus * V.tune.waves_per_cu * 64 / blk));
            launch_prepare(a, s);
            launch_batched(a, grid, blk, s);
            HIPCHK(hipGetLastError());
            launch_fill_u32(nfin.as<uint32_t>(), 0u, 1, s);
            launch_live_collect(S, list[cur].as<uint32_t>(), ctr_of(cur) + CTR_COUNT, (uint32_t)alive0,
                                fin.as<uint32_t>(), nfin.as<uint32_t>(), s);
            HIPCHK(hipGetLastError());
            uint32_t carried = 0;
            HIPCHK(hipMemcpyAsync(&nf, nfin.p, 4, hipMemcpyDeviceToHost, s));
            HIPCHK(hipMemcpyAsync(&carried, ctr_of(cur ^ 1) + CTR_COUNT, 4, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            if ((uint64_t)carried + nf != alive0)
                throw Err(-5, "internal error: live ions lost in a block (" + std::to_string(alive0) + " in, " +
                                  std::to_string(carried) + " carried, " + std::to_string(nf) + " finished)");
            fin_h.resize((size_t)4 * nf);
            if (nf) {
                HIPCHK(hipMemcpyAsync(fin_h.data(), fin.p, (size_t)nf * 16, hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
            }
            cur ^= 1;
        }
        const uint64_t step0 = step, step_end = step + B;
        // finished ions in the order of their birth (the reference walks its ion vector)
        std::vector<uint32_t> order(nf);
        for (uint32_t k = 0; k < nf; ++k) order[k] = k;
        std::sort(order.begin(), order.end(),
                  [&](uint32_t a, uint32_t b) { return slot_sid[fin_h[4 * a]] < slot_sid[fin_h[4 * b]]; });
        uint64_t done = (alive0 - nf) * (uint64_t)B;
        for (uint32_t k : order) {
            const uint32_t slot = fin_h[4 * k], status = fin_h[4 * k + 1], stp = fin_h[4 * k + 2], nc = fin_h[4 * k + 3];
            const uint64_t age0 = step0 - slot_birth[slot];
            done += (uint64_t)stp - age0;
            const int32_t t = slot_tag[slot];
            if (t >= 0 && t < o.max_tags) {
                uint64_t* c = &tally[((size_t)t * nsp + slot_sp[slot]) * 8];
                c[0] += 1;
                c[(status >= 1 && status <= 5) ? status : 5] += 1;
                c[6] += stp;
                c[7] += nc;
            }
            if (const uint32_t tone = slot_tone[slot]) {
                tagged_alive -= 1;
                // engine.rs:491-496: transmitted, crashed, returned, other
                const int cls = status == MQ_TRANSMITTED ? 0 : status == MQ_CRASHED ? 1 : status == MQ_RETURNED ? 2 : 3;
                if (tone - 1 < sw.f.size()) sw.acc[((size_t)(tone - 1) * nsp + slot_sp[slot]) * 4 + cls] += 1;
            }
            if (status == MQ_RETURNED && o.recirculate) {
                const double u = std::max(xs_u01(rng), 1e-12);
                const uint64_t delay = (uint64_t)std::ceil(-std::log(u) * o.residence_s / ph.dt);
                returns.push({slot_birth[slot] + stp + delay, slot_sp[slot]});
            }
            free_slots.push_back(slot);
        }
        st.finished += nf;
        st.alive -= nf;
        st.ion_steps += done;
        step = step_end;
        for (size_t t = 0; t < tone_ph.size(); ++t) tone_ph[t] = frac(tone_ph[t] + jp[t * nsp].w_ac * (double)B);
        // re-entries due by the end of the block, then fresh ions unless a sweep drains (engine.rs:512-535)
        while (!returns.empty() && returns.top().first <= step_end) {
            const uint32_t k = returns.top().second;
            returns.pop();
            spawn(k, true);
        }
        double f_tone0 = -1.0;  // the live tone's next frequency when a sweep point advances
        if (!sw.draining) {
            const uint64_t nnew = poisson(rng, o.inj_rate * (double)B * ph.dt);
            for (uint64_t i = 0; i < nnew; ++i) {
                const double u = xs_u01(rng);
                uint32_t k = 0;
                while (k + 1 < nsp && !(u <= cdf[k])) ++k;
                spawn(k, false);
                if (sw.running && ++sw.inj_in_point >= sw.cfg.ions_per_point) {
                    sw.inj_in_point = 0;
                    if (++sw.point >= sw.f.size()) {
                        sw.sweeps_done += 1;
                        sw.point = 0;
                        if (sw.sweeps_done >= sw.cfg.sweeps) sw.draining = true;
                    }
                    f_tone0 = sw.f[sw.point] * 1e3;
                }
            }
        }
        flush();
        // the sweep stops once the last ion on a sweep tone has ended
        if (sw.draining && sw.running && tagged_alive == 0) {
            sw.running = false;
            sw.draining = false;
            sw.point = (uint32_t)sw.f.size() - 1;
            f_tone0 = f_live;
        }
        if (f_tone0 >= 0.0) set_tone0(f_tone0);
        st.step = step;
    }

    // A new frequency of the live tone from `step` on: its jobs' w_ac, and the
    // phase offsets of the live ions on it.
    void set_tone0(double f_ac)
    {
        const double w = f_ac * ph.dt;
        for (uint32_t k = 0; k < nsp; ++k) jp[k].w_ac = w;
        hipStream_t s = V.stream;
        HIPCHK(hipMemcpyAsync(jobs.p, jp.data(), nsp * sizeof(JobP), hipMemcpyHostToDevice, s));
        if (st.alive)
            launch_live_rephase(S, list[cur].as<uint32_t>(), ctr_of(cur) + CTR_COUNT, (uint32_t)st.alive, tone_ph[0], w,
                                nsp, s);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
    }

    // apply_params (engine.rs:399): the live tone follows a running sweep.
    void set_ac(double f_ac, int32_t t)
    {
        f_live = f_ac;
        if (!sw.running) set_tone0(f_ac);
        tag = t;
    }

    // start_sweep (engine.rs:420-429): the points and their tones, the live tone
    // on the first point. As the clock's resize there, tone t keeps the phase it
    // had if the previous sweep had it and starts at 0 otherwise.
    void sweep_start(const mq_live_sweep_cfg& c)
    {
        if (sw.running) throw Err(-1, "a sweep is running; stop it and let it drain first");
        if (!std::isfinite(c.f_lo_khz) || !std::isfinite(c.f_hi_khz) || !(c.step_khz > 0.0) || !std::isfinite(c.step_khz))
            throw Err(-1, "sweep frequencies must be finite and step_khz > 0");
        if (c.ions_per_point == 0 || c.sweeps == 0) throw Err(-1, "ions_per_point and sweeps must be >= 1");
        const double span = std::round((c.f_hi_khz - c.f_lo_khz) / c.step_khz);
        const uint32_t n = span > 0.0 ? (uint32_t)std::min(span, (double)(MQ_LIVE_MAX_POINTS - 1)) : 0u;
        if (c.f_lo_khz < 0.0 || c.f_lo_khz + n * c.step_khz < 0.0) throw Err(-1, "sweep frequencies must be >= 0");
        sw = Sweep{};
        sw.cfg = c;
        for (uint32_t i = 0; i <= n; ++i) sw.f.push_back(c.f_lo_khz + i * c.step_khz);
        sw.acc.assign((size_t)sw.f.size() * nsp * 4, 0);
        const size_t ntone = 1 + sw.f.size();
        tone_ph.resize(ntone, 0.0);
        jp.resize(ntone * nsp);
        for (size_t t = 1; t < ntone; ++t)
            for (uint32_t k = 0; k < nsp; ++k) {
                jp[t * nsp + k] = jp[k];
                jp[t * nsp + k].w_ac = sw.f[t - 1] * 1e3 * ph.dt;
            }
        HIPCHK(hipMemcpyAsync(jobs.p, jp.data(), jp.size() * sizeof(JobP), hipMemcpyHostToDevice, V.stream));
        sw.running = true;
        set_tone0(sw.f[0] * 1e3);
    }
};

#define MQ_LIVE_GUARD(...)                 \
    try {                                  \
        __VA_ARGS__;                       \
        return 0;                          \
    } catch (const Err& e) {               \
        mq_set_last_error(e.what());       \
        return e.code;                     \
    } catch (const std::exception& e) {    \
        mq_set_last_error(e.what());       \
        return -9;                         \
    }

extern "C" {

int mq_live_create(mq_engine* eng, const mq_species* sp, const double* weights, size_t nsp, const mq_rfac* rf,
                   const double dc[4], const mq_physics* ph, const mq_live_opts* o, mq_live** out)
{
    mq_live* l = nullptr;
    try {
        if (!out || !rf || !dc || !ph || !o) throw Err(-1, "NULL argument");
        l = new mq_live();
        l->create(eng, sp, weights, nsp, *rf, dc, *ph, *o);
        *out = l;
        return 0;
    } catch (const Err& e) {
        mq_set_last_error(e.what());
        if (l) { l->destroy(); delete l; }
        return e.code;
    } catch (const std::exception& e) {
        mq_set_last_error(e.what());
        if (l) { l->destroy(); delete l; }
        return -9;
    }
}

void mq_live_destroy(mq_live* l)
{
    if (!l) return;
    l->destroy();
    delete l;
}

int mq_live_inject(mq_live* l, uint32_t species, uint32_t n)
{
    MQ_LIVE_GUARD({
        if (!l) throw Err(-1, "NULL argument");
        if (species >= l->nsp) throw Err(-1, "species index out of range");
        for (uint32_t i = 0; i < n; ++i) l->spawn(species, false);
        l->flush();
    })
}

int mq_live_run(mq_live* l, uint32_t nblocks, mq_live_state* out)
{
    MQ_LIVE_GUARD({
        if (!l) throw Err(-1, "NULL argument");
        for (uint32_t b = 0; b < nblocks; ++b) l->block();
        l->st.step = l->step;
        if (out) *out = l->st;
    })
}

int mq_live_set_ac(mq_live* l, double f_ac, int32_t tag)
{
    MQ_LIVE_GUARD({
        if (!l) throw Err(-1, "NULL argument");
        if (!(f_ac >= 0.0)) throw Err(-1, "AC frequency must be >= 0");
        l->set_ac(f_ac, tag);
    })
}

int mq_live_sweep_start(mq_live* l, const mq_live_sweep_cfg* c)
{
    MQ_LIVE_GUARD({
        if (!l || !c) throw Err(-1, "NULL argument");
        l->sweep_start(*c);
    })
}

int mq_live_sweep_stop(mq_live* l)
{
    MQ_LIVE_GUARD({
        if (!l) throw Err(-1, "NULL argument");
        // stop_sweep (engine.rs:431-432) also marks a finished sweep running and
        // draining: the next block ends it again, without fresh ions
        if (!l->sw.f.empty()) l->sw.draining = l->sw.running = true;
    })
}

int mq_live_sweep(const mq_live* l, mq_live_sweep_state* st, double* f_khz, uint64_t* acc)
{
    MQ_LIVE_GUARD({
        if (!l) throw Err(-1, "NULL argument");
        const auto& w = l->sw;
        const uint32_t np = (uint32_t)w.f.size();
        if (st) {
            mq_live_sweep_state r{};
            r.npoints = np;
            r.point = w.point;
            r.sweeps_done = w.sweeps_done;
            r.running = w.running;
            r.draining = w.draining;
            const double tot = (double)std::max<uint64_t>(1, (uint64_t)w.cfg.sweeps * np);
            r.progress = np ? std::min(1.0, ((double)w.sweeps_done * np + w.point) / tot) : 0.0;
            *st = r;
        }
        if (f_khz) std::copy(w.f.begin(), w.f.end(), f_khz);
        if (acc) std::copy(w.acc.begin(), w.acc.end(), acc);
    })
}

int mq_live_tally(const mq_live* l, int32_t tag, mq_outcome* per_species)
{
    MQ_LIVE_GUARD({
        if (!l || !per_species) throw Err(-1, "NULL argument");
        if (tag < 0 || tag >= l->o.max_tags) throw Err(-1, "tag out of range");
        for (uint32_t k = 0; k < l->nsp; ++k) {
            const uint64_t* c = &l->tally[((size_t)tag * l->nsp + k) * 8];
            mq_outcome r{};
            r.n = c[0];
            r.crashed = c[1]; r.returned = c[2]; r.transmitted = c[3]; r.left = c[4]; r.timeout = c[5];
            r.mean_coll = c[0] ? (double)c[7] / (double)c[0] : 0.0;
            r.mean_time = c[0] ? (double)c[6] * l->ph.dt / (double)c[0] : 0.0;
            per_species[k] = r;
        }
    })
}

int mq_live_slots(mq_live* l, const mq_ion_soa* s, uint32_t* sid, uint32_t* species, uint64_t* birth)
{
    MQ_LIVE_GUARD({
        if (!l) throw Err(-1, "NULL argument");
        const size_t n = l->cap;
        hipStream_t st = l->V.stream;
        auto d2h = [&](void* dst, const void* src, size_t bytes) {
            if (dst) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
        };
        if (s) {
            d2h(s->x, l->x.p, n * 8); d2h(s->y, l->y.p, n * 8); d2h(s->z, l->z.p, n * 8);
            d2h(s->steps, l->steps.p, n * 4); d2h(s->ncoll, l->ncoll.p, n * 4); d2h(s->status, l->status.p, n);
            d2h(s->phi_ac0, l->phi.p, n * 8); d2h(s->cd, l->cd.p, n * 4);
            std::vector<float> f(n);
            double* dst[3] = {s->vx, s->vy, s->vz};
            const void* src[3] = {l->fvx.p, l->fvy.p, l->fvz.p};
            for (int d = 0; d < 3; ++d) {
                if (!dst[d]) continue;
                HIPCHK(hipMemcpyAsync(f.data(), src[d], n * 4, hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
                for (size_t i = 0; i < n; ++i) dst[d][i] = (double)f[i];
            }
        }
        HIPCHK(hipStreamSynchronize(st));
        if (sid) std::copy(l->slot_sid.begin(), l->slot_sid.end(), sid);
        if (species) std::copy(l->slot_sp.begin(), l->slot_sp.end(), species);
        if (birth) std::copy(l->slot_birth.begin(), l->slot_birth.end(), birth);
    })
}

}  // extern "C"
