// kernels_plain.hip -- FP64 "plain" path and the kernels both paths share:
// injection, DC combination, summarize.
//
// The plain kernel is the FP64 form of chsim's direct per-ion kernel
// (simulation/chsim/src/channel.metal:70-124, launched by Engine::run_direct,
// channel.rs:478-493): one thread per ion, node-major FP64 basis grids read from
// global memory every step, relaunched for nsteps until every ion has finished.
// Built with -ffp-contract=off; fused multiply-adds are explicit fma().
#include <hip/hip_runtime.h>

#include "fp64_sample.h"
#include "kernels.h"
#include "mq_math.h"
#include "mq_types.h"

namespace mq {

constexpr double QE = 1.602176634e-19;
constexpr double KB = 1.380649e-23;

__device__ __forceinline__ uint32_t job_of(const JobP* jobs, uint32_t njobs, uint32_t i)
{
    uint32_t lo = 0, hi = njobs - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (jobs[mid].off <= (uint64_t)i) lo = mid; else hi = mid - 1;
    }
    return lo;
}


// ---------------------------------------------------------------- inject
// Engine::inject (channel.rs:342-374) with the Philox stream: three blocks
// (ion, 0, k, D_INJ), k = 0..2, give up to 12 uniforms.
// The initial state of ion `ion` of job J, written to slot i of S.
__device__ __forceinline__ void inject_one(const IonState& S, const JobP& J, const InjP& P, uint32_t ion, uint32_t i,
                                           uint32_t j)
{
    u32x4 w0 = philox(ion, 0, 0, D_INJ, J.key0, J.key1);
    u32x4 w1 = philox(ion, 0, 1, D_INJ, J.key0, J.key1);
    u32x4 w2 = philox(ion, 0, 2, D_INJ, J.key0, J.key1);
    const double u[12] = {u01(w0.a), u01(w0.b), u01(w0.c), u01(w0.d), u01(w1.a), u01(w1.b),
                          u01(w1.c), u01(w1.d), u01(w2.a), u01(w2.b), u01(w2.c), u01(w2.d)};
    const double m = J.m, ax = P.axis_x, ay = P.axis_y, z0 = P.z_inject;
    double x, y, z, vx, vy, vz, phi, ufirst, sn, cs;
    if (P.model == 2) {
        double sig = P.inj_sigma_mm * 1e-3, s = sqrt(KB * P.temp_k / m);
        double r1 = sqrt(-2.0 * log_d(u[0]));
        sincos2pi(u[1], sn, cs);
        x = ax + sig * (r1 * cs);
        y = ay + sig * (r1 * sn);
        z = z0;
        double r2 = sqrt(-2.0 * log_d(u[2]));
        sincos2pi(u[3], sn, cs);
        vx = s * (r2 * cs);
        vy = s * (r2 * sn);
        double r3 = sqrt(-2.0 * log_d(u[4]));
        sincos2pi(u[5], sn, cs);
        vz = s * (r3 * cs);
        phi = u[6];
        ufirst = u[7];
    } else if (P.model == 1) {
        // channel.rs:350-355: KE = 0.1 z (0.75 + 0.75 u) eV in a +-10 degree cone on axis
        double kin = 0.1 * (0.75 + 0.75 * u[0]) * J.zc;
        double vin = sqrt(2.0 * kin * QE / m);
        sincos2pi((-10.0 + 20.0 * u[1]) / 360.0, sn, cs);
        double vxy = 0.707 * vin * sn;
        x = ax; y = ay; z = z0;
        vx = vxy; vy = vxy; vz = vin * cs;
        phi = u[2];
        ufirst = u[3];
    } else {
        double s = sqrt(KB * P.temp_k / m);
        double r = P.r_inject_mm * 1e-3 * sqrt(u[0]);
        sincos2pi(u[1], sn, cs);
        vz = s * sqrt(-2.0 * log_d(u[2]));
        x = ax + r * cs;
        y = ay + r * sn;
        z = z0;
        double c3, s3;
        sincos2pi(u[4], s3, c3);
        vx = s * (sqrt(-2.0 * log_d(u[3])) * c3);
        sincos2pi(u[6], s3, c3);
        vy = s * (sqrt(-2.0 * log_d(u[5])) * c3);
        phi = u[7];
        ufirst = u[8];
    }
    double vmag = norm3(vx, vy, vz);
    double first_mfp = -log_d(fmax(ufirst, 1e-7)) * J.mfp_mean / 10.0;
    S.x[i] = x; S.y[i] = y; S.z[i] = z; S.phi[i] = phi;
    S.vx[i] = vx; S.vy[i] = vy; S.vz[i] = vz;
    S.fvx[i] = (float)vx; S.fvy[i] = (float)vy; S.fvz[i] = (float)vz;
    S.steps[i] = 0; S.ncoll[i] = 0;
    S.cd[i] = countdown(floor(first_mfp / vmag / P.dt) + 1.0);
    S.job[i] = j;
    S.status[i] = 0;
}

__global__ void k_inject(IonState S, const JobP* __restrict__ jobs, uint32_t njobs, InjP P, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = job_of(jobs, njobs, i);
    const JobP J = jobs[j];
    inject_one(S, J, P, (uint32_t)(i - J.off), i, j);
}

// ------------------------------------------------------- live population
// New ions of a live population (the injection of mqviz's Sim::block,
// engine.rs:512-535): ion k goes to slot rec[k].slot with the initial state of
// ion rec[k].sid of job rec[k].job (a species on one AC tone), the RF phase of
// the running clock at its birth and its tone's phase rec[k].pha; its slot is
// appended to the live list.
__global__ void k_live_inject(IonState S, const JobP* __restrict__ jobs, InjP P, const LiveNew* __restrict__ rec,
                              uint32_t k, double phr, uint32_t* __restrict__ list, uint32_t* __restrict__ count)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= k) return;
    const LiveNew r = rec[t];
    inject_one(S, jobs[r.job], P, r.sid, r.slot, r.job);
    S.phr[r.slot] = phr;
    S.phi[r.slot] = r.pha;
    S.sid[r.slot] = r.sid;
    list[count[0] + t] = r.slot;
}

// After the block's worker launch: the count of the carried list grows by the
// k new ions (one thread, after k_live_inject).
__global__ void k_live_count(uint32_t* __restrict__ count, uint32_t k) { count[0] += k; }

// The ions of the block's list that finished during it: {slot, status, steps,
// collisions} appended to `out` (the host tallies them and recycles their slots).
__global__ void k_live_collect(IonState S, const uint32_t* __restrict__ list, const uint32_t* __restrict__ count,
                               uint32_t* __restrict__ out, uint32_t* __restrict__ nout)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count[0]) return;
    const uint32_t id = list[i];
    const uint8_t st = S.status[id];
    if (st == 0) return;
    const uint32_t k = atomicAdd(nout, 1u);
    out[4 * k] = id;
    out[4 * k + 1] = st;
    out[4 * k + 2] = S.steps[id];
    out[4 * k + 3] = S.ncoll[id];
}

// A new frequency of one AC tone of the running clock (Clock, world.rs:437-465,
// whose phases are continuous): the AC phase offset of every live ion on that
// tone (job below job_end) so that its phase at its current step equals the
// tone's phase `pha` and then advances by w_ac.
__global__ void k_live_rephase(IonState S, const uint32_t* __restrict__ list, const uint32_t* __restrict__ count,
                               double pha, double w_ac, uint32_t job_end)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count[0]) return;
    const uint32_t id = list[i];
    if (S.job[id] >= job_end) return;
    const double p = pha - w_ac * (double)S.steps[id];
    S.phi[id] = p - floor(p);
}

// ------------------------------------------------------- DC combination
// Fields::dc_field (channel.rs:220-226) in FP64: e = sum_b v_b E_b, basis order.
__global__ void k_combine_dc(double* __restrict__ out, const double* d0, const double* d1, const double* d2,
                             const double* d3, double v0, double v1, double v2, double v3, uint32_t n3)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    double e = 0.0;
    if (d0 && v0 != 0.0) e = e + v0 * d0[i];
    if (d1 && v1 != 0.0) e = e + v1 * d1[i];
    if (d2 && v2 != 0.0) e = e + v2 * d2[i];
    if (d3 && v3 != 0.0) e = e + v3 * d3[i];
    out[i] = e;
}

// ---------------------------------------------------- FP64 plain step
// Per-cell FP64 records: the 8 corner node values (corner dx + 2 dy + 4 dz) of
// the RF, combined DC and AC tables, [basis][corner][component], 72 doubles in
// PLAIN_REC_D2 16-byte pieces.
__global__ void k_build_plain_recs(double2* __restrict__ rec, const double* __restrict__ rf,
                                   const double* __restrict__ dc, const double* __restrict__ ac, int nx, int ny, int nz)
{
    const uint32_t cx = (uint32_t)(nx - 1), cy = (uint32_t)(ny - 1), cz = (uint32_t)(nz - 1);
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (uint64_t)cx * cy * cz) return;
    const uint32_t i = (uint32_t)(c % cx), j = (uint32_t)((c / cx) % cy), k = (uint32_t)(c / ((uint64_t)cx * cy));
    const uint32_t sy = (uint32_t)nx, sz = (uint32_t)nx * (uint32_t)ny;
    const uint32_t b = i + sy * (j + (uint32_t)ny * k);
    const uint32_t node[8] = {b, b + 1, b + sy, b + 1 + sy, b + sz, b + 1 + sz, b + sy + sz, b + 1 + sy + sz};
    const double* F[3] = {rf, dc, ac};
    double v[72];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int kk = 0; kk < 8; ++kk)
#pragma unroll
            for (int d = 0; d < 3; ++d) v[24 * q + 3 * kk + d] = F[q][3 * node[kk] + d];
    double2* dst = rec + c * PLAIN_REC_D2;
#pragma unroll
    for (int q = 0; q < PLAIN_REC_D2; ++q) dst[q] = make_double2(v[2 * q], v[2 * q + 1]);
}

// channel.metal:75-123 in FP64: kick with E at the current position, drift,
// advance the step counter, collide, classify (return, exit, box, crash, timeout).
// REC: the field comes from per-cell records (k_build_plain_recs) that a lane
// keeps in registers and reloads, 36 16-byte loads off one address, only when
// its ion has entered another cell; otherwise it is gathered from the node
// tables, 72 separately addressed values per step. The arithmetic is the same.
template <bool REC>
__global__ __launch_bounds__(256) void k_plain(IonState S, const JobP* __restrict__ jobs, BatchP B,
                                               const double* __restrict__ rf, const double* __restrict__ dc,
                                               const double* __restrict__ ac, const double2* __restrict__ rec,
                                               const uint8_t* __restrict__ mask, uint32_t n, uint32_t nsteps,
                                               uint32_t* __restrict__ live)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool alive = false;
    if (i < n && S.status[i] == 0) {
        const JobP J = jobs[S.job[i]];
        const uint32_t ion = (uint32_t)(i - J.off);
        double x = S.x[i], y = S.y[i], z = S.z[i], vx = S.vx[i], vy = S.vy[i], vz = S.vz[i];
        const double phi = S.phi[i];
        uint32_t st = S.steps[i], ncoll = S.ncoll[i], cd = S.cd[i];
        const uint32_t sy = (uint32_t)B.nx, sz = (uint32_t)B.nx * (uint32_t)B.ny;
        const double dt = B.dt;
        int status = 0;
        u32x4 cw = {0, 0, 0, 0};
        uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;  // xoshiro128** state (compat streams)
        if (B.compat) {
            s0 = S.rng[4 * (size_t)i]; s1 = S.rng[4 * (size_t)i + 1];
            s2 = S.rng[4 * (size_t)i + 2]; s3 = S.rng[4 * (size_t)i + 3];
        }
        uint32_t kc = 0xFFFFFFFFu;  // the cell whose record kv holds
        double kv[REC ? 72 : 1];
        for (uint32_t s = 0; s < nsteps; ++s) {
            int ii, jj, kk;
            double tx, ty, tz;
            loc_d(x * B.inv_h, B.nx, ii, tx, B.nearest);
            loc_d(y * B.inv_h, B.ny, jj, ty, B.nearest);
            loc_d(z * B.inv_h, B.nz, kk, tz, B.nearest);
            double er[3], ed[3], ea[3];
            if constexpr (REC) {
                const uint32_t cell =
                    (uint32_t)ii + (uint32_t)(B.nx - 1) * ((uint32_t)jj + (uint32_t)(B.ny - 1) * (uint32_t)kk);
                if (cell != kc) {
                    kc = cell;
                    const double2* p = rec + (size_t)cell * PLAIN_REC_D2;
#pragma unroll
                    for (int q = 0; q < PLAIN_REC_D2; ++q) {
                        const double2 t = p[q];
                        kv[2 * q] = t.x;
                        kv[2 * q + 1] = t.y;
                    }
                }
                sample_rec(kv, tx, ty, tz, er, ed, ea);
            } else {
                const uint32_t b = (uint32_t)ii + sy * ((uint32_t)jj + (uint32_t)B.ny * (uint32_t)kk);
                sample_d(rf, b, sy, sz, tx, ty, tz, er);
                sample_d(dc, b, sy, sz, tx, ty, tz, ed);
                sample_d(ac, b, sy, sz, tx, ty, tz, ea);
            }
            const double a_rf = J.v_rf * sin2pi(frac_d(J.w_rf * (double)st));
            const double a_ac = J.v_ac * sin2pi(frac_d(fma(J.w_ac, (double)st, phi)));
            double E;
            E = (a_rf * er[0] + J.s_dc * ed[0]) + a_ac * ea[0]; vx = fma(J.kq, E, vx); x = fma(vx, dt, x);
            E = (a_rf * er[1] + J.s_dc * ed[1]) + a_ac * ea[1]; vy = fma(J.kq, E, vy); y = fma(vy, dt, y);
            E = (a_rf * er[2] + J.s_dc * ed[2]) + a_ac * ea[2]; vz = fma(J.kq, E, vz); z = fma(vz, dt, z);
            const uint32_t cur = st;
            st += 1;
            if (B.compat) {
                // the reference's per-ion xoshiro128** stream in the draw order of
                // channel.metal:93-113 (rand01, gauss3, iso_dir), in FP64
                if (!B.legacy_coll) {
                    const double bound = norm3(vx, vy, vz) + J.v_bound;
                    if (xo_u01(xo_next(s0, s1, s2, s3)) < J.nsig_dt * bound) {
                        const double u1 = fmax(xo_u01(xo_next(s0, s1, s2, s3)), 1e-7);
                        const double u2 = xo_u01(xo_next(s0, s1, s2, s3));
                        const double u3 = fmax(xo_u01(xo_next(s0, s1, s2, s3)), 1e-7);
                        const double u4 = xo_u01(xo_next(s0, s1, s2, s3));
                        const double r1 = sqrt(-2.0 * log_d(u1)), r2 = sqrt(-2.0 * log_d(u3));
                        double s1n, c1n, s2n, c2n;
                        sincos2pi(u2, s1n, c1n);
                        sincos2pi(u4, s2n, c2n);
                        const double gx = vx - J.s_gas * (r1 * c1n), gy = vy - J.s_gas * (r1 * s1n),
                                     gz = vz - J.s_gas * (r2 * c2n);
                        const double gl = norm3(gx, gy, gz);
                        if (xo_u01(xo_next(s0, s1, s2, s3)) * bound < gl) {
                            const double c = 2.0 * xo_u01(xo_next(s0, s1, s2, s3)) - 1.0;
                            double sp, cp;
                            sincos2pi(xo_u01(xo_next(s0, s1, s2, s3)), sp, cp);
                            const double sn = sqrt(fmax(0.0, 1.0 - c * c));
                            const double mg = J.mr * gl;
                            vx = (vx - J.mr * gx) + mg * (sn * cp);
                            vy = (vy - J.mr * gy) + mg * (sn * sp);
                            vz = (vz - J.mr * gz) + mg * c;
                            ncoll += 1;
                        }
                    }
                } else {
                    cd -= 1;
                    if (cd == 0) {
                        const double vm = norm3(vx, vy, vz);
                        if (vm > J.floor_v) {
                            const double c = 2.0 * xo_u01(xo_next(s0, s1, s2, s3)) - 1.0;
                            double sp, cp;
                            sincos2pi(xo_u01(xo_next(s0, s1, s2, s3)), sp, cp);
                            const double sn = sqrt(fmax(0.0, 1.0 - c * c));
                            const double mv = J.mr * vm;
                            vx = vx * J.one_m_mr + mv * (sn * cp);
                            vy = vy * J.one_m_mr + mv * (sn * sp);
                            vz = vz * J.one_m_mr + mv * c;
                            ncoll += 1;
                        }
                        const double mfp = -log_d(fmax(xo_u01(xo_next(s0, s1, s2, s3)), 1e-7)) * J.mfp_mean;
                        cd = countdown(floor(mfp / fmax(norm3(vx, vy, vz), 1e-3) / dt));
                    }
                }
            } else if (!B.legacy_coll) {
                if (J.nsig_dt > 0.0) {
                    const double bound = norm3(vx, vy, vz) + J.v_bound;
                    if ((cur & 3) == 0 || s == 0) cw = philox(ion, cur >> 2, 0, D_CAND, J.key0, J.key1);
                    if (u01(pick(cw, cur & 3)) < J.nsig_dt * bound) {
                        u32x4 w = philox(ion, cur, 0, D_COLL, J.key0, J.key1);
                        double r1 = sqrt(-2.0 * log_d(u01(w.a))), r2 = sqrt(-2.0 * log_d(u01(w.c)));
                        double s1, c1, s2, c2;
                        sincos2pi(u01(w.b), s1, c1);
                        sincos2pi(u01(w.d), s2, c2);
                        const double gx = vx - J.s_gas * (r1 * c1), gy = vy - J.s_gas * (r1 * s1),
                                     gz = vz - J.s_gas * (r2 * c2);
                        const double gl = norm3(gx, gy, gz);
                        u32x4 w2 = philox(ion, cur, 1, D_COLL, J.key0, J.key1);
                        if (u01(w2.a) * bound < gl) {
                            double c = 2.0 * u01(w2.b) - 1.0, sp, cp;
                            sincos2pi(u01(w2.c), sp, cp);
                            double sn = sqrt(fmax(0.0, 1.0 - c * c));
                            double mg = J.mr * gl;
                            vx = (vx - J.mr * gx) + mg * (sn * cp);
                            vy = (vy - J.mr * gy) + mg * (sn * sp);
                            vz = (vz - J.mr * gz) + mg * c;
                            ncoll += 1;
                        }
                    }
                }
            } else {
                cd -= 1;
                if (cd == 0) {
                    u32x4 w = philox(ion, cur, 0, D_COLL, J.key0, J.key1);
                    double vm = norm3(vx, vy, vz);
                    if (vm > J.floor_v) {
                        double c = 2.0 * u01(w.a) - 1.0, sp, cp;
                        sincos2pi(u01(w.b), sp, cp);
                        double sn = sqrt(fmax(0.0, 1.0 - c * c));
                        double mv = J.mr * vm;
                        vx = vx * J.one_m_mr + mv * (sn * cp);
                        vy = vy * J.one_m_mr + mv * (sn * sp);
                        vz = vz * J.one_m_mr + mv * c;
                        ncoll += 1;
                    }
                    double mfp = -log_d(fmax(u01(w.c), 1e-7)) * J.mfp_mean;
                    cd = countdown(floor(mfp / fmax(norm3(vx, vy, vz), 1e-3) / dt));
                }
            }
            if (z < B.z_ret) status = 2;
            else if (z > B.z_exit) status = 3;
            else if (x < 0.0 || y < 0.0 || z < 0.0 || x > B.lim0 || y > B.lim1 || z > B.lim2) status = 4;
            else if (z > B.cz0 && z < B.cz1 &&
                     mask[(uint32_t)rint(x * B.inv_h) +
                          sy * ((uint32_t)rint(y * B.inv_h) + (uint32_t)B.ny * (uint32_t)rint(z * B.inv_h))])
                status = 1;
            else if (st >= B.max_steps) status = 5;
            if (status) break;
        }
        S.x[i] = x; S.y[i] = y; S.z[i] = z; S.vx[i] = vx; S.vy[i] = vy; S.vz[i] = vz;
        S.steps[i] = st; S.ncoll[i] = ncoll; S.cd[i] = cd; S.status[i] = (uint8_t)status;
        if (B.compat) {
            S.rng[4 * (size_t)i] = s0; S.rng[4 * (size_t)i + 1] = s1;
            S.rng[4 * (size_t)i + 2] = s2; S.rng[4 * (size_t)i + 3] = s3;
        }
        alive = status == 0;
    }
    const uint64_t m = __ballot(alive);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(live, (uint32_t)__popcll(m));
}

// ------------------------------------------------------------ summarize
// summarize (channel.rs:496-506): per-job counts, sum of steps and collisions.
// Per-job counters. The ions of a job are contiguous, so the lanes of a
// wavefront nearly always share one job: each group of lanes with the same job
// (peeled off by ballot) is reduced in the wavefront and added by one lane,
// instead of four atomics per ion on the same few counters. The sums are
// integers, so the order of the additions does not matter.
__global__ void k_summarize(const uint8_t* __restrict__ status, const uint32_t* __restrict__ steps,
                            const uint32_t* __restrict__ ncoll, const uint32_t* __restrict__ job,
                            unsigned long long* __restrict__ acc, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const bool act = i < n;
    uint32_t j = 0, st = 0;
    unsigned long long sv = 0, cv = 0;
    if (act) {
        j = job[i];
        st = status[i];
        st = st >= 1 && st <= 5 ? st : (uint32_t)ACC_UNFINISHED;
        sv = steps[i];
        cv = ncoll[i];
    }
    uint64_t todo = __ballot(act);
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const uint32_t lj = (uint32_t)__shfl((int)j, leader);
        const uint64_t same = __ballot(act && j == lj) & todo;
        const bool in = (same >> lane) & 1ull;
        unsigned long long s6 = in ? sv : 0ull, s7 = in ? cv : 0ull;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            s6 += __shfl_xor(s6, o);
            s7 += __shfl_xor(s7, o);
        }
        uint32_t cls[6];
#pragma unroll
        for (uint32_t c = 1; c <= 5; ++c) cls[c] = (uint32_t)__popcll(__ballot(in && st == c));
        const uint32_t unfinished = (uint32_t)__popcll(__ballot(in && st == (uint32_t)ACC_UNFINISHED));
        if (lane == (uint32_t)leader) {
            unsigned long long* a = acc + (size_t)lj * ACC;
            atomicAdd(&a[0], (unsigned long long)__popcll(same));
#pragma unroll
            for (uint32_t c = 1; c <= 5; ++c)
                if (cls[c]) atomicAdd(&a[c], (unsigned long long)cls[c]);
            if (unfinished) atomicAdd(&a[ACC_UNFINISHED], (unsigned long long)unfinished);
            atomicAdd(&a[6], s6);
            atomicAdd(&a[7], s7);
        }
        todo &= ~same;
    }
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

static inline uint32_t nblk(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

void launch_inject(const IonState& S, const JobP* jobs, uint32_t njobs, const InjP& P, uint32_t n, hipStream_t st)
{
    if (n) k_inject<<<nblk(n, 256), 256, 0, st>>>(S, jobs, njobs, P, n);
}

void launch_live_inject(const IonState& S, const JobP* jobs, const InjP& P, const LiveNew* rec, uint32_t k, double phr,
                        uint32_t* list, uint32_t* count, hipStream_t st)
{
    if (!k) return;
    k_live_inject<<<nblk(k, 256), 256, 0, st>>>(S, jobs, P, rec, k, phr, list, count);
    k_live_count<<<1, 1, 0, st>>>(count, k);
}

void launch_live_collect(const IonState& S, const uint32_t* list, const uint32_t* count, uint32_t nmax, uint32_t* out,
                         uint32_t* nout, hipStream_t st)
{
    if (nmax) k_live_collect<<<nblk(nmax, 256), 256, 0, st>>>(S, list, count, out, nout);
}

void launch_live_rephase(const IonState& S, const uint32_t* list, const uint32_t* count, uint32_t nmax, double pha,
                         double w_ac, uint32_t job_end, hipStream_t st)
{
    if (nmax) k_live_rephase<<<nblk(nmax, 256), 256, 0, st>>>(S, list, count, pha, w_ac, job_end);
}

void launch_combine_dc(double* out, const double* const d[4], const double v[4], uint32_t n3, hipStream_t st)
{
    k_combine_dc<<<nblk(n3, 256), 256, 0, st>>>(out, d[0], d[1], d[2], d[3], v[0], v[1], v[2], v[3], n3);
}

void launch_plain(const IonState& S, const JobP* jobs, const BatchP& B, const double* rf, const double* dc,
                  const double* ac, const double2* rec, const uint8_t* mask, uint32_t n, uint32_t nsteps,
                  uint32_t* live, hipStream_t st)
{
    if (rec) k_plain<true><<<nblk(n, 256), 256, 0, st>>>(S, jobs, B, rf, dc, ac, rec, mask, n, nsteps, live);
    else k_plain<false><<<nblk(n, 256), 256, 0, st>>>(S, jobs, B, rf, dc, ac, rec, mask, n, nsteps, live);
}
void launch_build_plain_recs(double2* rec, const double* rf, const double* dc, const double* ac, int nx, int ny,
                             int nz, hipStream_t st)
{
    const uint64_t nc = (uint64_t)(nx - 1) * (ny - 1) * (nz - 1);
    if (nc) k_build_plain_recs<<<(uint32_t)((nc + 127) / 128), 128, 0, st>>>(rec, rf, dc, ac, nx, ny, nz);
}

void launch_summarize(const IonState& S, unsigned long long* acc, uint32_t n, hipStream_t st)
{
    if (n) k_summarize<<<nblk(n, 256), 256, 0, st>>>(S.status, S.steps, S.ncoll, S.job, acc, n);
}

void launch_fill_u64(unsigned long long* p, unsigned long long v, uint32_t n, hipStream_t st)
{
    if (n) k_fill_u64<<<nblk(n, 256), 256, 0, st>>>(p, v, n);
}

void launch_fill_u32(uint32_t* p, uint32_t v, uint32_t n, hipStream_t st)
{
    if (n) k_fill_u32<<<nblk(n, 256), 256, 0, st>>>(p, v, n);
}

}  // namespace mq
