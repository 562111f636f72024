// kernels_tandem.hip -- tandem multi-domain batch: one clock, a splitter,
// nine channel ports and a combiner (include/mq_tandem.h).
//
// Reference: trapsim Model::advance, enter, leave
// (simulation/trapsim/src/lib.rs:176-385). Per step an ion samples the grid
// of its region (cube for splitter/combiner, channel for ports), takes chsim's
// kick-drift step and thermal collision, then is classified in its region;
// a crossing moves it to the next region with the coordinate map of
// enter/leave and counts the crossing in per-port u64 counters (integer
// atomics: exact and order-independent).
//
//  * k_tandem_plain: FP64, one thread per ion, relaunched every nsteps
//    (the plain path; bit-identical to the oracle).
//  * k_tandem_batched: FP32 fields and push with FP64 position and clock, in
//    persistent wavefronts with lane refill and carry, as k_batched; the cell
//    record is reloaded from the new region's table on a crossing.
#include <hip/hip_runtime.h>

#include "fp32_cells.h"
#include "fp64_sample.h"
#include "kernels.h"
#include "mq_math.h"
#include "mq_types.h"

namespace mq {

// a[k] without a dynamically indexed load (which would copy the parameter
// block to scratch): an unrolled select chain.
__device__ __forceinline__ double sel9(const double (&a)[9], uint32_t k)
{
    double r = a[0];
#pragma unroll
    for (uint32_t j = 1; j < 9; ++j) r = (k == j) ? a[j] : r;
    return r;
}

// Classification and region crossing after a step (lib.rs:206-265 with
// enter 326-356 and leave 357-377). `crash` tests the conductor mask at the
// nearest node of the ion's grid (clamped to the grid). Returns the new status
// (0 alive, 1 lost, 2 back, 3 detected).
// mxp, myp: the mouth centre of the ion's port when it is in a port (T.mx[p],
// T.my[p]; the caller keeps them per region), ignored in the cube.
template <class V, class Crash>
__device__ __forceinline__ int tandem_cross(const TandemP& T, const TGrid& cube, const TGrid& port, uint32_t& reg,
                                            uint32_t& last, double& x, double& y, double& z, V& vz, uint32_t job,
                                            unsigned long long* __restrict__ ctr, Crash crash, double mxp, double myp)
{
    if (reg == REG_SPLIT || reg == REG_COMBINE) {
        if (crash(false)) return 1;  // a conductor node of the cube (classify, lib.rs:208-212)
        if (z > T.seam) {
            uint32_t p = 0;
            double best = 0.0, mx = 0.0, my = 0.0;
#pragma unroll
            for (uint32_t k = 0; k < 9; ++k) {
                const double dx = x - T.mx[k], dy = y - T.my[k];
                const double d2 = dx * dx + dy * dy;
                if (k == 0 || d2 < best) { best = d2; p = k; mx = T.mx[k]; my = T.my[k]; }
            }
            if (best < T.ap_r2) {
                unsigned long long* c = ctr + (size_t)p * TC_PORT;
                x = x + (T.ax_p - mx);
                y = y + (T.ay_p - my);
                if (reg == REG_SPLIT) {
                    z = T.z_ret + (z - T.seam);
                    atomicAdd(&c[TC_ENTERED], 1ull);
                    if (last != NO_PORT) {
                        atomicAdd(&c[TC_REVISITS], 1ull);
                        atomicAdd(&c[TC_RETRY + last], 1ull);
                        if (last != p) atomicAdd(&c[TC_FROM_OTHER], 1ull);
                    }
                } else {
                    z = T.z_exit - (z - T.seam);
                    vz = -vz;
                    atomicAdd(&c[TC_REVERSE], 1ull);
                }
                reg = 1u + p;
                return 0;
            }
            // outside every mouth the ion keeps moving above the board, up to the cube's top (lib.rs:226-229)
            if (z > cube.lim2) return 1;
        }
        if (y > cube.lim1) {
            // the entrance face: outside the aperture a wall (lib.rs:231-236); through it, the tube, and
            // l_ext beyond the face the ion is back or detected (lib.rs:237-243)
            const double ex = x - T.ent_x, ez = z - T.ent_z;
            if (!(ex * ex + ez * ez < T.ap_r2)) return 1;
            if (y > cube.lim1 + T.l_ext) return reg == REG_SPLIT ? 2 : 3;
        }
        if (x < 0.0 || y < 0.0 || z < 0.0 || x > cube.lim0) return 1;
        return 0;
    }
    const uint32_t p = reg - 1u;
    unsigned long long* c = ctr + (size_t)p * TC_PORT;
    const double mx = mxp, my = myp;
    if (x < 0.0 || y < 0.0 || z < 0.0 || x > port.lim0 || y > port.lim1 || z > port.lim2 || crash(true)) {
        atomicAdd(&c[TC_LOST], 1ull);
        return 1;
    }
    if (z < T.z_ret) {
        x = x + (mx - T.ax_p);
        y = y + (my - T.ay_p);
        z = T.seam + (z - T.z_ret);
        last = p;
        reg = REG_SPLIT;
        atomicAdd(&c[TC_RETURNED], 1ull);
    } else if (z > T.z_exit) {
        x = x + (mx - T.ax_p);
        y = y + (my - T.ay_p);
        z = T.seam - (z - T.z_exit);
        vz = -vz;
        last = NO_PORT;
        reg = REG_COMBINE;
        atomicAdd(&c[TC_FORWARDED], 1ull);
        if (job < (uint32_t)TC_MAX_JOBS) atomicAdd(&c[TC_BYJOB + job], 1ull);
    }
    return 0;
}

// Excitation of a region in cycles per step (0 for the cube and idle ports).
__device__ __forceinline__ double w_ac_of(const TandemP& T, uint32_t reg)
{
    return (reg == REG_SPLIT || reg == REG_COMBINE) ? 0.0 : sel9(T.w_ac, reg - 1u);
}

__device__ __forceinline__ bool port_driven(const TandemP& T, uint32_t reg) { return w_ac_of(T, reg) > 0.0; }

// --------------------------------------------------------------- FP64 plain
struct TPlainArgs {
    IonState S;
    const JobP* jobs;
    TandemP T;
    TGrid cube, port;
    const double* dc[4];  // combined DC: splitter, combiner, driven port, idle port
    // with REC: per-cell FP64 records of the same four field sets (k_build_plain_recs; the cube's
    // records carry an unused AC slot)
    const double2* rec[4];
    unsigned long long* ctr;
    uint32_t n, nsteps;
    uint32_t* live;
};

// REC: the field comes from the per-cell records of the ion's field set, kept
// per lane while the ion stays in one cell of one set; otherwise from the node
// tables. The lerp chain is the same, so the results are too.
template <bool REC>
__global__ __launch_bounds__(256) void k_tandem_plain(TPlainArgs A)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const IonState& S = A.S;
    const TandemP& T = A.T;
    bool alive = false;
    if (i < A.n && S.status[i] == 0) {
        const uint32_t jk = S.job[i];
        const JobP J = A.jobs[jk];
        const uint32_t ion = (uint32_t)(i - J.off);
        double x = S.x[i], y = S.y[i], z = S.z[i], vx = S.vx[i], vy = S.vy[i], vz = S.vz[i];
        uint32_t st = S.steps[i], ncoll = S.ncoll[i];
        uint32_t reg = S.region[i], last = S.last[i];
        const double dt = T.dt;
        int status = 0;
        u32x4 cw = {0, 0, 0, 0};
        uint32_t kc = 0xFFFFFFFFu;  // 4 x cell + field set of the record in kv
        double kv[REC ? 72 : 1];
        for (uint32_t s = 0; s < A.nsteps; ++s) {
            const bool cube = reg == REG_SPLIT || reg == REG_COMBINE;
            const TGrid& G = cube ? A.cube : A.port;
            const double* dcp = reg == REG_SPLIT ? A.dc[0] : reg == REG_COMBINE ? A.dc[1]
                              : port_driven(T, reg) ? A.dc[2] : A.dc[3];
            int ii, jj, kk;
            double tx, ty, tz;
            loc_d(x * G.inv_h, G.nx, ii, tx, (int)T.nearest);
            loc_d(y * G.inv_h, G.ny, jj, ty, (int)T.nearest);
            loc_d(z * G.inv_h, G.nz, kk, tz, (int)T.nearest);
            double er[3], ed[3], ea[3] = {0.0, 0.0, 0.0};
            if constexpr (REC) {
                const uint32_t set = reg == REG_SPLIT ? 0u : reg == REG_COMBINE ? 1u : port_driven(T, reg) ? 2u : 3u;
                const uint32_t cell =
                    (uint32_t)ii + (uint32_t)(G.nx - 1) * ((uint32_t)jj + (uint32_t)(G.ny - 1) * (uint32_t)kk);
                if (4u * cell + set != kc) {
                    kc = 4u * cell + set;
                    const double2* p = (set == 0u ? A.rec[0] : set == 1u ? A.rec[1] : set == 2u ? A.rec[2] : A.rec[3]) +
                                       (size_t)cell * PLAIN_REC_D2;
#pragma unroll
                    for (int q = 0; q < PLAIN_REC_D2; ++q) {
                        const double2 t = p[q];
                        kv[2 * q] = t.x;
                        kv[2 * q + 1] = t.y;
                    }
                }
                double eac[3];
                sample_rec(kv, tx, ty, tz, er, ed, eac);
                if (!cube) { ea[0] = eac[0]; ea[1] = eac[1]; ea[2] = eac[2]; }
            } else {
                const uint32_t sy = (uint32_t)G.nx, sz = (uint32_t)G.nx * (uint32_t)G.ny;
                const uint32_t b = (uint32_t)ii + sy * ((uint32_t)jj + (uint32_t)G.ny * (uint32_t)kk);
                sample_d(G.rf, b, sy, sz, tx, ty, tz, er);
                sample_d(dcp, b, sy, sz, tx, ty, tz, ed);
                if (!cube) sample_d(G.ac, b, sy, sz, tx, ty, tz, ea);
            }
            const double a_rf = T.v_rf * sin2pi(frac_d(T.w_rf * (double)st));
            double a_ac = 0.0;
            if (!cube) {
                const double wa = w_ac_of(T, reg);
                if (wa > 0.0) a_ac = T.v_ac * sin2pi(frac_d(wa * (double)st));
            }
            double E;
            E = (a_rf * er[0] + ed[0]) + a_ac * ea[0]; vx = fma(J.kq, E, vx); x = fma(vx, dt, x);
            E = (a_rf * er[1] + ed[1]) + a_ac * ea[1]; vy = fma(J.kq, E, vy); y = fma(vy, dt, y);
            E = (a_rf * er[2] + ed[2]) + a_ac * ea[2]; vz = fma(J.kq, E, vz); z = fma(vz, dt, z);
            const uint32_t cur = st;
            st += 1;
            // thermal bath (channel.metal:93-99 in the form of k_plain)
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
            // conductor bit of the nearest node; the index is clamped to the grid, since a cube ion may be
            // past its y face inside the entrance tube
            auto crash = [&](bool in_port) {
                const TGrid& H = in_port ? A.port : A.cube;
                const uint32_t i0 = (uint32_t)fmin(fmax(rint(x * H.inv_h), 0.0), (double)(H.nx - 1));
                const uint32_t i1 = (uint32_t)fmin(fmax(rint(y * H.inv_h), 0.0), (double)(H.ny - 1));
                const uint32_t i2 = (uint32_t)fmin(fmax(rint(z * H.inv_h), 0.0), (double)(H.nz - 1));
                return H.mask[i0 + (uint32_t)H.nx * (i1 + (uint32_t)H.ny * i2)] != 0;
            };
            const bool in_port = reg != REG_SPLIT && reg != REG_COMBINE;
            status = tandem_cross(T, A.cube, A.port, reg, last, x, y, z, vz, jk, A.ctr, crash,
                                  in_port ? sel9(T.mx, reg - 1u) : 0.0, in_port ? sel9(T.my, reg - 1u) : 0.0);
            if (status == 0 && st >= T.max_steps) status = 5;
            if (status) break;
        }
        S.x[i] = x; S.y[i] = y; S.z[i] = z; S.vx[i] = vx; S.vy[i] = vy; S.vz[i] = vz;
        S.steps[i] = st; S.ncoll[i] = ncoll; S.status[i] = (uint8_t)status;
        S.region[i] = (uint8_t)reg; S.last[i] = (uint8_t)last;
        alive = status == 0;
    }
    const uint64_t m = __ballot(alive);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(A.live, (uint32_t)__popcll(m));
}

// ------------------------------------------------------------- FP32 batched
struct TLane {
    double x, y, z;
    float vx, vy, vz;
    float kq, nsig, sgas, vb, mr;
    uint32_t st, ncoll, k0, k1, ion, id, job, visit, reg, last;
    float tx, ty, tz, fx, fy, fz, a_rf, a_ac;
    uint32_t cw1, cw2, cw3;
    int32_t cell, tab;
    // constants of the current region, set on entry: cell table, mouth centre and
    // excitation (cycles per step; 0 in the cube and idle ports)
    int32_t rtab;
    double mx, my, wac;
};

__device__ __forceinline__ void set_region(const TandemP& T, TLane& L)
{
    const bool in_port = L.reg != REG_SPLIT && L.reg != REG_COMBINE;
    L.wac = w_ac_of(T, L.reg);
    L.rtab = L.reg == REG_SPLIT ? 0 : L.reg == REG_COMBINE ? 1 : (L.wac > 0.0 ? 2 : 3);
    L.mx = in_port ? sel9(T.mx, L.reg - 1u) : 0.0;
    L.my = in_port ? sel9(T.my, L.reg - 1u) : 0.0;
}

struct TBatchArgs {
    IonState S;
    const JobP* jobs;
    TandemP T;
    TGrid cube, port;
    const float4* cells[4];  // splitter, combiner, driven port, idle port
    unsigned long long* ctr;
    const uint32_t* list_in;
    uint32_t* list_out;
    uint32_t* ctr_in;
    uint32_t* ctr_out;
    uint32_t chunk, period;
    uint32_t nqueues;  // 1, or MAX_QUEUES contiguous segments of a re-sorted list (one per XCD)
};

__device__ __forceinline__ int32_t tab_of(const TandemP& T, uint32_t reg)
{
    return reg == REG_SPLIT ? 0 : reg == REG_COMBINE ? 1 : (port_driven(T, reg) ? 2 : 3);
}

__device__ __forceinline__ int32_t tlocate(const TGrid& cube, const TGrid& port, uint32_t nearest, TLane& L)
{
    const bool in_cube = L.reg == REG_SPLIT || L.reg == REG_COMBINE;
    const double inv_h = in_cube ? cube.inv_h : port.inv_h;
    const int nx = in_cube ? cube.nx : port.nx, ny = in_cube ? cube.ny : port.ny, nz = in_cube ? cube.nz : port.nz;
    const double gx = fmin(fmax(L.x * inv_h, 0.0), (double)(nx - 1));
    const double gy = fmin(fmax(L.y * inv_h, 0.0), (double)(ny - 1));
    const double gz = fmin(fmax(L.z * inv_h, 0.0), (double)(nz - 1));
    const int i = min((int)gx, nx - 2), j = min((int)gy, ny - 2), k = min((int)gz, nz - 2);
    L.fx = (float)(gx - (double)i);
    L.fy = (float)(gy - (double)j);
    L.fz = (float)(gz - (double)k);
    L.tx = nearest ? 0.0f : L.fx;
    L.ty = nearest ? 0.0f : L.fy;
    L.tz = nearest ? 0.0f : L.fz;
    return i + (nx - 1) * (j + (ny - 1) * k);
}

// The four cell tables: splitter, combiner, driven port, idle port.
struct Tables {
    const float4 *c0, *c1, *c2, *c3;
};

// Locate in the lane's region, reload the record on a cell or table change.
__device__ __forceinline__ void tfetch(const TandemP& T, const TGrid& cube, const TGrid& port, const Tables& C,
                                       TLane& L, float4 R[CELL_F4])
{
    const int32_t t = L.rtab;
    const int32_t c = tlocate(cube, port, T.nearest, L);
    if (c != L.cell || t != L.tab) {
        L.cell = c;
        L.tab = t;
        load_cell(t == 0 ? C.c0 : t == 1 ? C.c1 : t == 2 ? C.c2 : C.c3, c, R);
    }
}

__device__ __forceinline__ void tphases(const TandemP& T, TLane& L)
{
    double pr = T.w_rf * (double)L.st;
    pr = pr - floor(pr);
    L.a_rf = (float)T.v_rf * sin_rev((float)pr);
    L.a_ac = 0.0f;
    const double wa = L.wac;
    if (wa > 0.0) {
        double pa = wa * (double)L.st;
        pa = pa - floor(pa);
        L.a_ac = (float)T.v_ac * sin_rev((float)pa);
    }
}

__global__ __launch_bounds__(256) void k_tandem_batched(TBatchArgs A)
{
    const IonState& S = A.S;
    const TandemP& T = A.T;
    const TGrid& cube = A.cube;
    const TGrid& port = A.port;
    const Tables C = {A.cells[0], A.cells[1], A.cells[2], A.cells[3]};
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t count = A.ctr_in[CTR_COUNT];
    const uint32_t nq = A.nqueues;
    uint32_t q = nq > 1 ? blockIdx.x % nq : 0, tried = 0;
    const double dt = T.dt;
    TLane L;
    float4 R[CELL_F4];
    bool have = false, drained = false;
    L.cell = -1;
    L.tab = -1;
    for (;;) {
        // refill idle lanes: with several queues the (re-sorted) list is cut into
        // contiguous segments and the wavefronts of one XCD start on their own
        // segment, then steal, as in k_batched
        while (!drained) {
            const uint64_t idle = __ballot(!have);
            if (!idle) break;
            const uint32_t nidle = (uint32_t)__popcll(idle);
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)idle) - 1;
            const uint32_t s0 = (uint32_t)((uint64_t)count * q / nq), s1 = (uint32_t)((uint64_t)count * (q + 1) / nq);
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&A.ctr_in[q * CTR_Q], nidle);
            base = __shfl(base, (int)leader);
            const uint32_t avail = s1 - s0;
            if (base >= avail) {
                // queue q is used up for this launch: move on. A wavefront with no
                // ion left claims from the next queue at once; one that still
                // steps ions tries it at its next refill.
                if (++tried >= nq) drained = true;
                else q = (q + 1 == nq) ? 0 : q + 1;
                if (__any(have)) break;
                continue;
            }
            if ((idle >> lane) & 1ull) {  // exactly the lanes counted in the claim
                const uint32_t slot = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                if (slot < avail) {
                    const uint32_t id = A.list_in[s0 + slot];
                    const uint32_t jk = S.job[id];
                    const JobP& J = A.jobs[jk];
                    L.id = id;
                    L.job = jk;
                    L.ion = (uint32_t)(id - J.off);
                    L.x = S.x[id]; L.y = S.y[id]; L.z = S.z[id];
                    L.vx = S.fvx[id]; L.vy = S.fvy[id]; L.vz = S.fvz[id];
                    L.st = S.steps[id]; L.ncoll = S.ncoll[id];
                    L.reg = S.region[id]; L.last = S.last[id];
                    L.kq = (float)J.kq; L.nsig = (float)J.nsig_dt; L.sgas = (float)J.s_gas;
                    L.vb = (float)J.v_bound; L.mr = (float)J.mr;
                    L.k0 = J.key0; L.k1 = J.key1;
                    L.visit = 0;
                    L.cw1 = L.cw2 = L.cw3 = 0;
                    set_region(T, L);
                    tfetch(T, cube, port, C, L, R);
                    tphases(T, L);
                    have = true;
                }
            }
            break;
        }
        if (!__any(have)) break;
        bool stop = false;
        int status = 0;
        for (uint32_t r = 0; r < A.period; ++r) {
            if (have && !stop) {
                const float pxy = L.tx * L.ty, pyz = L.ty * L.tz, pxz = L.tx * L.tz, pxyz = pxy * L.tz;
                float e[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const float er = poly(R, 0, d, L.tx, L.ty, L.tz, pxy, pyz, pxz, pxyz);
                    const float ed = poly(R, 1, d, L.tx, L.ty, L.tz, pxy, pyz, pxz, pxyz);
                    const float ea = poly(R, 2, d, L.tx, L.ty, L.tz, pxy, pyz, pxz, pxyz);
                    e[d] = fmaf(L.a_ac, ea, fmaf(L.a_rf, er, ed));
                }
                L.vx = fmaf(L.kq, e[0], L.vx);
                L.vy = fmaf(L.kq, e[1], L.vy);
                L.vz = fmaf(L.kq, e[2], L.vz);
                L.x = fma((double)L.vx, dt, L.x);
                L.y = fma((double)L.vy, dt, L.y);
                L.z = fma((double)L.vz, dt, L.z);
                const uint32_t cur = L.st;
                L.st += 1;
                tfetch(T, cube, port, C, L, R);  // record at the new position (same region): mask bits and next kick
                if (L.nsig > 0.0f) {
                    const float bound = sqrtf(fmaf(L.vz, L.vz, fmaf(L.vy, L.vy, L.vx * L.vx))) + L.vb;
                    uint32_t cw;
                    const uint32_t ph = cur & 3;
                    if (ph == 0 || L.visit == 0) {
                        const u32x4 w = philox(L.ion, cur >> 2, 0, D_CAND, L.k0, L.k1);
                        cw = pick(w, ph);
                        L.cw1 = w.b; L.cw2 = w.c; L.cw3 = w.d;
                    } else {
                        cw = ph == 1 ? L.cw1 : (ph == 2 ? L.cw2 : L.cw3);
                    }
                    if (u01f(cw) < L.nsig * bound) {
                        const u32x4 w = philox(L.ion, cur, 0, D_COLL, L.k0, L.k1);
                        const float r1 = sqrtf(-2.0f * log_f(u01f(w.a))), r2 = sqrtf(-2.0f * log_f(u01f(w.c)));
                        const float a1 = u01f(w.b), a2 = u01f(w.d);
                        const float gx = L.vx - L.sgas * (r1 * cos_rev(a1));
                        const float gy = L.vy - L.sgas * (r1 * sin_rev(a1));
                        const float gz = L.vz - L.sgas * (r2 * cos_rev(a2));
                        const float gl = sqrtf(fmaf(gz, gz, fmaf(gy, gy, gx * gx)));
                        const u32x4 w2 = philox(L.ion, cur, 1, D_COLL, L.k0, L.k1);
                        if (u01f(w2.a) * bound < gl) {
                            const float cth = 2.0f * u01f(w2.b) - 1.0f, a3 = u01f(w2.c);
                            const float sn = sqrtf(fmaxf(0.0f, 1.0f - cth * cth));
                            const float mg = L.mr * gl;
                            L.vx = fmaf(mg, sn * cos_rev(a3), fmaf(-L.mr, gx, L.vx));
                            L.vy = fmaf(mg, sn * sin_rev(a3), fmaf(-L.mr, gy, L.vy));
                            L.vz = fmaf(mg, cth, fmaf(-L.mr, gz, L.vz));
                            L.ncoll += 1;
                        }
                    }
                }
                L.visit += 1;
                const uint32_t corner = (L.fx >= 0.5f ? 1u : 0u) | (L.fy >= 0.5f ? 2u : 0u) | (L.fz >= 0.5f ? 4u : 0u);
                const bool on_conductor = ((__float_as_uint(R[18].x) >> corner) & 1u) != 0u;
                auto crash = [on_conductor](bool) { return on_conductor; };
                const uint32_t reg0 = L.reg;
                status = tandem_cross(T, cube, port, L.reg, L.last, L.x, L.y, L.z, L.vz, L.job, A.ctr, crash, L.mx, L.my);
                if (status == 0 && L.reg != reg0) {
                    set_region(T, L);
                    tfetch(T, cube, port, C, L, R);
                }
                if (status == 0) tphases(T, L);
                if (status == 0 && L.st >= T.max_steps) status = 5;
                if (status != 0 || L.visit >= A.chunk) stop = true;
            }
        }
        const bool carry = have && stop && status == 0;
        const uint64_t cm = __ballot(carry);
        if (cm) {
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)cm) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&A.ctr_out[CTR_COUNT], (uint32_t)__popcll(cm));
            base = __shfl(base, (int)leader);
            if (carry) A.list_out[base + (uint32_t)__popcll(cm & ((1ull << lane) - 1ull))] = L.id;
        }
        if (have && stop) {
            const uint32_t id = L.id;
            S.x[id] = L.x; S.y[id] = L.y; S.z[id] = L.z;
            S.fvx[id] = L.vx; S.fvy[id] = L.vy; S.fvz[id] = L.vz;
            S.steps[id] = L.st; S.ncoll[id] = L.ncoll;
            S.status[id] = (uint8_t)status;
            S.region[id] = (uint8_t)L.reg; S.last[id] = (uint8_t)L.last;
            have = false;
        }
    }
}

// Region and status of every ion at the start, and the final totals.
__global__ void k_tandem_init(IonState S, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    S.region[i] = (uint8_t)REG_SPLIT;
    S.last[i] = (uint8_t)NO_PORT;
    S.status[i] = 0;
}

// End totals: each 256-thread block counts its ions per class (ballots) and
// sums their steps (wavefront sums), then adds the block's eight figures once,
// instead of three atomics per ion on the same eight counters.
__global__ __launch_bounds__(256) void k_tandem_totals(IonState S, uint32_t n, unsigned long long* __restrict__ tot)
{
    __shared__ unsigned long long part[4][TT_N];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const bool act = i < n;
    int k = 0;
    unsigned long long sv = 0;
    if (act) {
        const uint32_t st = S.status[i], reg = S.region[i];
        if (st == 1) k = 4;
        else if (st == 2) k = 5;
        else if (st == 3) k = 6;
        else if (st == 5) k = reg == REG_SPLIT ? 1 : reg == REG_COMBINE ? 3 : 2;
        else k = TT_N;  // never finished: reported by the host as an error
        sv = S.steps[i];
    }
    unsigned long long c[TT_N];
    c[0] = (unsigned long long)__popcll(__ballot(act));
#pragma unroll
    for (int q = 1; q <= 6; ++q) c[q] = (unsigned long long)__popcll(__ballot(act && k == q));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sv += __shfl_xor(sv, o);
    c[7] = sv;
    const unsigned long long unfinished = (unsigned long long)__popcll(__ballot(act && k == TT_N));
    if (lane == 0 && unfinished) atomicAdd(&tot[TT_N], unfinished);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < TT_N; ++q) part[w][q] = c[q];
    __syncthreads();
    if (threadIdx.x < (uint32_t)TT_N) {
        const uint32_t q = threadIdx.x;
        const unsigned long long v = part[0][q] + part[1][q] + part[2][q] + part[3][q];
        if (v) atomicAdd(&tot[q], v);
    }
}

// ------------------------------------------------------------- launchers
void launch_tandem_init(const IonState& S, uint32_t n, hipStream_t st)
{
    if (n) k_tandem_init<<<(n + 255) / 256, 256, 0, st>>>(S, n);
}

void launch_tandem_plain(const IonState& S, const JobP* jobs, const TandemP& T, const TGrid& cube, const TGrid& port,
                         const double* const dc[4], const double2* const rec[4], unsigned long long* ctr, uint32_t n,
                         uint32_t nsteps, uint32_t* live, hipStream_t st)
{
    TPlainArgs a;
    a.S = S; a.jobs = jobs; a.T = T; a.cube = cube; a.port = port;
    for (int k = 0; k < 4; ++k) {
        a.dc[k] = dc[k];
        a.rec[k] = rec ? rec[k] : nullptr;
    }
    a.ctr = ctr; a.n = n; a.nsteps = nsteps; a.live = live;
    if (!n) return;
    if (rec) k_tandem_plain<true><<<(n + 255) / 256, 256, 0, st>>>(a);
    else k_tandem_plain<false><<<(n + 255) / 256, 256, 0, st>>>(a);
}

void launch_tandem_batched(const IonState& S, const JobP* jobs, const TandemP& T, const TGrid& cube,
                           const TGrid& port, const float4* const cells[4], unsigned long long* ctr,
                           const uint32_t* list_in, uint32_t* list_out, uint32_t* ctr_in, uint32_t* ctr_out,
                           uint32_t chunk, uint32_t period, uint32_t nqueues, uint32_t grid, uint32_t block,
                           hipStream_t st)
{
    TBatchArgs a;
    a.S = S; a.jobs = jobs; a.T = T; a.cube = cube; a.port = port;
    for (int k = 0; k < 4; ++k) a.cells[k] = cells[k];
    a.ctr = ctr; a.list_in = list_in; a.list_out = list_out; a.ctr_in = ctr_in; a.ctr_out = ctr_out;
    a.chunk = chunk; a.period = period; a.nqueues = nqueues;
    k_tandem_batched<<<grid, block, 0, st>>>(a);
}

// Re-sort key of a live tandem ion: its cell table (splitter, combiner, driven
// port, idle port: ions of different ports that share a table and a cell read
// the same record) and the 4 x 4 x 4-cell bin of its cell in that table's grid.
__global__ void k_tandem_bin(IonState S, TandemP T, TGrid cube, TGrid port, const uint32_t* __restrict__ list,
                             const uint32_t* __restrict__ ctr, uint32_t* __restrict__ bin_of,
                             uint32_t* __restrict__ rank_of, uint32_t* __restrict__ hist, uint32_t nbc, uint32_t nbp)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ctr[CTR_COUNT]) return;
    const uint32_t id = list[i];
    const uint32_t reg = S.region[id];
    const int32_t t = tab_of(T, reg);
    const TGrid& g = t < 2 ? cube : port;
    const int ci = min((int)fmin(fmax(S.x[id] * g.inv_h, 0.0), (double)(g.nx - 1)), g.nx - 2);
    const int cj = min((int)fmin(fmax(S.y[id] * g.inv_h, 0.0), (double)(g.ny - 1)), g.ny - 2);
    const int ck = min((int)fmin(fmax(S.z[id] * g.inv_h, 0.0), (double)(g.nz - 1)), g.nz - 2);
    const uint32_t nbx = (uint32_t)((g.nx - 2) >> 2) + 1, nby = (uint32_t)((g.ny - 2) >> 2) + 1;
    const uint32_t off = t == 0 ? 0u : t == 1 ? nbc : t == 2 ? 2u * nbc : 2u * nbc + nbp;
    const uint32_t bin = off + (((uint32_t)ck >> 2) * nby + ((uint32_t)cj >> 2)) * nbx + ((uint32_t)ci >> 2);
    bin_of[i] = bin;
    bin_rank(hist, bin, i, rank_of);
}

void launch_tandem_bin(const IonState& S, const TandemP& T, const TGrid& cube, const TGrid& port, const uint32_t* list,
                       const uint32_t* ctr, uint32_t* bin_of, uint32_t* rank_of, uint32_t* hist, uint32_t nbc,
                       uint32_t nbp, uint32_t nmax, hipStream_t st)
{
    if (nmax) k_tandem_bin<<<(nmax + 255) / 256, 256, 0, st>>>(S, T, cube, port, list, ctr, bin_of, rank_of, hist, nbc, nbp);
}

void launch_tandem_totals(const IonState& S, uint32_t n, unsigned long long* tot, hipStream_t st)
{
    if (n) k_tandem_totals<<<(n + 255) / 256, 256, 0, st>>>(S, n, tot);
}

}  // namespace mq
