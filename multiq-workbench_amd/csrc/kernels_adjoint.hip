// kernels_adjoint.hip -- the differentiable form: gradients of a smoothed
// transmission / leakage objective through the FP64 step, collision-free or
// with the thermal collisions of the Philox streams held at their draws.
//
// The reference has no gradients (its Optimize search is a grid of sweeps with
// Lorentzian fits, src-tauri/src/research.rs:133-163); this is the hand-written
// discrete adjoint of the step the plain kernel integrates (channel.metal:82-92
// without collisions or termination):
//     v' = v + kq E(x, n; theta),   x' = x + v' dt,
//     E  = a_rf(n) E_rf(x) + s_dc E_dc(x) + a_ac(n) E_ac(x),
//     a_rf = V_rf sin(2 pi frac(f_rf dt n)),  a_ac = V_ac sin(2 pi frac(f_ac dt n + phi)).
// Backward step (lambda' = dJ/d(x', v')):
//     mu = lambda_v' + dt lambda_x',  G = kq mu,
//     lambda_x = lambda_x' + (dE/dx)^T G,  lambda_v = mu,  dJ/dtheta += (dE/dtheta)^T G.
// With gas, a step that ends in a collision maps the post-kick velocity v by
// v' = v - mr g + mr |g| d with g and d from the step's Philox draws (frozen:
// the decisions are those of the primal run and carry no derivative); its
// adjoint lambda_v' <- (1 - mr) lambda_v' + mr (g / |g|) (d . lambda_v') comes
// before the step's own adjoint. The recompute pass then also stores each
// step's post-kick velocity, from which the walk redraws the collision.
// sqrt(N) checkpointing: the forward stores (x, v) every K steps; the backward
// recomputes each K-step segment from its checkpoint into a per-ion buffer and
// walks it backwards, so memory is O((N/K + K) ions) and work is 2 forwards
// plus one backward. Where the whole trajectory fits in device memory (24 B per
// ion-step, 48 B with gas: 31 GB for 65,536 ions x 20,000 steps) the forward
// writes every step's position itself and the backward walks it without a
// recompute (AdjObj.full): one forward plus one backward.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fp64_sample.h"
#include "kernels.h"
#include "mq_math.h"
#include "mq_types.h"

namespace mq {

struct Cell3 {
    uint32_t b;     // node index of the lower corner
    uint32_t cell;  // cell index i + (nx - 1) (j + (ny - 1) k)
    double tx, ty, tz;
    double dx, dy, dz;  // d(t)/d(x): inv_h inside the grid, 0 where the coordinate is clamped
};

__device__ __forceinline__ void locate_d(double x, int n, double inv_h, int nearest, int& i, double& t, double& dtdx)
{
    const double g = x * inv_h;
    double gc = g;
    bool clamped = false;
    if (!(gc >= 0.0)) { gc = 0.0; clamped = true; }
    if (gc > (double)(n - 1)) { gc = (double)(n - 1); clamped = true; }
    int ii = (int)gc;
    if (ii > n - 2) ii = n - 2;
    i = ii;
    t = nearest ? 0.0 : gc - (double)ii;
    dtdx = (nearest || clamped) ? 0.0 : inv_h;
}

__device__ __forceinline__ Cell3 cell_of(const BatchP& B, double x, double y, double z)
{
    Cell3 c;
    int i, j, k;
    locate_d(x, B.nx, B.inv_h, B.nearest, i, c.tx, c.dx);
    locate_d(y, B.ny, B.inv_h, B.nearest, j, c.ty, c.dy);
    locate_d(z, B.nz, B.inv_h, B.nearest, k, c.tz, c.dz);
    c.b = (uint32_t)i + (uint32_t)B.nx * ((uint32_t)j + (uint32_t)B.ny * (uint32_t)k);
    c.cell = (uint32_t)i + (uint32_t)(B.nx - 1) * ((uint32_t)j + (uint32_t)(B.ny - 1) * (uint32_t)k);
    return c;
}

// Trilinear value of one component from its 8 corner values (corner index
// dx + 2 dy + 4 dz), in the lerp order of the forward step, and the derivatives
// by x, y, z.
__device__ __forceinline__ double tri_val(const double f[8], const Cell3& c)
{
    const double c00 = lerp_d(f[0], f[1], c.tx), c10 = lerp_d(f[2], f[3], c.tx);
    const double c01 = lerp_d(f[4], f[5], c.tx), c11 = lerp_d(f[6], f[7], c.tx);
    return lerp_d(lerp_d(c00, c10, c.ty), lerp_d(c01, c11, c.ty), c.tz);
}

__device__ __forceinline__ double tri_jac(const double f[8], const Cell3& c, double jac[3])
{
    const double c000 = f[0], c100 = f[1], c010 = f[2], c110 = f[3];
    const double c001 = f[4], c101 = f[5], c011 = f[6], c111 = f[7];
    {
        const double c00 = lerp_d(c000, c100, c.tx), c10 = lerp_d(c010, c110, c.tx);
        const double c01 = lerp_d(c001, c101, c.tx), c11 = lerp_d(c011, c111, c.tx);
        const double lo = lerp_d(c00, c10, c.ty), hi = lerp_d(c01, c11, c.ty);
        const double dtx = lerp_d(lerp_d(c100 - c000, c110 - c010, c.ty), lerp_d(c101 - c001, c111 - c011, c.ty), c.tz);
        const double dty = lerp_d(c10 - c00, c11 - c01, c.tz);
        const double dtz = hi - lo;
        jac[0] = dtx * c.dx;
        jac[1] = dty * c.dy;
        jac[2] = dtz * c.dz;
        return lerp_d(lo, hi, c.tz);
    }
}

// The fields of the three bases at a cell: from the double node tables (any
// grid), or from per-cell float records of the 8 corner node values (grids
// whose values are all exactly floats: each value converts back to the same
// double, so the arithmetic and results are those of the node tables, from 18
// contiguous 16-byte loads instead of 72 scattered ones).
struct NodeGrids {
    const double *rf, *dc, *ac;
    uint32_t sy, sz;
};
struct CellRecs {
    const float4* rec;   // ADJ_REC_F4 per cell: [basis][corner][component]
};

__device__ __forceinline__ void corners(const double* __restrict__ F, uint32_t b, uint32_t sy, uint32_t sz, int d,
                                        double f[8])
{
    f[0] = F[3 * b + d]; f[1] = F[3 * (b + 1) + d];
    f[2] = F[3 * (b + sy) + d]; f[3] = F[3 * (b + 1 + sy) + d];
    f[4] = F[3 * (b + sz) + d]; f[5] = F[3 * (b + 1 + sz) + d];
    f[6] = F[3 * (b + sy + sz) + d]; f[7] = F[3 * (b + 1 + sy + sz) + d];
}

__device__ __forceinline__ void load_rec(const CellRecs& G, const Cell3& c, float v[72])
{
    const float4* p = G.rec + (size_t)c.cell * ADJ_REC_F4;
#pragma unroll
    for (int q = 0; q < ADJ_REC_F4; ++q) {
        const float4 t = p[q];
        v[4 * q] = t.x; v[4 * q + 1] = t.y; v[4 * q + 2] = t.z; v[4 * q + 3] = t.w;
    }
}

// Per-lane cache of the grid data of one cell. With float records the lane
// keeps its cell's record in registers and reloads it only when the ion has
// moved to another cell (the loads are masked to those lanes), which takes the
// per-step record traffic through the vector L1 down to the cell changes. The
// node tables are read at each use.
template <class G_> struct Cache;
template <> struct Cache<CellRecs> {
    uint32_t cell = 0xFFFFFFFFu;
    float v[72];
};
template <> struct Cache<NodeGrids> {};

__device__ __forceinline__ void refresh(const CellRecs& G, const Cell3& c, Cache<CellRecs>& k)
{
    if (c.cell != k.cell) {
        k.cell = c.cell;
        load_rec(G, c, k.v);
    }
}
__device__ __forceinline__ void refresh(const NodeGrids&, const Cell3&, Cache<NodeGrids>&) {}

__device__ __forceinline__ void sample3(const NodeGrids& G, const Cell3& c, const Cache<NodeGrids>&, double e[3][3])
{
    const double* F[3] = {G.rf, G.dc, G.ac};
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double f[8];
            corners(F[q], c.b, G.sy, G.sz, d, f);
            e[q][d] = tri_val(f, c);
        }
}

__device__ __forceinline__ void sample3(const CellRecs&, const Cell3& c, const Cache<CellRecs>& k, double e[3][3])
{
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double f[8];
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) f[kk] = (double)k.v[24 * q + 3 * kk + d];
            e[q][d] = tri_val(f, c);
        }
}

__device__ __forceinline__ void sample3_jac(const NodeGrids& G, const Cell3& c, const Cache<NodeGrids>&, double e[3][3],
                                            double j[3][3][3])
{
    const double* F[3] = {G.rf, G.dc, G.ac};
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double f[8];
            corners(F[q], c.b, G.sy, G.sz, d, f);
            e[q][d] = tri_jac(f, c, j[q][d]);
        }
}

__device__ __forceinline__ void sample3_jac(const CellRecs&, const Cell3& c, const Cache<CellRecs>& k, double e[3][3],
                                            double j[3][3][3])
{
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double f[8];
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) f[kk] = (double)k.v[24 * q + 3 * kk + d];
            e[q][d] = tri_jac(f, c, j[q][d]);
        }
}

// One forward step (identical arithmetic to k_plain without collisions). The
// drive amplitudes do not depend on the cell, so they are computed while the
// record loads are in flight.
template <class G_>
__device__ __forceinline__ void step_fwd(const BatchP& B, const JobP& J, const G_& G, Cache<G_>& K, double phi,
                                         uint32_t n, double& x, double& y, double& z, double& vx, double& vy,
                                         double& vz)
{
    const Cell3 c = cell_of(B, x, y, z);
    refresh(G, c, K);
    const double a_rf = J.v_rf * sin2pi_sel(frac_d(J.w_rf * (double)n));
    const double a_ac = J.v_ac * sin2pi_sel(frac_d(fma(J.w_ac, (double)n, phi)));
    double e3[3][3];
    sample3(G, c, K, e3);
    const double* er = e3[0];
    const double* ed = e3[1];
    const double* ea = e3[2];
    double E;
    E = (a_rf * er[0] + J.s_dc * ed[0]) + a_ac * ea[0]; vx = fma(J.kq, E, vx); x = fma(vx, B.dt, x);
    E = (a_rf * er[1] + J.s_dc * ed[1]) + a_ac * ea[1]; vy = fma(J.kq, E, vy); y = fma(vy, B.dt, y);
    E = (a_rf * er[2] + J.s_dc * ed[2]) + a_ac * ea[2]; vz = fma(J.kq, E, vz); z = fma(vz, B.dt, z);
}

// Thermal collision of step n on the post-kick velocity (vx, vy, vz): the
// Philox branch of k_plain, operation for operation, so the decisions and the
// new velocity are the plain path's. With the draws frozen the collision is
// the map v' = v - mr g + mr |g| d (g = v minus the drawn gas velocity, d the
// drawn direction); on a collision it returns true with g / |g| and d, the
// factors of its Jacobian (1 - mr) I + mr d (g / |g|)^T.
__device__ __forceinline__ bool collide_frozen(const JobP& J, uint32_t ion, uint32_t n, double& vx, double& vy,
                                               double& vz, double gh[3], double d[3])
{
    const double bound = norm3(vx, vy, vz) + J.v_bound;
    const u32x4 cw = philox(ion, n >> 2, 0, D_CAND, J.key0, J.key1);
    if (!(u01(pick(cw, n & 3)) < J.nsig_dt * bound)) return false;
    const u32x4 w = philox(ion, n, 0, D_COLL, J.key0, J.key1);
    const double r1 = sqrt(-2.0 * log_d(u01(w.a))), r2 = sqrt(-2.0 * log_d(u01(w.c)));
    double s1, c1, s2, c2;
    sincos2pi(u01(w.b), s1, c1);
    sincos2pi(u01(w.d), s2, c2);
    const double gx = vx - J.s_gas * (r1 * c1), gy = vy - J.s_gas * (r1 * s1), gz = vz - J.s_gas * (r2 * c2);
    const double gl = norm3(gx, gy, gz);
    const u32x4 w2 = philox(ion, n, 1, D_COLL, J.key0, J.key1);
    if (!(u01(w2.a) * bound < gl)) return false;
    const double c = 2.0 * u01(w2.b) - 1.0;
    double sp, cp;
    sincos2pi(u01(w2.c), sp, cp);
    const double sn = sqrt(fmax(0.0, 1.0 - c * c));
    const double mg = J.mr * gl;
    d[0] = sn * cp; d[1] = sn * sp; d[2] = c;
    gh[0] = gx / gl; gh[1] = gy / gl; gh[2] = gz / gl;
    vx = (vx - J.mr * gx) + mg * d[0];
    vy = (vy - J.mr * gy) + mg * d[1];
    vz = (vz - J.mr * gz) + mg * d[2];
    return true;
}

__device__ __forceinline__ double sigm(double u) { return 1.0 / (1.0 + exp(-u)); }

// Objective of one ion at its final state and dJ_i/d(x, y, z).
__device__ __forceinline__ double objective(const AdjObj& O, double x, double y, double z, double g[3])
{
    const double u = (z - O.z_exit) / O.scale;
    const double st = sigm(u);
    double Ji = st;
    g[0] = 0.0; g[1] = 0.0;
    g[2] = st * (1.0 - st) / O.scale;
    if (O.w_leak != 0.0) {
        const double rx = x - O.axis_x, ry = y - O.axis_y;
        const double r = sqrt(rx * rx + ry * ry);
        const double ur = (r - O.r_leak) / O.scale;
        const double sl = sigm(ur);
        Ji -= O.w_leak * sl;
        if (r > 0.0) {
            const double k = -O.w_leak * sl * (1.0 - sl) / O.scale / r;
            g[0] = k * rx;
            g[1] = k * ry;
        }
    }
    return Ji;
}

// Forward: N steps per ion, (x, v) stored every K steps (slot s holds step s*K).
// GAS: thermal collisions with frozen draws (a separate instantiation, so the
// collision-free kernels keep their register budget).
template <class G_, bool GAS>
__global__ __launch_bounds__(256) void k_adj_forward(IonState S, const JobP* __restrict__ jobs, BatchP B, G_ G,
                                                    AdjObj O, AdjBuf A, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const JobP J = jobs[S.job[i]];
    double x = S.x[i], y = S.y[i], z = S.z[i], vx = S.vx[i], vy = S.vy[i], vz = S.vz[i];
    const double phi = S.phi[i];
    Cache<G_> K;
    // The job parameters must have arrived before the loop: loads still in
    // flight at the loop head make the wait before their first use in each
    // step also wait for that step's record loads.
    asm volatile("" ::"v"(J.kq), "v"(J.v_rf), "v"(J.w_rf), "v"(J.v_ac), "v"(J.w_ac), "v"(J.s_dc), "v"(phi));
    const uint32_t ion = (uint32_t)(i - J.off);
    for (uint32_t s = 0; s < O.nsteps; ++s) {
        if (s % O.ckpt == 0) {
            const size_t o = (size_t)(s / O.ckpt) * n + i;
            A.ck[0][o] = x; A.ck[1][o] = y; A.ck[2][o] = z; A.ck[3][o] = vx; A.ck[4][o] = vy; A.ck[5][o] = vz;
        }
        // the trajectory slot of step s (with O.full), written with non-temporal stores: the stream of
        // positions (31 GB for 65,536 ions x 20,000 steps) would otherwise evict the cell records from L2
        const size_t o = (size_t)s * n + i;
        if (O.full) {
            __builtin_nontemporal_store(x, &A.seg[0][o]);
            __builtin_nontemporal_store(y, &A.seg[1][o]);
            __builtin_nontemporal_store(z, &A.seg[2][o]);
        }
        step_fwd(B, J, G, K, phi, s, x, y, z, vx, vy, vz);
        if constexpr (GAS) {
            if (O.full) {
                __builtin_nontemporal_store(vx, &A.seg[3][o]);
                __builtin_nontemporal_store(vy, &A.seg[4][o]);
                __builtin_nontemporal_store(vz, &A.seg[5][o]);
            }
            double gh[3], d[3];
            (void)collide_frozen(J, ion, s, vx, vy, vz, gh, d);
        }
    }
    double g[3];
    A.Ji[i] = objective(O, x, y, z, g);
    A.fin[0][i] = x; A.fin[1][i] = y; A.fin[2][i] = z;
}

// Adjoint variables and per-ion parameter gradients of the backward walk.
struct Adj {
    double lx, ly, lz, lvx, lvy, lvz;
    double gvrf, gfrf, gvac, gfac, gdc;
};

// One backward step s at the cell c of the recomputed position x_s, whose
// record the caller has just refreshed: the drive phases and the adjoint
// momentum come first, while the record loads are in flight.
template <class G_>
__device__ __forceinline__ void walk_step(const BatchP& B, const JobP& J, const G_& G, const Cache<G_>& K,
                                          const Cell3& c, double phi, double two_pi_dt, uint32_t s, Adj& a)
{
    double srf, crf, sac, cac;
    sincos2pi_sel(frac_d(J.w_rf * (double)s), srf, crf);
    sincos2pi_sel(frac_d(fma(J.w_ac, (double)s, phi)), sac, cac);
    const double a_rf = J.v_rf * srf, a_ac = J.v_ac * sac;
    const double mx = a.lvx + B.dt * a.lx, my = a.lvy + B.dt * a.ly, mz = a.lvz + B.dt * a.lz;
    const double Gx = J.kq * mx, Gy = J.kq * my, Gz = J.kq * mz;
    double e3[3][3], j3[3][3][3];
    sample3_jac(G, c, K, e3, j3);
    const double* er = e3[0];
    const double* ed = e3[1];
    const double* ea = e3[2];
    const double (&jr)[3][3] = j3[0];
    const double (&jd)[3][3] = j3[1];
    const double (&ja)[3][3] = j3[2];
    // lambda_x += (dE/dx)^T G
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const double t = Gx * ((a_rf * jr[0][e] + J.s_dc * jd[0][e]) + a_ac * ja[0][e]) +
                         Gy * ((a_rf * jr[1][e] + J.s_dc * jd[1][e]) + a_ac * ja[1][e]) +
                         Gz * ((a_rf * jr[2][e] + J.s_dc * jd[2][e]) + a_ac * ja[2][e]);
        if (e == 0) a.lx += t; else if (e == 1) a.ly += t; else a.lz += t;
    }
    a.lvx = mx; a.lvy = my; a.lvz = mz;
    const double gr = Gx * er[0] + Gy * er[1] + Gz * er[2];
    const double ga = Gx * ea[0] + Gy * ea[1] + Gz * ea[2];
    a.gvrf += srf * gr;
    a.gfrf += J.v_rf * crf * two_pi_dt * (double)s * gr;
    a.gvac += sac * ga;
    a.gfac += J.v_ac * cac * two_pi_dt * (double)s * ga;
    a.gdc += Gx * ed[0] + Gy * ed[1] + Gz * ed[2];
}

// A step's stored state as the backward walk reads it: the position before the
// step and, with gas, the post-kick velocity.
struct Stored {
    double x, y, z;
    double ux = 0.0, uy = 0.0, uz = 0.0;
};

// Backward: per ion, segments last to first; per-ion partial gradients and the
// gradient with respect to the initial state. In the walk down a segment the
// stored state of step s - 2 is read during step s.
template <class G_, bool GAS>
__global__ __launch_bounds__(256) void k_adj_backward(IonState S, const JobP* __restrict__ jobs, BatchP B, G_ G,
                                                     AdjObj O, AdjBuf A, uint32_t n, double scale)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const JobP J = jobs[S.job[i]];
    const double phi = S.phi[i];
    double g[3];
    (void)objective(O, A.fin[0][i], A.fin[1][i], A.fin[2][i], g);
    Adj a{scale * g[0], scale * g[1], scale * g[2], 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double two_pi_dt = 6.283185307179586 * B.dt;
    const uint32_t nseg = (O.nsteps + O.ckpt - 1) / O.ckpt;
    Cache<G_> K;  // record cache, kept across segments (records are keyed by the cell index)
    asm volatile("" ::"v"(J.kq), "v"(J.v_rf), "v"(J.w_rf), "v"(J.v_ac), "v"(J.w_ac), "v"(J.s_dc), "v"(phi));  // as in k_adj_forward
    const uint32_t ion = (uint32_t)(i - J.off);  // with GAS, seg[3..5] hold each step's post-kick velocity
    for (int seg = (int)nseg - 1; seg >= 0; --seg) {
        const uint32_t s0 = (uint32_t)seg * O.ckpt;
        const uint32_t s1 = min(O.nsteps, s0 + O.ckpt);
        const size_t oc = (size_t)seg * n + i;
        double x = A.ck[0][oc], y = A.ck[1][oc], z = A.ck[2][oc], vx = A.ck[3][oc], vy = A.ck[4][oc], vz = A.ck[5][oc];
        // recompute the segment's positions x_s, s0 <= s < s1 (with O.full the forward has stored them)
        for (uint32_t s = s0; !O.full && s < s1; ++s) {
            const size_t o = (size_t)(s - s0) * n + i;
            A.seg[0][o] = x; A.seg[1][o] = y; A.seg[2][o] = z;
            step_fwd(B, J, G, K, phi, s, x, y, z, vx, vy, vz);
            if constexpr (GAS) {
                A.seg[3][o] = vx; A.seg[4][o] = vy; A.seg[5][o] = vz;
                double gh[3], d[3];
                (void)collide_frozen(J, ion, s, vx, vy, vz, gh, d);
            }
        }
        // Walk it backwards. Each step's stored position (and with GAS its post-kick velocity) is read at the
        // top of the step two steps before its own, so that the HBM read has about two steps to arrive: read
        // one step ahead, as before, the walk took 44 ms for 65,536 ions x 20,000 steps, two steps ahead 32 ms.
        // Look-ahead indices below s0 are clamped to s0 (redundant reads in bounds).
        const auto load = [&](int s) {
            const size_t o = (size_t)((uint32_t)max(s, (int)s0) - s0) * n + i;
            Stored r;
            r.x = __builtin_nontemporal_load(&A.seg[0][o]);
            r.y = __builtin_nontemporal_load(&A.seg[1][o]);
            r.z = __builtin_nontemporal_load(&A.seg[2][o]);
            if constexpr (GAS) {
                r.ux = __builtin_nontemporal_load(&A.seg[3][o]);
                r.uy = __builtin_nontemporal_load(&A.seg[4][o]);
                r.uz = __builtin_nontemporal_load(&A.seg[5][o]);
            }
            return r;
        };
        Stored p = load((int)s1 - 1), q = load((int)s1 - 2);
        for (int s = (int)s1 - 1; s >= (int)s0; --s) {
            const Stored r = load(s - 2);
            const Cell3 c = cell_of(B, p.x, p.y, p.z);
            refresh(G, c, K);
            if constexpr (GAS) {
                // the collision that ended step s: lambda_v <- ((1 - mr) I + mr d gh^T)^T lambda_v
                double ux = p.ux, uy = p.uy, uz = p.uz, gh[3], d[3];
                if (collide_frozen(J, ion, (uint32_t)s, ux, uy, uz, gh, d)) {
                    const double dl = J.mr * ((d[0] * a.lvx + d[1] * a.lvy) + d[2] * a.lvz);
                    a.lvx = (1.0 - J.mr) * a.lvx + gh[0] * dl;
                    a.lvy = (1.0 - J.mr) * a.lvy + gh[1] * dl;
                    a.lvz = (1.0 - J.mr) * a.lvz + gh[2] * dl;
                }
            }
            walk_step(B, J, G, K, c, phi, two_pi_dt, (uint32_t)s, a);
            p = q;
            q = r;
        }
    }
    A.gp[0][i] = a.gvrf; A.gp[1][i] = a.gfrf; A.gp[2][i] = a.gvac; A.gp[3][i] = a.gfac; A.gp[4][i] = a.gdc;
    A.gs[0][i] = a.lx; A.gs[1][i] = a.ly; A.gs[2][i] = a.lz; A.gs[3][i] = a.lvx; A.gs[4][i] = a.lvy; A.gs[5][i] = a.lvz;
}

// Deterministic sum of `cols` columns of length n (fixed order: 1024-element
// blocks summed by a fixed tree, then the block sums in order on one thread).
__global__ __launch_bounds__(256) void k_colsum_blocks(const double* const* __restrict__ cols, uint32_t n,
                                                      double* __restrict__ part)
{
    __shared__ double sh[256];
    const uint32_t c = blockIdx.y, t = threadIdx.x, base = blockIdx.x * 1024;
    double s = 0.0;
    for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t i = base + t * 4 + k;
        if (i < n) s += cols[c][i];
    }
    sh[t] = s;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    if (t == 0) part[(size_t)c * gridDim.x + blockIdx.x] = sh[0];
}

__global__ void k_colsum_final(const double* __restrict__ part, uint32_t nblocks, uint32_t ncols, double* __restrict__ out)
{
    const uint32_t c = threadIdx.x;
    if (c >= ncols) return;
    double s = 0.0;
    for (uint32_t b = 0; b < nblocks; ++b) s += part[(size_t)c * nblocks + b];
    out[c] = s;
}

static inline uint32_t nblk(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

static NodeGrids node_grids(const BatchP& B, const AdjGrids& G)
{
    return NodeGrids{G.rf, G.dc, G.ac, (uint32_t)B.nx, (uint32_t)B.nx * (uint32_t)B.ny};
}

void launch_adj_forward(const IonState& S, const JobP* jobs, const BatchP& B, const AdjGrids& G, const AdjObj& O,
                        const AdjBuf& A, uint32_t n, hipStream_t st)
{
    if (!n) return;
    if (O.gas) {
        if (G.rec) k_adj_forward<CellRecs, true><<<nblk(n, 256), 256, 0, st>>>(S, jobs, B, CellRecs{G.rec}, O, A, n);
        else k_adj_forward<NodeGrids, true><<<nblk(n, 256), 256, 0, st>>>(S, jobs, B, node_grids(B, G), O, A, n);
    } else {
        if (G.rec) k_adj_forward<CellRecs, false><<<nblk(n, 256), 256, 0, st>>>(S, jobs, B, CellRecs{G.rec}, O, A, n);
        else k_adj_forward<NodeGrids, false><<<nblk(n, 256), 256, 0, st>>>(S, jobs, B, node_grids(B, G), O, A, n);
    }
}

void launch_adj_backward(const IonState& S, const JobP* jobs, const BatchP& B, const AdjGrids& G, const AdjObj& O,
                         const AdjBuf& A, uint32_t n, double scale, hipStream_t st)
{
    if (!n) return;
    if (O.gas) {
        if (G.rec)
            k_adj_backward<CellRecs, true><<<nblk(n, 256), 256, 0, st>>>(S, jobs, B, CellRecs{G.rec}, O, A, n, scale);
        else
            k_adj_backward<NodeGrids, true><<<nblk(n, 256), 256, 0, st>>>(S, jobs, B, node_grids(B, G), O, A, n, scale);
    } else {
        if (G.rec)
            k_adj_backward<CellRecs, false><<<nblk(n, 256), 256, 0, st>>>(S, jobs, B, CellRecs{G.rec}, O, A, n, scale);
        else
            k_adj_backward<NodeGrids, false><<<nblk(n, 256), 256, 0, st>>>(S, jobs, B, node_grids(B, G), O, A, n, scale);
    }
}

// flag = 0 when some value of a is not exactly a float (flag starts at 1).
__global__ void k_exact_f32(const double* __restrict__ a, uint64_t n, uint32_t* __restrict__ flag)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && (double)(float)a[i] != a[i]) *flag = 0u;
}

// Per-cell float records of the corner node values of the three bases
// (float index 24 basis + 3 corner + component, corner = dx + 2 dy + 4 dz).
__global__ void k_build_adj_recs(float4* __restrict__ rec, const double* __restrict__ rf, const double* __restrict__ dc,
                                 const double* __restrict__ ac, int nx, int ny, int nz)
{
    const uint32_t cx = (uint32_t)(nx - 1), cy = (uint32_t)(ny - 1), cz = (uint32_t)(nz - 1);
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (uint64_t)cx * cy * cz) return;
    const uint32_t i = (uint32_t)(c % cx), j = (uint32_t)((c / cx) % cy), k = (uint32_t)(c / ((uint64_t)cx * cy));
    const uint32_t sy = (uint32_t)nx, sz = (uint32_t)nx * (uint32_t)ny;
    const uint32_t b = i + sy * (j + (uint32_t)ny * k);
    const uint32_t node[8] = {b, b + 1, b + sy, b + 1 + sy, b + sz, b + 1 + sz, b + sy + sz, b + 1 + sy + sz};
    const double* F[3] = {rf, dc, ac};
    float v[72];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int kk = 0; kk < 8; ++kk)
#pragma unroll
            for (int d = 0; d < 3; ++d) v[24 * q + 3 * kk + d] = (float)F[q][3 * node[kk] + d];
    float4* dst = rec + c * ADJ_REC_F4;
#pragma unroll
    for (int q = 0; q < ADJ_REC_F4; ++q) dst[q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
}

void launch_exact_f32(const double* a, uint64_t n, uint32_t* flag, hipStream_t st)
{
    if (n) k_exact_f32<<<(uint32_t)((n + 255) / 256), 256, 0, st>>>(a, n, flag);
}

void launch_build_adj_recs(float4* rec, const double* rf, const double* dc, const double* ac, int nx, int ny, int nz,
                           hipStream_t st)
{
    const uint64_t nc = (uint64_t)(nx - 1) * (ny - 1) * (nz - 1);
    if (nc) k_build_adj_recs<<<(uint32_t)((nc + 127) / 128), 128, 0, st>>>(rec, rf, dc, ac, nx, ny, nz);
}

void launch_colsum(const double* const* cols_dev, uint32_t ncols, uint32_t n, double* part, double* out, hipStream_t st)
{
    const uint32_t nb = std::max(1u, nblk(n, 1024));
    k_colsum_blocks<<<dim3(nb, ncols), 256, 0, st>>>(cols_dev, n, part);
    k_colsum_final<<<1, 64, 0, st>>>(part, nb, ncols, out);
}

}  // namespace mq
