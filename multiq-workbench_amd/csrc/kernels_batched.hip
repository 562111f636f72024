// kernels_batched.hip -- batched FP32 path: per-cell coefficient tables and the
// persistent wavefront worker.
//
// Counterparts in the reference: tc_brick (channel_tiled.metal:54-76), the
// persistent worker tc_tail (:386-482) with the register corner cache
// (:97-124), and the live-list compaction tc_filter / tc_size_args (:341-374).
//
// Design for CDNA4:
//  * Fields: one 304-byte record per grid cell (at a 320-byte stride) holding
//    the trilinear polynomial coefficients of the RF, combined-DC and AC bases
//    (72 FP32) and the conductor bits of its 8 corners. A cell change costs 19
//    contiguous 16-byte loads; the record stays in VGPRs while the ion stays in
//    the cell.
//  * Precision: FP32 field values, interpolation, kick and collision sampling;
//    FP64 position, drift, time (step counter) and RF/AC phase.
//  * Scheduling: each wavefront keeps 64 ions in flight; every `period` steps
//    the lanes whose ion finished (or used up the launch's chunk) are refilled
//    from the live list with one atomic per wavefront (ballot + popcount
//    prefix), and ions that used up the chunk are appended to the next live
//    list the same way.
#include <hip/hip_runtime.h>

#include "fp32_cells.h"
#include "kernels.h"
#include "mq_math.h"
#include "mq_types.h"

namespace mq {

// ------------------------------------------------------------- cell table
// 128-thread launches: the bound lets a thread keep the 72 FP64 corner values without spilling.
__global__ __launch_bounds__(128) void k_build_cells(float4* __restrict__ cells, const double* __restrict__ rf,
                              const double* __restrict__ dc, const double* __restrict__ ac,
                              const uint8_t* __restrict__ mask, int nx, int ny, int nz)
{
    const uint32_t cx = (uint32_t)(nx - 1), cy = (uint32_t)(ny - 1), cz = (uint32_t)(nz - 1);
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cx * cy * cz) return;
    const uint32_t i = c % cx, j = (c / cx) % cy, k = c / (cx * cy);
    const uint32_t sy = (uint32_t)nx, sz = (uint32_t)nx * (uint32_t)ny;
    const uint32_t b = i + sy * (j + (uint32_t)ny * k);
    const uint32_t node[8] = {b, b + 1, b + sy, b + 1 + sy, b + sz, b + 1 + sz, b + sy + sz, b + 1 + sy + sz};
    float out[76];
    const double* F[3] = {rf, dc, ac};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            // corner f_abc, a = dx, b = dy, c = dz
            const double f000 = F[q][3 * node[0] + d], f100 = F[q][3 * node[1] + d];
            const double f010 = F[q][3 * node[2] + d], f110 = F[q][3 * node[3] + d];
            const double f001 = F[q][3 * node[4] + d], f101 = F[q][3 * node[5] + d];
            const double f011 = F[q][3 * node[6] + d], f111 = F[q][3 * node[7] + d];
            float* o = out + 24 * d + 8 * q;  // component-major (fp32_cells.h)
            o[0] = (float)f000;
            o[1] = (float)(f100 - f000);
            o[2] = (float)(f010 - f000);
            o[3] = (float)(f001 - f000);
            o[4] = (float)((f110 - f100) - (f010 - f000));
            o[5] = (float)((f011 - f010) - (f001 - f000));
            o[6] = (float)((f101 - f100) - (f001 - f000));
            o[7] = (float)(((f111 - f110) - (f101 - f100)) - ((f011 - f010) - (f001 - f000)));
        }
    }
    uint32_t bits = 0;
    if (mask) {
#pragma unroll
        for (int q = 0; q < 8; ++q) bits |= (mask[node[q]] ? 1u : 0u) << q;
    }
    out[72] = __uint_as_float(bits);
    out[73] = out[74] = out[75] = 0.0f;
    float4* dst = cells + (size_t)c * CELL_STRIDE_F4;
#pragma unroll
    for (int q = 0; q < CELL_F4; ++q) dst[q] = make_float4(out[4 * q], out[4 * q + 1], out[4 * q + 2], out[4 * q + 3]);
#pragma unroll
    for (int q = CELL_F4; q < CELL_STRIDE_F4; ++q) dst[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ void k_fill_hist(uint32_t* hist, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) hist[i] = 0u;
}

__global__ void k_iota(uint32_t* list, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) list[i] = i;
}

// Counter block: queue heads at [q * CTR_Q] (one 128-byte line each), the list
// length at [CTR_COUNT]. Zero the heads of the launch about to run and the
// whole block of the next one.
__global__ void k_prepare(uint32_t* ctr_in, uint32_t* ctr_out)
{
    const uint32_t t = threadIdx.x;
    if (t < MAX_QUEUES) {
        ctr_in[t * CTR_Q] = 0;
        ctr_out[t * CTR_Q] = 0;
    }
    if (t == 0) ctr_out[CTR_COUNT] = 0;
}

// Re-sort bin of a position (coarse cell bins, z-major, x fastest).
__device__ __forceinline__ uint32_t sort_bin(const BatchP& B, const SortGeom& G, double x, double y, double z)
{
    const int ci = min((int)fmin(fmax(x * B.inv_h, 0.0), (double)(B.nx - 1)), B.nx - 2);
    const int cj = min((int)fmin(fmax(y * B.inv_h, 0.0), (double)(B.ny - 1)), B.ny - 2);
    const int ck = min((int)fmin(fmax(z * B.inv_h, 0.0), (double)(B.nz - 1)), B.nz - 2);
    return (((uint32_t)ck >> G.sz) * G.nby + ((uint32_t)cj >> G.sy)) * G.nbx + ((uint32_t)ci >> G.sx);
}

// ------------------------------------------------------------- the worker
struct Lane {
    double x, y, z, phi, phr, w_rf, w_ac;
    float vx, vy, vz;
    float kq, vrf, vac, sdc, nsig, sgas, vb, mr, omr, floorv, mfpm;
    uint32_t st, ncoll, cd, k0, k1, ion, id, visit;
    float tx, ty, tz, a_rf, a_ac;
    uint32_t corner;  // nearest corner of the cell (bit dx + 2 dy + 4 dz), for the crash test
    uint32_t cw1, cw2, cw3;
    uint32_t xs0, xs1, xs2, xs3;  // xoshiro128** state (compat streams)
    int32_t cell;
};

__device__ __forceinline__ int32_t locate(const BatchP& B, Lane& L)
{
    double gx = fmin(fmax(L.x * B.inv_h, 0.0), (double)(B.nx - 1));
    double gy = fmin(fmax(L.y * B.inv_h, 0.0), (double)(B.ny - 1));
    double gz = fmin(fmax(L.z * B.inv_h, 0.0), (double)(B.nz - 1));
    int i = min((int)gx, B.nx - 2), j = min((int)gy, B.ny - 2), k = min((int)gz, B.nz - 2);
    const float fx = (float)(gx - (double)i), fy = (float)(gy - (double)j), fz = (float)(gz - (double)k);
    L.corner = (fx >= 0.5f ? 1u : 0u) | (fy >= 0.5f ? 2u : 0u) | (fz >= 0.5f ? 4u : 0u);
    L.tx = B.nearest ? 0.0f : fx;
    L.ty = B.nearest ? 0.0f : fy;
    L.tz = B.nearest ? 0.0f : fz;
    return i + (B.nx - 1) * (j + (B.ny - 1) * k);
}

// Field of the record R at the lane's (tx, ty, tz): a_rf E_rf + s_dc E_dc + a_ac E_ac.
__device__ __forceinline__ void cell_field(const float4 R[CELL_F4], const Lane& L, float e[3])
{
    const float pxy = L.tx * L.ty, pyz = L.ty * L.tz, pxz = L.tx * L.tz, pxyz = pxy * L.tz;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float er = poly(R, 0, d, L.tx, L.ty, L.tz, pxy, pyz, pxz, pxyz);
        const float ed = poly(R, 1, d, L.tx, L.ty, L.tz, pxy, pyz, pxz, pxyz);
        const float ea = poly(R, 2, d, L.tx, L.ty, L.tz, pxy, pyz, pxz, pxyz);
        e[d] = fmaf(L.a_ac, ea, fmaf(L.sdc, ed, L.a_rf * er));
    }
}

// ---- staged tiles (the reference's tiled rounds, tc_push, channel_tiled.metal:237-334):
// a workgroup takes 256 consecutive entries of the re-sorted list and stages the
// records of the 6 x 6 x 6 cells around the bin of its first ion in LDS. A reload
// inside the tile reads LDS; outside it, global memory.
constexpr int TILE_N = 6;
constexpr int TILE_CELLS = TILE_N * TILE_N * TILE_N;
constexpr uint32_t TILE_LDS_BYTES = TILE_CELLS * (CELL_F4 - 1) * 16 + TILE_CELLS * 4 + 16;  // 63 088 B

struct Tile {
    float4* rec;    // [TILE_CELLS][CELL_F4 - 1]: the coefficients
    uint32_t* msk;  // [TILE_CELLS]: the corner conductor bits
    int32_t* org;   // [4]: cell index of the tile's corner (i, j, k), then the chunk's first entry
};

__device__ __forceinline__ void reload_tiled(const BatchP& B, const float4* __restrict__ cells, const Tile& T,
                                             int32_t c, float4 R[CELL_F4])
{
    const int32_t cx = B.nx - 1, cy = B.ny - 1;
    const int32_t i = c % cx, j = (c / cx) % cy, k = c / (cx * cy);
    const int32_t ti = i - T.org[0], tj = j - T.org[1], tk = k - T.org[2];
    if ((uint32_t)ti < (uint32_t)TILE_N && (uint32_t)tj < (uint32_t)TILE_N && (uint32_t)tk < (uint32_t)TILE_N) {
        const int32_t t = ti + TILE_N * (tj + TILE_N * tk);
#pragma unroll
        for (int q = 0; q < CELL_F4 - 1; ++q) R[q] = T.rec[t * (CELL_F4 - 1) + q];
        R[CELL_F4 - 1] = make_float4(__uint_as_float(T.msk[t]), 0.0f, 0.0f, 0.0f);
    } else {
        load_cell(cells, c, R);
    }
}

__device__ __forceinline__ void phases(Lane& L)
{
    double pr = fma(L.w_rf, (double)L.st, L.phr);  // phr = 0 unless the ion was born on a running clock
    pr = pr - floor(pr);
    double pa = fma(L.w_ac, (double)L.st, L.phi);
    pa = pa - floor(pa);
    L.a_rf = L.vrf * sin_rev((float)pr);
    L.a_ac = L.vac * sin_rev((float)pa);
}

__device__ __forceinline__ uint32_t countdown_f(float x)
{
    if (!(x < 2147483520.0f)) return 2147483647u;
    if (x < 1.0f) return 1u;
    return (uint32_t)x;
}

// Collision of a lane whose candidate fired at step pcur, on the Philox streams
// (channel.metal:93-113): the thermal bath's acceptance test and elastic
// scatter, or the legacy scatter and the next free path. Draws are keyed by
// (ion, step), so where and when a worker runs it does not change the result.
template <bool LEGACY>
__device__ __forceinline__ void collide_philox(Lane& L, uint32_t pcur, float dtf)
{
    if constexpr (!LEGACY) {
        const float bound = sqrtf(fmaf(L.vz, L.vz, fmaf(L.vy, L.vy, L.vx * L.vx))) + L.vb;
        const u32x4 w = philox(L.ion, pcur, 0, D_COLL, L.k0, L.k1);
        const float r1 = sqrtf(-2.0f * log_f(u01f(w.a))), r2 = sqrtf(-2.0f * log_f(u01f(w.c)));
        const float a1 = u01f(w.b), a2 = u01f(w.d);
        const float gx = L.vx - L.sgas * (r1 * cos_rev(a1));
        const float gy = L.vy - L.sgas * (r1 * sin_rev(a1));
        const float gz = L.vz - L.sgas * (r2 * cos_rev(a2));
        const float gl = sqrtf(fmaf(gz, gz, fmaf(gy, gy, gx * gx)));
        const u32x4 w2 = philox(L.ion, pcur, 1, D_COLL, L.k0, L.k1);
        if (u01f(w2.a) * bound < gl) {
            const float cth = 2.0f * u01f(w2.b) - 1.0f, a3 = u01f(w2.c);
            const float sn = sqrtf(fmaxf(0.0f, 1.0f - cth * cth));
            const float mg = L.mr * gl;
            L.vx = fmaf(mg, sn * cos_rev(a3), fmaf(-L.mr, gx, L.vx));
            L.vy = fmaf(mg, sn * sin_rev(a3), fmaf(-L.mr, gy, L.vy));
            L.vz = fmaf(mg, cth, fmaf(-L.mr, gz, L.vz));
            L.ncoll += 1;
        }
    } else {
        const u32x4 w = philox(L.ion, pcur, 0, D_COLL, L.k0, L.k1);
        const float vm = sqrtf(fmaf(L.vz, L.vz, fmaf(L.vy, L.vy, L.vx * L.vx)));
        if (vm > L.floorv) {
            const float cth = 2.0f * u01f(w.a) - 1.0f, a3 = u01f(w.b);
            const float sn = sqrtf(fmaxf(0.0f, 1.0f - cth * cth));
            const float mv = L.mr * vm;
            L.vx = fmaf(mv, sn * cos_rev(a3), L.vx * L.omr);
            L.vy = fmaf(mv, sn * sin_rev(a3), L.vy * L.omr);
            L.vz = fmaf(mv, cth, L.vz * L.omr);
            L.ncoll += 1;
        }
        const float mfp = -log_f(fmaxf(u01f(w.c), 1e-7f)) * L.mfpm;
        const float vm2 = sqrtf(fmaf(L.vz, L.vz, fmaf(L.vy, L.vy, L.vx * L.vx)));
        L.cd = countdown_f(floorf(mfp / fmaxf(vm2, 1e-3f) / dtf));
    }
}

// Status of a lane after its step (channel.metal:114-120), in the reference's
// order: return, exit, box, crash (the conductor bit of the nearest corner,
// inside the crash window), then timeout. k_batched writes the same tests out
// in its step loop.
// Evaluated without branches: the tests are selected in reverse priority.
__device__ __forceinline__ int classify(const BatchP& B, const Lane& L, uint32_t bits)
{
    int status = L.st >= B.max_steps ? 5 : 0;
    status = (L.z > B.cz0 && L.z < B.cz1 && ((bits >> L.corner) & 1u)) ? 1 : status;
    status = (L.x < 0.0 || L.y < 0.0 || L.z < 0.0 || L.x > B.lim0 || L.y > B.lim1 || L.z > B.lim2) ? 4 : status;
    status = L.z > B.z_exit ? 3 : status;
    status = L.z < B.z_ret ? 2 : status;
    return status;
}

// LEGACY selects the collision model at compile time (legacy countdown or the
// thermal bath), so each instantiation keeps only its own values live.
// COMPAT selects the reference's per-ion xoshiro128** stream (draw order of
// channel.metal:93-113) instead of Philox.
// STAGED: workgroup chunks of the re-sorted list with their tile staged in LDS
// (no refill inside a chunk; the launch's chunk is one refill period).
template <bool LEGACY, bool COMPAT, bool STAGED>
__global__ __launch_bounds__(256) void k_batched(WorkerArgs A)
{
    static_assert(!(STAGED && COMPAT), "staged chunks use the Philox streams");
    const BatchP& B = A.B;
    const IonState& S = A.S;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t count = A.ctr_in[CTR_COUNT];
    const uint32_t nq = A.nqueues;
    uint32_t q = nq > 1 ? blockIdx.x % nq : 0, tried = 0;
    const double dt = B.dt;
    const float dtf = (float)B.dt;
    Lane L;
    float4 R[CELL_F4];
    bool have = false, drained = false, pending = false;
    uint32_t pcur = 0;
    L.cell = -1;
    extern __shared__ float4 lds_tile[];
    Tile T{};
    if constexpr (STAGED) {
        T.rec = lds_tile;
        T.msk = reinterpret_cast<uint32_t*>(lds_tile + TILE_CELLS * (CELL_F4 - 1));
        T.org = reinterpret_cast<int32_t*>(T.msk + TILE_CELLS);
    }
    for (;;) {
        if constexpr (STAGED) {
            // ---- claim the next 256 entries for the workgroup and stage the tile around
            // the re-sort bin of the first one (the entries of a chunk are bin-sorted)
            __syncthreads();  // the previous chunk's last reads of the tile are done
            if (threadIdx.x == 0) {
                const uint32_t b0 = atomicAdd(&A.ctr_in[0], blockDim.x);
                T.org[3] = (int32_t)b0;
                if (b0 < count) {
                    const uint32_t id0 = A.list_in[b0];
                    const double g[3] = {S.x[id0] * B.inv_h, S.y[id0] * B.inv_h, S.z[id0] * B.inv_h};
                    const int32_t n[3] = {B.nx, B.ny, B.nz};
                    const uint32_t sh[3] = {A.sg.sx, A.sg.sy, A.sg.sz};
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const int32_t ci = min((int32_t)fmin(fmax(g[d], 0.0), (double)(n[d] - 1)), n[d] - 2);
                        T.org[d] = (int32_t)(((uint32_t)ci >> sh[d]) << sh[d]) - 1;  // one cell of halo
                    }
                }
            }
            __syncthreads();
            const uint32_t b0 = (uint32_t)T.org[3];
            if (b0 >= count) break;  // the same value in every thread of the workgroup
            const int32_t cx = B.nx - 1, cy = B.ny - 1, cz = B.nz - 1;
            for (uint32_t q = threadIdx.x; q < (uint32_t)(TILE_CELLS * CELL_F4); q += blockDim.x) {
                const int32_t t = (int32_t)(q / CELL_F4), f = (int32_t)(q % CELL_F4);
                const int32_t gi = T.org[0] + t % TILE_N, gj = T.org[1] + (t / TILE_N) % TILE_N,
                              gk = T.org[2] + t / (TILE_N * TILE_N);
                float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (gi >= 0 && gj >= 0 && gk >= 0 && gi < cx && gj < cy && gk < cz)
                    v = A.cells[(size_t)(gi + cx * (gj + cy * gk)) * CELL_STRIDE_F4 + f];
                if (f < CELL_F4 - 1) T.rec[t * (CELL_F4 - 1) + f] = v;
                else T.msk[t] = __float_as_uint(v.x);
            }
            __syncthreads();
            const uint32_t slot = b0 + threadIdx.x;
            if (slot < count) {
                const uint32_t id = A.list_in[slot];
                const JobP& J = A.jobs[S.job[id]];
                L.id = id;
                L.ion = S.sid ? S.sid[id] : (uint32_t)(id - J.off);
                L.phr = S.phr ? S.phr[id] : 0.0;
                L.x = S.x[id]; L.y = S.y[id]; L.z = S.z[id]; L.phi = S.phi[id];
                L.vx = S.fvx[id]; L.vy = S.fvy[id]; L.vz = S.fvz[id];
                L.st = S.steps[id]; L.ncoll = S.ncoll[id];
                if constexpr (LEGACY) L.cd = S.cd[id];
                L.w_rf = J.w_rf; L.w_ac = J.w_ac;
                L.kq = (float)J.kq; L.vrf = (float)J.v_rf; L.vac = (float)J.v_ac; L.sdc = (float)J.s_dc;
                L.nsig = (float)J.nsig_dt; L.sgas = (float)J.s_gas; L.vb = (float)J.v_bound;
                L.mr = (float)J.mr; L.omr = (float)J.one_m_mr; L.floorv = (float)J.floor_v;
                L.mfpm = (float)J.mfp_mean;
                L.k0 = J.key0; L.k1 = J.key1;
                L.visit = 0;
                L.cw1 = L.cw2 = L.cw3 = 0;
                L.cell = locate(B, L);
                reload_tiled(B, A.cells, T, L.cell, R);
                phases(L);
                have = true;
            }
        } else {
        // ---- refill idle lanes from the live list (one atomic per wavefront). With
        // several queues the list is cut into contiguous (spatially sorted) segments
        // and the wavefronts of one XCD start on their own segment, then steal.
        while (!drained) {
            const uint64_t idle = __ballot(!have && lane < A.lanes);
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
                    const JobP& J = A.jobs[S.job[id]];
                    L.id = id;
                    L.ion = S.sid ? S.sid[id] : (uint32_t)(id - J.off);
                    L.phr = S.phr ? S.phr[id] : 0.0;
                    L.x = S.x[id]; L.y = S.y[id]; L.z = S.z[id]; L.phi = S.phi[id];
                    L.vx = S.fvx[id]; L.vy = S.fvy[id]; L.vz = S.fvz[id];
                    L.st = S.steps[id]; L.ncoll = S.ncoll[id];
                    if constexpr (LEGACY) L.cd = S.cd[id];
                    L.w_rf = J.w_rf; L.w_ac = J.w_ac;
                    L.kq = (float)J.kq; L.vrf = (float)J.v_rf; L.vac = (float)J.v_ac; L.sdc = (float)J.s_dc;
                    L.nsig = (float)J.nsig_dt; L.sgas = (float)J.s_gas; L.vb = (float)J.v_bound;
                    L.mr = (float)J.mr; L.omr = (float)J.one_m_mr; L.floorv = (float)J.floor_v;
                    L.mfpm = (float)J.mfp_mean;
                    L.k0 = J.key0; L.k1 = J.key1;
                    L.visit = 0;
                    L.cw1 = L.cw2 = L.cw3 = 0;
                    if constexpr (COMPAT) {
                        L.xs0 = S.rng[4 * (size_t)id]; L.xs1 = S.rng[4 * (size_t)id + 1];
                        L.xs2 = S.rng[4 * (size_t)id + 2]; L.xs3 = S.rng[4 * (size_t)id + 3];
                    }
                    const int32_t c = locate(B, L);
                    if (c != L.cell) { L.cell = c; load_cell(A.cells, c, R); }
                    phases(L);
                    have = true;
                }
            }
            break;
        }
        if (!__any(have)) break;
        }
        // ---- advance every held ion by up to `period` steps. A lane whose collision
        // candidate fires (thermal) or whose countdown expires (legacy) becomes
        // `pending` and pauses; the collision code then runs once for every pending
        // lane of the wavefront when at least `coll_batch` lanes wait, when no lane
        // can step, or at the end of the period (the thermal analogue of tc_tail_leg,
        // channel_tiled.metal:485-585). An ion's own sequence of operations is
        // unchanged: its classification depends on its position only and runs at
        // the step, and its next kick waits for its collision.
        bool stop = false;
        int status = 0;
        // A staged chunk runs until every ion of the workgroup has stopped.
        for (uint32_t r = 0; STAGED || r < A.period; ++r) {
            if (have && !stop && !pending) {
                float e[3];
                cell_field(R, L, e);
                L.vx = fmaf(L.kq, e[0], L.vx);
                L.vy = fmaf(L.kq, e[1], L.vy);
                L.vz = fmaf(L.kq, e[2], L.vz);
                L.x = fma((double)L.vx, dt, L.x);
                L.y = fma((double)L.vy, dt, L.y);
                L.z = fma((double)L.vz, dt, L.z);
                const uint32_t cur = L.st;
                L.st += 1;
                // next-step field cell (hoisted above the collision code: it depends on x only)
                const int32_t c = locate(B, L);
                if (c != L.cell) {
                    L.cell = c;
                    if constexpr (STAGED) reload_tiled(B, A.cells, T, c, R);
                    else load_cell(A.cells, c, R);
                }
                phases(L);
                // ---- collision candidate of this step
                if constexpr (!LEGACY && COMPAT) {
                    // the reference draws a candidate uniform every step (channel.metal:94-95)
                    const float bound = sqrtf(fmaf(L.vz, L.vz, fmaf(L.vy, L.vy, L.vx * L.vx))) + L.vb;
                    if (xo_u01f(xo_next(L.xs0, L.xs1, L.xs2, L.xs3)) < L.nsig * bound) { pending = true; pcur = cur; }
                } else if constexpr (!LEGACY) {
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
                        if (u01f(cw) < L.nsig * bound) { pending = true; pcur = cur; }
                    }
                } else {
                    L.cd -= 1;
                    if (L.cd == 0) { pending = true; pcur = cur; }
                }
                L.visit += 1;
                // ---- classification (channel.metal:114-120)
                // (written out here rather than through classify(): reading the record's
                // conductor word only inside the crash window keeps the compiler from
                // copying freshly loaded record registers, which would wait on the loads)
                if (L.z < B.z_ret) status = 2;
                else if (L.z > B.z_exit) status = 3;
                else if (L.x < 0.0 || L.y < 0.0 || L.z < 0.0 || L.x > B.lim0 || L.y > B.lim1 || L.z > B.lim2) status = 4;
                else if (L.z > B.cz0 && L.z < B.cz1) {
                    const uint32_t bits = __float_as_uint(R[18].x);
                    if ((bits >> L.corner) & 1u) status = 1;
                }
                if (status == 0 && L.st >= B.max_steps) status = 5;
                if (status != 0 || L.visit >= A.chunk) stop = true;
            }
            // ---- batched collisions of the pending lanes
            const uint64_t pm = __ballot(pending);
            if (pm) {
                const uint32_t npend = (uint32_t)__popcll(pm);
                const uint64_t am = __ballot(have && !stop && !pending);
                if (npend >= A.coll_batch || am == 0 || (!STAGED && r + 1 == A.period)) {
                    if (pending && COMPAT) {
                        if constexpr (!LEGACY) {
                            const float bound = sqrtf(fmaf(L.vz, L.vz, fmaf(L.vy, L.vy, L.vx * L.vx))) + L.vb;
                            const float u1 = fmaxf(xo_u01f(xo_next(L.xs0, L.xs1, L.xs2, L.xs3)), 1e-7f);
                            const float u2 = xo_u01f(xo_next(L.xs0, L.xs1, L.xs2, L.xs3));
                            const float u3 = fmaxf(xo_u01f(xo_next(L.xs0, L.xs1, L.xs2, L.xs3)), 1e-7f);
                            const float u4 = xo_u01f(xo_next(L.xs0, L.xs1, L.xs2, L.xs3));
                            const float r1 = sqrtf(-2.0f * log_f(u1)), r2 = sqrtf(-2.0f * log_f(u3));
                            const float gx = L.vx - L.sgas * (r1 * cos_rev(u2));
                            const float gy = L.vy - L.sgas * (r1 * sin_rev(u2));
                            const float gz = L.vz - L.sgas * (r2 * cos_rev(u4));
                            const float gl = sqrtf(fmaf(gz, gz, fmaf(gy, gy, gx * gx)));
                            if (xo_u01f(xo_next(L.xs0, L.xs1, L.xs2, L.xs3)) * bound < gl) {
                                const float cth = 2.0f * xo_u01f(xo_next(L.xs0, L.xs1, L.xs2, L.xs3)) - 1.0f;
                                const float a3 = xo_u01f(xo_next(L.xs0, L.xs1, L.xs2, L.xs3));
                                const float sn = sqrtf(fmaxf(0.0f, 1.0f - cth * cth));
                                const float mg = L.mr * gl;
                                L.vx = fmaf(mg, sn * cos_rev(a3), fmaf(-L.mr, gx, L.vx));
                                L.vy = fmaf(mg, sn * sin_rev(a3), fmaf(-L.mr, gy, L.vy));
                                L.vz = fmaf(mg, cth, fmaf(-L.mr, gz, L.vz));
                                L.ncoll += 1;
                            }
                        } else {
                            const float vm = sqrtf(fmaf(L.vz, L.vz, fmaf(L.vy, L.vy, L.vx * L.vx)));
                            if (vm > L.floorv) {
                                const float cth = 2.0f * xo_u01f(xo_next(L.xs0, L.xs1, L.xs2, L.xs3)) - 1.0f;
                                const float a3 = xo_u01f(xo_next(L.xs0, L.xs1, L.xs2, L.xs3));
                                const float sn = sqrtf(fmaxf(0.0f, 1.0f - cth * cth));
                                const float mv = L.mr * vm;
                                L.vx = fmaf(mv, sn * cos_rev(a3), L.vx * L.omr);
                                L.vy = fmaf(mv, sn * sin_rev(a3), L.vy * L.omr);
                                L.vz = fmaf(mv, cth, L.vz * L.omr);
                                L.ncoll += 1;
                            }
                            const float mfp = -log_f(fmaxf(xo_u01f(xo_next(L.xs0, L.xs1, L.xs2, L.xs3)), 1e-7f)) * L.mfpm;
                            const float vm2 = sqrtf(fmaf(L.vz, L.vz, fmaf(L.vy, L.vy, L.vx * L.vx)));
                            L.cd = countdown_f(floorf(mfp / fmaxf(vm2, 1e-3f) / dtf));
                        }
                        pending = false;
                    } else if (pending) {
                        collide_philox<LEGACY>(L, pcur, dtf);
                        pending = false;
                    }
                }
            }
            // with no lane left to step, the batch above has run every pending collision
            if constexpr (STAGED) {
                if (!__any(have && !stop)) break;
            }
        }
        // ---- write back stopped ions; carry unfinished ones to the next launch
        const bool carry = have && stop && status == 0;
        const uint64_t cm = __ballot(carry);
        if (cm) {
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)cm) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&A.ctr_out[CTR_COUNT], (uint32_t)__popcll(cm));
            base = __shfl(base, (int)leader);
            if (carry) {
                const uint32_t pos = base + (uint32_t)__popcll(cm & ((1ull << lane) - 1ull));
                A.list_out[pos] = L.id;
                if (A.bin_out) {
                    const uint32_t bin = sort_bin(A.B, A.sg, L.x, L.y, L.z);
                    A.bin_out[pos] = bin;
                    // the next re-sort's counts and ranks, one atomic per bin and wavefront
                    if (A.hist_out) bin_rank(A.hist_out, bin, pos, A.rank_out);
                }
            }
        }
        if (have && stop) {
            const uint32_t id = L.id;
            S.x[id] = L.x; S.y[id] = L.y; S.z[id] = L.z;
            S.fvx[id] = L.vx; S.fvy[id] = L.vy; S.fvz[id] = L.vz;
            S.steps[id] = L.st; S.ncoll[id] = L.ncoll;
            if constexpr (LEGACY) S.cd[id] = L.cd;
            if constexpr (COMPAT) {
                S.rng[4 * (size_t)id] = L.xs0; S.rng[4 * (size_t)id + 1] = L.xs1;
                S.rng[4 * (size_t)id + 2] = L.xs2; S.rng[4 * (size_t)id + 3] = L.xs3;
            }
            S.status[id] = (uint8_t)status;
            have = false;
        }
    }
}

// ------------------------------------------------- the tail worker, four lanes per ion
// Once few ions are left, a wavefront's step costs the latency of its own
// instruction chain (about 400 instructions), however few lanes hold ions. This
// worker spreads each ion's step over a quad of lanes, 16 ions per wavefront:
//  * lane d of the quad (d = 0, 1, 2; lane 3 repeats d = 2) holds component d
//    of the cell record (the RF, DC and AC polynomials of E_d: six float4, the
//    component-major layout of fp32_cells.h) plus the conductor bits, and
//    evaluates E_d;
//  * lanes 0 and 1 evaluate the RF and AC drive sines;
//  * at each refill period lane k draws Philox candidate block (step / 4) + k of
//    the ion, so the quad holds the candidate words of the period's steps;
//  * results reach the quad by DPP quad broadcasts; everything else (kick,
//    drift, cell location, collisions, classification) runs on all four lanes
//    with identical inputs.
// Every value an ion's step computes is the same operation on the same operands
// as in k_batched, so the two workers give the same results bit for bit. Philox
// streams only (the reference's streams draw sequentially per ion); one queue;
// the refill period is at most 8 steps (three candidate blocks cover it).
template <int S>
__device__ __forceinline__ uint32_t qbu(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, S * 0x55, 0xF, 0xF, false);
}
template <int S>
__device__ __forceinline__ float qbf(float v)
{
    return __uint_as_float(qbu<S>(__float_as_uint(v)));
}

// The quad's drive amplitudes at the lane's step: lane 0 evaluates the RF sine,
// lane 1 the AC sine, from the tone (w, p0, v) each picked when the ion arrived
// (RF: w_rf, phr, vrf; AC: w_ac, phi, vac).
struct Tone {
    double w, p0;
    float v;
};

__device__ __forceinline__ void phases4(Lane& L, const Tone& T)
{
    double p = fma(T.w, (double)L.st, T.p0);
    p = p - floor(p);
    const float a = T.v * sin_rev((float)p);
    L.a_rf = qbf<0>(a);
    L.a_ac = qbf<1>(a);
}

__device__ __forceinline__ void load_part(const float4* __restrict__ cells, int32_t c, uint32_t d, float4 Rc[6],
                                          uint32_t& bits)
{
    const float4* src = cells + (size_t)c * CELL_STRIDE_F4;
    // the conductor bits first: the step's crash test then waits for that load only
    bits = reinterpret_cast<const uint32_t*>(src + 18)[0];
#pragma unroll
    for (int q = 0; q < 6; ++q) Rc[q] = src[6 * d + q];
}

template <bool LEGACY>
__global__ __launch_bounds__(256) void k_tail4(WorkerArgs A)
{
    const BatchP& B = A.B;
    const IonState& S = A.S;
    const uint32_t lane = threadIdx.x & 63, role = lane & 3, dc = role < 3 ? role : 2u;
    const uint32_t count = A.ctr_in[CTR_COUNT];
    const double dt = B.dt;
    const float dtf = (float)B.dt;
    Lane L;
    Tone T{};
    float4 Rc[6];
    uint32_t bits = 0;
    bool have = false, drained = false, pending = false;
    // chk: the crash test of the quad's last step is outstanding; it needs the
    // conductor bits of the cell just entered and is resolved at the next
    // iteration (before that step's field, which needs the cell's record anyway),
    // so the record load's latency overlaps the candidate and collision work
    bool chk = false;
    uint32_t pcur = 0;
    L.cell = -1;
    for (;;) {
        // ---- refill idle quads from the live list (one atomic per wavefront)
        if (!drained) {
            const uint64_t idle = __ballot(role == 0 && !have && (lane >> 2) < A.lanes);
            if (idle) {
                const uint32_t nidle = (uint32_t)__popcll(idle);
                const uint32_t leader = (uint32_t)__ffsll((unsigned long long)idle) - 1;
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(&A.ctr_in[0], nidle);
                base = __shfl(base, (int)leader);
                if (base >= count) drained = true;
                uint32_t slot = 0xFFFFFFFFu;
                if ((idle >> lane) & 1ull) slot = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                slot = qbu<0>(slot);  // the quad's claim, made by its lane 0
                if (slot < count) {
                    const uint32_t id = A.list_in[slot];
                    const JobP& J = A.jobs[S.job[id]];
                    L.id = id;
                    L.ion = S.sid ? S.sid[id] : (uint32_t)(id - J.off);
                    L.phr = S.phr ? S.phr[id] : 0.0;
                    L.x = S.x[id]; L.y = S.y[id]; L.z = S.z[id]; L.phi = S.phi[id];
                    L.vx = S.fvx[id]; L.vy = S.fvy[id]; L.vz = S.fvz[id];
                    L.st = S.steps[id]; L.ncoll = S.ncoll[id];
                    if constexpr (LEGACY) L.cd = S.cd[id];
                    L.w_rf = J.w_rf; L.w_ac = J.w_ac;
                    L.kq = (float)J.kq; L.vrf = (float)J.v_rf; L.vac = (float)J.v_ac; L.sdc = (float)J.s_dc;
                    L.nsig = (float)J.nsig_dt; L.sgas = (float)J.s_gas; L.vb = (float)J.v_bound;
                    L.mr = (float)J.mr; L.omr = (float)J.one_m_mr; L.floorv = (float)J.floor_v;
                    L.mfpm = (float)J.mfp_mean;
                    L.k0 = J.key0; L.k1 = J.key1;
                    L.visit = 0;
                    const int32_t c = locate(B, L);
                    if (c != L.cell) { L.cell = c; load_part(A.cells, c, dc, Rc, bits); }
                    T = role == 1 ? Tone{L.w_ac, L.phi, L.vac} : Tone{L.w_rf, L.phr, L.vrf};
                    phases4(L, T);
                    have = true;
                }
            }
        }
        if (!__any(have)) break;
        // ---- the candidate words of this period: lane k draws block (st / 4) + k
        uint32_t cb0 = 0;
        u32x4 W{0u, 0u, 0u, 0u};
        if constexpr (!LEGACY) {
            if (have && L.nsig > 0.0f) {
                cb0 = L.st >> 2;
                W = philox(L.ion, cb0 + role, 0, D_CAND, L.k0, L.k1);
            }
        }
        bool stop = false;
        int status = 0;
        for (uint32_t r = 0; r < A.period; ++r) {
            if (chk) {
                chk = false;
                status = classify(B, L, bits);
                if (status != 0 || L.visit >= A.chunk) stop = true;
            }
            if (have && !stop && !pending) {  // the same in the four lanes of a quad
                const float pxy = L.tx * L.ty, pyz = L.ty * L.tz, pxz = L.tx * L.tz, pxyz = pxy * L.tz;
                const float er = poly8(Rc[0], Rc[1], L.tx, L.ty, L.tz, pxy, pyz, pxz, pxyz);
                const float ed = poly8(Rc[2], Rc[3], L.tx, L.ty, L.tz, pxy, pyz, pxz, pxyz);
                const float ea = poly8(Rc[4], Rc[5], L.tx, L.ty, L.tz, pxy, pyz, pxz, pxyz);
                const float e = fmaf(L.a_ac, ea, fmaf(L.sdc, ed, L.a_rf * er));
                L.vx = fmaf(L.kq, qbf<0>(e), L.vx);
                L.vy = fmaf(L.kq, qbf<1>(e), L.vy);
                L.vz = fmaf(L.kq, qbf<2>(e), L.vz);
                L.x = fma((double)L.vx, dt, L.x);
                L.y = fma((double)L.vy, dt, L.y);
                L.z = fma((double)L.vz, dt, L.z);
                const uint32_t cur = L.st;
                L.st += 1;
                const int32_t c = locate(B, L);
                if (c != L.cell) { L.cell = c; load_part(A.cells, c, dc, Rc, bits); }
                phases4(L, T);
                if constexpr (!LEGACY) {
                    if (L.nsig > 0.0f) {
                        const float bound = sqrtf(fmaf(L.vz, L.vz, fmaf(L.vy, L.vy, L.vx * L.vx))) + L.vb;
                        // word cur & 3 of block cur / 4, held by quad lane cur / 4 - cb0 (0..2)
                        const uint32_t mine = pick(W, cur & 3), src = (cur >> 2) - cb0;
                        const uint32_t w0 = qbu<0>(mine), w1 = qbu<1>(mine), w2 = qbu<2>(mine);
                        const uint32_t cw = src == 0 ? w0 : (src == 1 ? w1 : w2);
                        if (u01f(cw) < L.nsig * bound) { pending = true; pcur = cur; }
                    }
                } else {
                    L.cd -= 1;
                    if (L.cd == 0) { pending = true; pcur = cur; }
                }
                L.visit += 1;
                chk = true;
            }
            // ---- batched collisions of the pending ions (counted once per quad)
            const uint64_t pm = __ballot(pending && role == 0);
            if (pm) {
                const uint32_t npend = (uint32_t)__popcll(pm);
                const uint64_t am = __ballot(have && !stop && !pending);
                if (npend >= A.coll_batch || am == 0 || r + 1 == A.period) {
                    if (pending) {
                        collide_philox<LEGACY>(L, pcur, dtf);
                        pending = false;
                    }
                }
            }
        }
        if (chk) {  // the period's last step
            chk = false;
            status = classify(B, L, bits);
            if (status != 0 || L.visit >= A.chunk) stop = true;
        }
        // ---- write back stopped ions; carry unfinished ones to the next launch
        const bool carry = have && stop && status == 0;
        const uint64_t cm = __ballot(carry && role == 0);
        if (cm) {
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)cm) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&A.ctr_out[CTR_COUNT], (uint32_t)__popcll(cm));
            base = __shfl(base, (int)leader);
            if (carry && role == 0) A.list_out[base + (uint32_t)__popcll(cm & ((1ull << lane) - 1ull))] = L.id;
        }
        if (have && stop) {
            if (role == 0) {
                const uint32_t id = L.id;
                S.x[id] = L.x; S.y[id] = L.y; S.z[id] = L.z;
                S.fvx[id] = L.vx; S.fvy[id] = L.vy; S.fvz[id] = L.vz;
                S.steps[id] = L.st; S.ncoll[id] = L.ncoll;
                if constexpr (LEGACY) S.cd[id] = L.cd;
                S.status[id] = (uint8_t)status;
            }
            have = false;
        }
    }
}

// ------------------------------------------------- field probe
// The worker's field at given positions: the same cell location, record load
// and polynomial, with fixed drive amplitudes; crash = the conductor bit of the
// nearest corner node (the worker's crash test without its z window).
__global__ __launch_bounds__(256) void k_field_probe(const float4* __restrict__ cells, BatchP B, uint32_t n,
                                                     const double* __restrict__ px, const double* __restrict__ py,
                                                     const double* __restrict__ pz, float a_rf, float s_dc,
                                                     float a_ac, float* __restrict__ ex, float* __restrict__ ey,
                                                     float* __restrict__ ez, uint8_t* __restrict__ crash)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Lane L;
    L.x = px[i]; L.y = py[i]; L.z = pz[i];
    L.a_rf = a_rf; L.sdc = s_dc; L.a_ac = a_ac;
    const int32_t c = locate(B, L);
    float4 R[CELL_F4];
    load_cell(cells, c, R);
    float e[3];
    cell_field(R, L, e);
    ex[i] = e[0]; ey[i] = e[1]; ez[i] = e[2];
    crash[i] = (uint8_t)((__float_as_uint(R[18].x) >> L.corner) & 1u);
}

void launch_field_probe(const float4* cells, const BatchP& B, uint32_t n, const double* x, const double* y,
                        const double* z, float a_rf, float s_dc, float a_ac, float* ex, float* ey, float* ez,
                        uint8_t* crash, hipStream_t st)
{
    if (n) k_field_probe<<<(n + 255) / 256, 256, 0, st>>>(cells, B, n, x, y, z, a_rf, s_dc, a_ac, ex, ey, ez, crash);
}

// ------------------------------------------------- gather probe
// Random whole-record reads over the first `window` records: each lane reads
// `iters` records at hashed indices (splitmix-style 64-bit mixer of lane and
// iteration) and folds them into one float, so no load is dead code.
__global__ __launch_bounds__(256) void k_gather_probe(const float4* __restrict__ cells, uint64_t window, uint32_t iters,
                                                      float* __restrict__ out)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float acc = 0.0f;
    for (uint32_t it = 0; it < iters; ++it) {
        uint64_t h = t * 0x9E3779B97F4A7C15ull + (uint64_t)it * 0xD1B54A32D192ED03ull;
        h ^= h >> 31;
        h *= 0xBF58476D1CE4E5B9ull;
        h ^= h >> 29;
        const uint64_t c = h % window;
        float4 R[CELL_F4];
        load_cell(cells, (int32_t)c, R);
#pragma unroll
        for (int q = 0; q < CELL_F4; ++q) acc += (R[q].x + R[q].y) + (R[q].z + R[q].w);
    }
    out[t] = acc;
}

void launch_gather_probe(const float4* cells, uint64_t window, uint32_t iters, float* out, uint32_t grid,
                         uint32_t block, hipStream_t st)
{
    k_gather_probe<<<grid, block, 0, st>>>(cells, window, iters, out);
}

// ------------------------------------------------- re-sort by grid cell
// Counting sort of the live list by coarse cell bin (z-major, x fastest):
// k_bin counts (or the worker launch that wrote the list has counted and
// ranked the ions it carried), the three scan kernels turn counts into
// offsets, k_scatter places each ion. Order inside a bin follows the atomics;
// per-ion results do not depend on list order.
__global__ void k_bin(IonState S, BatchP B, const uint32_t* __restrict__ list, const uint32_t* __restrict__ ctr,
                      uint32_t* __restrict__ bin_of, uint32_t* __restrict__ rank_of, uint32_t* __restrict__ hist,
                      SortGeom G)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ctr[CTR_COUNT]) return;
    const uint32_t id = list[i];
    const uint32_t bin = sort_bin(B, G, S.x[id], S.y[id], S.z[id]);
    bin_of[i] = bin;
    bin_rank(hist, bin, i, rank_of);
}

// Exclusive scan of hist in place: blocks of 1024 entries (256 threads x 4).
__global__ __launch_bounds__(256) void k_scan_blocks(uint32_t* __restrict__ hist, uint32_t* __restrict__ bsum, uint32_t n)
{
    __shared__ uint32_t wsum[4];
    const uint32_t t = threadIdx.x, base = blockIdx.x * 1024 + t * 4;
    uint32_t v[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) { v[k] = base + k < n ? hist[base + k] : 0u; s += v[k]; }
    const uint32_t lane = t & 63, w = t >> 6;
    uint32_t incl = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o);
        if (lane >= (uint32_t)o) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t woff = 0;
    for (uint32_t k = 0; k < w; ++k) woff += wsum[k];
    uint32_t run = woff + incl - s;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) hist[base + k] = run;
        run += v[k];
    }
    if (t == 255) bsum[blockIdx.x] = woff + incl;
}

// Exclusive scan of the block sums (at most 1024) by one block.
__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t* __restrict__ bsum, uint32_t nb)
{
    __shared__ uint32_t wsum[16];
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t s = t < nb ? bsum[t] : 0u;
    uint32_t incl = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o);
        if (lane >= (uint32_t)o) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t woff = 0;
    for (uint32_t k = 0; k < w; ++k) woff += wsum[k];
    if (t < nb) bsum[t] = woff + incl - s;
}

__global__ void k_scan_add(uint32_t* __restrict__ hist, const uint32_t* __restrict__ bsum, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) hist[i] += bsum[i >> 10];
}

__global__ void k_scatter(const uint32_t* __restrict__ list, const uint32_t* __restrict__ ctr,
                          const uint32_t* __restrict__ bin_of, const uint32_t* __restrict__ rank_of,
                          const uint32_t* __restrict__ offs, uint32_t* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ctr[CTR_COUNT]) return;
    out[offs[bin_of[i]] + rank_of[i]] = list[i];
}

static inline uint32_t nblk(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

void launch_fill_hist(uint32_t* hist, uint32_t nb, hipStream_t st)
{
    k_fill_hist<<<nblk(nb, 256), 256, 0, st>>>(hist, nb);
}

void launch_sort_place(const uint32_t* list, const uint32_t* ctr, uint32_t* out, const uint32_t* bin_of,
                       const uint32_t* rank_of, uint32_t* hist, uint32_t* bsum, uint32_t nb, uint32_t nmax, hipStream_t st)
{
    const uint32_t nblocks = nblk(nb, 1024);
    k_scan_blocks<<<nblocks, 256, 0, st>>>(hist, bsum, nb);
    k_scan_sums<<<1, 1024, 0, st>>>(bsum, nblocks);
    k_scan_add<<<nblk(nb, 256), 256, 0, st>>>(hist, bsum, nb);
    if (nmax) k_scatter<<<nblk(nmax, 256), 256, 0, st>>>(list, ctr, bin_of, rank_of, hist, out);
}

void launch_sort(const IonState& S, const BatchP& B, const uint32_t* list, const uint32_t* ctr, uint32_t* out,
                 uint32_t* bin_of, uint32_t* rank_of, uint32_t* hist, uint32_t* bsum, const SortGeom& G,
                 uint32_t nmax, bool ranked, hipStream_t st)
{
    const uint32_t nb = G.nbins;
    if (!ranked) {
        launch_fill_hist(hist, nb, st);
        if (nmax) k_bin<<<nblk(nmax, 256), 256, 0, st>>>(S, B, list, ctr, bin_of, rank_of, hist, G);
    }
    launch_sort_place(list, ctr, out, bin_of, rank_of, hist, bsum, nb, nmax, st);
}

void launch_build_cells(float4* cells, const double* rf, const double* dc, const double* ac, const uint8_t* mask,
                        int nx, int ny, int nz, hipStream_t st)
{
    const uint64_t nc = (uint64_t)(nx - 1) * (ny - 1) * (nz - 1);
    k_build_cells<<<nblk(nc, 128), 128, 0, st>>>(cells, rf, dc, ac, mask, nx, ny, nz);
}

void launch_iota(uint32_t* list, uint32_t n, hipStream_t st)
{
    if (n) k_iota<<<nblk(n, 256), 256, 0, st>>>(list, n);
}

void launch_prepare(const WorkerArgs& a, hipStream_t st) { k_prepare<<<1, 64, 0, st>>>(a.ctr_in, a.ctr_out); }

void launch_batched(const WorkerArgs& a, uint32_t grid, uint32_t block, hipStream_t st)
{
    if (a.B.compat) {
        if (a.B.legacy_coll) k_batched<true, true, false><<<grid, block, 0, st>>>(a);
        else k_batched<false, true, false><<<grid, block, 0, st>>>(a);
    } else {
        if (a.B.legacy_coll) k_batched<true, false, false><<<grid, block, 0, st>>>(a);
        else k_batched<false, false, false><<<grid, block, 0, st>>>(a);
    }
}

void launch_tail4(const WorkerArgs& a, uint32_t grid, hipStream_t st)
{
    if (a.B.legacy_coll) k_tail4<true><<<grid, 256, 0, st>>>(a);
    else k_tail4<false><<<grid, 256, 0, st>>>(a);
}

void launch_staged(const WorkerArgs& a, uint32_t grid, hipStream_t st)
{
    if (a.B.legacy_coll) k_batched<true, false, true><<<grid, 256, TILE_LDS_BYTES, st>>>(a);
    else k_batched<false, false, true><<<grid, 256, TILE_LDS_BYTES, st>>>(a);
}

}  // namespace mq
