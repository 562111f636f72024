// The legacy Java potentials, regenerated on the GPU: the ejection sketch's
// float32 electrode rasterization and its float32 lexicographic Gauss-Seidel
// sweeps with the check-point stop rule, bit-identical to running the sketch's
// loops one node after the other.
//
// Reference (reference/MQ_MZ_ejection_singleQ/): GEOMETRY.pde:11-47, 54-318,
// 362-414, 463-467, 512-550 (rasterization); POTENTIAL.pde:31-84 (sweeps and
// stop rule); IO.pde:180-226 (called with 0.001 % and 6000, export);
// MQ_MZ_ejection_singleQ.pde:27-73, 118, 152-162 (dimensions);
// CONVERSIONS.pde:12-15 (check point). Fields::from_theirs
// (simulation/chsim/src/channel.rs:187-202) is the consumer.
//
// Float semantics: the sketch computes in Java float (IEEE single, source
// order, no fused operations; Processing makes every decimal literal a float
// literal). This file is compiled with -ffp-contract=off, divisions and
// square roots are the correctly rounded ones, and denormals are kept.
//
// Sweeps as wavefronts. The sequential sweep visits z outer, x, y inner, so
// when node (i, j, k) is updated in sweep t its lower neighbours (i-1, j-1,
// k-1) already hold sweep t and its upper neighbours still hold sweep t-1.
// Step tau updates, in place, every free node with i + j + k + 2t = tau for
// its own t < T: all nodes of one parity of i + j + k, each in its own sweep.
// Its lower neighbours were written at step tau - 1 (sweep t) and its upper
// neighbours at tau - 1 too (sweep t - 1, since (s + 1) + 2(t - 1) = tau - 1),
// and nothing writes them at tau (they have the other parity). So every
// update reads exactly the values the sequential sweep reads. The stop rule
// needs the check point's value after each sweep; a first pass logs it and
// finds the sweep count T, a second pass reruns from the start state with that
// T when it is below the cap (the first pass has by then run nodes below the
// check point beyond T).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/mq_legacy.h"

void mq_set_last_error(const std::string& msg);  // engine.cpp

namespace {

struct LgErr : std::runtime_error {
    int code;
    LgErr(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define LGCHK(expr)                                                                                      \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            throw LgErr(-3, std::string("HIP error ") + hipGetErrorString(e_) + " at " #expr);           \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) { LGCHK(hipMalloc(&p, bytes < 256 ? 256 : bytes)); }
    ~DevBuf() { if (p) (void)hipFree(p); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    template <class T> T* as() const { return (T*)p; }
};

struct LgStream {
    hipStream_t s = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    LgStream()
    {
        LGCHK(hipStreamCreate(&s));
        LGCHK(hipEventCreate(&a));
        LGCHK(hipEventCreate(&b));
    }
    ~LgStream()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        if (s) (void)hipStreamDestroy(s);
    }
};

constexpr int TB = 256;

// ------------------------------------------------------------ geometry

// The sketch's globals at scl nodes per mm (MQ_MZ_ejection_singleQ.pde:27-73, 152-162).
struct JGeo {
    float scl, step, hx, hy, qstart, q1, gap, plate, q2, dh, de, spacing, zdev;
    int nx, ny, nz;
};

JGeo jgeo_of(float scl)
{
    JGeo g;
    g.scl = scl;
    const float xdev = 16.0f, ydev = 16.0f;
    g.qstart = 3.0f;
    g.q1 = 12.0f;
    g.gap = 1.0f;
    g.plate = 1.6f;
    g.q2 = 3.0f;
    g.dh = 7.0f;
    g.de = 6.35f;
    g.spacing = 0.323f * g.de;
    g.zdev = g.qstart + g.q1 + g.gap + g.plate + g.gap + g.q2 + g.gap + g.plate;  // float, left to right
    g.hx = xdev / 2.0f;
    g.hy = ydev / 2.0f;
    g.nx = (int)(xdev * scl);
    g.ny = (int)(ydev * scl);
    g.nz = (int)(g.zdev * scl);
    g.step = 1.0f / scl;
    return g;
}

void check_point(const JGeo& g, int32_t c[3])
{
    const float xc = g.hx + 1.0f, yc = g.hy + 0.5f, zc = g.qstart + g.q1 - 1.0f;
    c[0] = (int)(xc * g.scl);
    c[1] = (int)(yc * g.scl);
    c[2] = (int)(zc * g.scl);
}

__device__ int layer_of(const JGeo& g, float z)
{
    const float l1 = 0.0f + g.qstart, l2 = l1 + g.q1, l3 = l2 + g.gap, l4 = l3 + g.plate;
    const float l5 = l4 + g.gap, l6 = l5 + g.q2, l7 = l6 + g.gap, l8 = l7 + g.plate;
    if (0.0f < z && z <= l1) return 1;
    if (l1 < z && z <= l2) return 2;
    if (l2 < z && z <= l3) return 3;
    if (l3 < z && z <= l4) return 4;
    if (l4 < z && z <= l5) return 5;
    if (l5 < z && z <= l6) return 6;
    if (l6 < z && z <= l7) return 7;
    if (l7 < z && z <= l8) return 8;
    return 0;
}

// Rod of the quadrant of (x, y), its centre shifted by sq in y toward the
// x axis of the box; true if (x, y) is strictly inside diameter d.
__device__ bool in_rod(const JGeo& g, float x, float y, float sq, float d)
{
    const float c = g.spacing / 2.0f + g.de / 2.0f;
    float cx, cy;
    if (x < g.hx && y < g.hy) { cx = g.hx - c; cy = g.hy - c + sq; }
    else if (x > g.hx && y < g.hy) { cx = g.hx + c; cy = g.hy - c + sq; }
    else if (x < g.hx && y > g.hy) { cx = g.hx - c; cy = g.hy + c - sq; }
    else if (x > g.hx && y > g.hy) { cx = g.hx + c; cy = g.hy + c - sq; }
    else return false;
    const float dx = x - cx, dy = y - cy;
    return sqrtf(dx * dx + dy * dy) < d / 2.0f;
}

__global__ void k_legacy_raster(JGeo g, int basis, float* phi, uint8_t* el)
{
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    const size_t id = (size_t)blockIdx.x * TB + threadIdx.x;
    if (id >= n) return;
    const int i = (int)(id % g.nx), j = (int)((id / g.nx) % g.ny), k = (int)(id / ((size_t)g.nx * g.ny));
    const float x = (float)i * g.step, y = (float)j * g.step, z = (float)k * g.step;
    const int layer = layer_of(g, z);
    bool e = false;
    float v = 0.0f;
    if (layer == 2 || layer == 6) {
        // withinSqueezedQuadElectrode, also in Q2 (the squeeze formula extends past Q1)
        e = in_rod(g, x, y, 0.02f * (g.qstart + g.q1 - z), g.de);
    } else if (layer == 3 || layer == 5 || layer == 7) {
        e = in_rod(g, x, y, 0.0f, 4.6f);  // withinQuadElectrode(x, y, 4.6): the stems
    } else if (layer == 4) {
        const float X = g.hx - x, Y = g.hy - y, r = g.dh / 2.0f;
        e = !((X * X + Y * Y) < r * r);  // the plate outside the hole
    } else if (layer == 8) {
        e = true;
    }
    if (e) {
        const bool up = layer <= 3;  // Q1 and its stems
        if (basis == MQ_LEGACY_RF) {
            const bool neg = (x <= g.hx && y <= g.hy) || (x >= g.hx && y >= g.hy);  // SIGN
            v = (layer == 4 || layer == 8) ? 0.0f : (neg ? -1.0f : 1.0f);
        } else if (basis == MQ_LEGACY_DC) {
            v = layer == 4 ? 1.0f : up ? 0.0f : -1.0f;  // plate +1, everything downstream -1
        } else {
            v = (up && layer != 1) ? (x >= g.hx ? 1.0f : -1.0f) : 0.0f;  // right rods +1, left -1, upstream only
        }
    }
    phi[id] = v;
    el[id] = e ? 1 : 0;
}

// ------------------------------------------------------------ sweeps

// Red-black split storage for the sweeps: colour c = (i + j + k) & 1 keeps its
// nodes of row (j, k) at (k ny + j) hx + (i >> 1), hx = (nx + 1) / 2. A step
// updates one colour from the other's values only (a node's new value does not
// read its old one), so it reads one array and writes the other, both
// contiguous along a row: half the bytes of the interleaved layout.
struct SweepGrid {
    int nx, ny, nz, hx;
    int jgroups;       // groups of LEG_R interior rows along j
    uint32_t rows;     // row groups: jgroups (nz - 2)
    int ch_color;      // check point: colour and split index
    size_t ch_idx;
};

__host__ __device__ inline size_t split_idx(int hx, int ny, int i, int j, int k)
{
    return ((size_t)k * ny + j) * hx + (size_t)(i >> 1);
}

__global__ void k_legacy_split(const float* __restrict__ phi, const uint8_t* __restrict__ el, int nx, int ny,
                               int nz, int hx, float* c0, float* c1, uint8_t* e0, uint8_t* e1)
{
    const size_t n = (size_t)nx * ny * nz;
    const size_t id = (size_t)blockIdx.x * TB + threadIdx.x;
    if (id >= n) return;
    const int i = (int)(id % nx), j = (int)((id / nx) % ny), k = (int)(id / ((size_t)nx * ny));
    const size_t q = split_idx(hx, ny, i, j, k);
    if ((i + j + k) & 1) { c1[q] = phi[id]; e1[q] = el[id]; }
    else { c0[q] = phi[id]; e0[q] = el[id]; }
}

__global__ void k_legacy_merge(const float* __restrict__ c0, const float* __restrict__ c1, int nx, int ny, int nz,
                               int hx, float* phi)
{
    const size_t n = (size_t)nx * ny * nz;
    const size_t id = (size_t)blockIdx.x * TB + threadIdx.x;
    if (id >= n) return;
    const int i = (int)(id % nx), j = (int)((id / nx) % ny), k = (int)(id / ((size_t)nx * ny));
    const size_t q = split_idx(hx, ny, i, j, k);
    phi[id] = ((i + j + k) & 1) ? c1[q] : c0[q];
}

// Step tau of the wavefront: colour tau & 1, the free interior nodes with
// i + j + k + 2t = tau and t < T, from the other colour's values. A thread
// takes column m of LEG_R consecutive rows j at one k and issues all their
// loads before any store (the arrays do not alias). Workgroups are dealt to
// the 8 XCDs round robin (workgroup b runs on XCD b mod 8); the launch has a
// multiple of 8 workgroups and workgroup b takes slab (b mod 8) of the z
// range, so the neighbour rows a workgroup reads were mostly fetched into the
// same XCD's L2 by its neighbours. log_ch[t] receives the check point's value
// after sweep t.
constexpr int LEG_R = 4;
constexpr int N_XCD = 8;

__global__ void __launch_bounds__(TB) k_legacy_step(SweepGrid G, const float* __restrict__ other,
                                                     float* __restrict__ own, const uint8_t* __restrict__ el_own,
                                                     int tau, int T, float* __restrict__ log_ch)
{
    const uint32_t per_xcd = gridDim.x / N_XCD;
    const uint32_t wg = (blockIdx.x % N_XCD) * per_xcd + blockIdx.x / N_XCD;
    const uint32_t id = wg * TB + threadIdx.x;
    const uint32_t m = id % (uint32_t)G.hx, g = id / (uint32_t)G.hx;
    if (g >= G.rows) return;
    const int j0 = 1 + LEG_R * (int)(g % (uint32_t)G.jgroups), k = 1 + (int)(g / (uint32_t)G.jgroups);
    const size_t sy = (size_t)G.hx, sz = (size_t)G.ny * G.hx;
    float v[LEG_R];
    bool upd[LEG_R];
    int tt[LEG_R];
    size_t oo[LEG_R];
#pragma unroll
    for (int r = 0; r < LEG_R; ++r) {
        const int j = j0 + r;
        const int p = (tau + j + k) & 1;  // parity of i in this colour
        const int i = 2 * (int)m + p;
        const int s = i + j + k;
        const int t = (tau - s) >> 1;
        const bool ok = j <= G.ny - 2 && i >= 1 && i <= G.nx - 2 && s <= tau && t < T;
        const size_t row = ((size_t)k * G.ny + (ok ? j : 1)) * G.hx, o = row + (ok ? m : 1);
        // (i-1) >> 1 = m - 1 + p, (i+1) >> 1 = m + p; the j and k neighbours share i >> 1 = m
        const size_t a = ok ? row + m - 1 + p : o;
        v[r] = (other[a] + other[row + (ok ? m + p : 1)] + other[o - sy] + other[o + sy] + other[o - sz] +
                other[o + sz]) / 6.0f;
        upd[r] = ok && el_own[o] == 0;
        tt[r] = t;
        oo[r] = o;
    }
#pragma unroll
    for (int r = 0; r < LEG_R; ++r) {
        if (!upd[r]) continue;
        own[oo[r]] = v[r];
        if ((tau & 1) == G.ch_color && oo[r] == G.ch_idx) log_ch[tt[r]] = v[r];
    }
}

// The stop rule's change at the check point, in float.
float dif_of(float prev, float now) { return std::fabs(100.0f * ((prev + 1e-10f) - now) / (prev + 1e-10f)); }

struct PinnedLog {
    float* p = nullptr;
    explicit PinnedLog(size_t n) { LGCHK(hipHostMalloc((void**)&p, (n ? n : 1) * 4, hipHostMallocDefault)); }
    ~PinnedLog() { if (p) (void)hipHostFree(p); }
};

// calculatePotential's sweeps on the device buffer phi (start state), in place.
mq_legacy_stats run_sweeps(int nx, int ny, int nz, float* phi, const uint8_t* el, const int32_t check[3],
                           float precision, uint32_t max_iter, hipStream_t s, hipEvent_t ea, hipEvent_t eb)
{
    if (nx < 3 || ny < 3 || nz < 3) throw LgErr(-1, "grid needs at least 3 nodes per axis");
    if ((double)nx * ny * nz > 2.0e9) throw LgErr(-1, "grid too large");
    for (int d = 0; d < 3; ++d)
        if (check[d] < 0 || check[d] >= (d == 0 ? nx : d == 1 ? ny : nz)) throw LgErr(-1, "check point outside the grid");
    if (max_iter > (1u << 28)) throw LgErr(-1, "max_iter too large");
    const size_t N = (size_t)nx * ny * nz;
    SweepGrid G;
    G.nx = nx;
    G.ny = ny;
    G.nz = nz;
    G.hx = (nx + 1) / 2;
    G.jgroups = (ny - 2 + LEG_R - 1) / LEG_R;
    G.rows = (uint32_t)G.jgroups * (uint32_t)(nz - 2);
    G.ch_color = (check[0] + check[1] + check[2]) & 1;
    G.ch_idx = split_idx(G.hx, ny, check[0], check[1], check[2]);
    if ((uint64_t)G.hx * G.rows > 0xFFFFFFFFull - TB) throw LgErr(-1, "grid too large");
    const unsigned blocks =
        (unsigned)((((uint64_t)G.hx * G.rows + TB - 1) / TB + N_XCD - 1) / N_XCD * N_XCD);  // multiple of 8
    const size_t NS = (size_t)nz * ny * G.hx;  // one colour array

    mq_legacy_stats st{};
    st.check[0] = check[0];
    st.check[1] = check[1];
    st.check[2] = check[2];

    std::vector<uint8_t> h_el(N);
    LGCHK(hipMemcpyAsync(h_el.data(), el, N, hipMemcpyDeviceToHost, s));
    const size_t nch = (size_t)check[0] + (size_t)nx * ((size_t)check[1] + (size_t)ny * check[2]);
    float c0 = 0.0f;
    LGCHK(hipMemcpyAsync(&c0, phi + nch, 4, hipMemcpyDeviceToHost, s));
    LGCHK(hipStreamSynchronize(s));
    uint64_t nfree = 0;
    for (int k = 1; k < nz - 1; ++k)
        for (int j = 1; j < ny - 1; ++j)
            for (int i = 1; i < nx - 1; ++i) nfree += h_el[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)] == 0;
    const bool ch_fixed = h_el[nch] || check[0] == 0 || check[0] == nx - 1 || check[1] == 0 || check[1] == ny - 1 ||
                          check[2] == 0 || check[2] == nz - 1;

    const int s_lo = 3, s_hi = (nx - 2) + (ny - 2) + (nz - 2);
    const int s_ch = check[0] + check[1] + check[2];
    DevBuf col(2 * NS * 4), ecol(2 * NS), log((size_t)(max_iter > 0 ? max_iter : 1) * 4);
    float* cc[2] = {col.as<float>(), col.as<float>() + NS};
    uint8_t* ec[2] = {ecol.as<uint8_t>(), ecol.as<uint8_t>() + NS};
    const unsigned nblk = (unsigned)((N + TB - 1) / TB);
    auto split = [&]() {
        LGCHK(hipMemsetAsync(col.p, 0, 2 * NS * 4, s));
        LGCHK(hipMemsetAsync(ecol.p, 0, 2 * NS, s));
        k_legacy_split<<<nblk, TB, 0, s>>>(phi, el, nx, ny, nz, G.hx, cc[0], cc[1], ec[0], ec[1]);
        LGCHK(hipGetLastError());
    };
    auto run_pass = [&](int T, int from, int to) {
        for (int tau = from; tau <= to; ++tau)
            k_legacy_step<<<blocks, TB, 0, s>>>(G, cc[(tau & 1) ^ 1], cc[tau & 1], ec[tau & 1], tau, T,
                                                log.as<float>());
        LGCHK(hipGetLastError());
    };

    // the stop rule over the check point's values after each sweep
    float dif = 100.0f, prev = 1.0f;
    uint32_t nit = 0;
    auto feed = [&](float now) {
        dif = dif_of(prev, now);
        prev = now;
        ++nit;
    };
    LGCHK(hipEventRecord(ea, s));
    split();
    uint32_t T = 0;
    if (max_iter > 0 && ch_fixed) {
        // the check point never changes: its start value decides the rule
        while (nit < max_iter && dif > precision) feed(c0);
        T = nit;
        run_pass((int)T, s_lo, s_hi + 2 * ((int)T - 1));
    } else if (max_iter > 0) {
        // pass one: run the wavefronts, read the check point's log one chunk behind the launches
        PinnedLog h_log(max_iter);
        const int chunk = 256;
        const int tau_end = s_hi + 2 * ((int)max_iter - 1);
        int tau = s_lo;
        int pending = -1;  // sweeps of the check point whose log copy is in flight
        hipEvent_t ev = nullptr;
        LGCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        bool decided = false;
        try {
            while (!decided) {
                int have = -1;
                if (tau <= tau_end) {
                    const int upto = std::min(tau_end, tau + chunk - 1);
                    run_pass((int)max_iter, tau, upto);
                    tau = upto + 1;
                    // sweeps t of the check point complete: s_ch + 2t <= tau - 1
                    have = tau - 1 >= s_ch ? std::min<int>((int)max_iter, (tau - 1 - s_ch) / 2 + 1) : 0;
                }
                if (pending >= 0) {
                    LGCHK(hipEventSynchronize(ev));
                    while (nit < (uint32_t)pending && dif > precision) feed(h_log.p[nit]);
                    if (!(nit < max_iter && dif > precision)) decided = true;
                    pending = -1;
                }
                if (decided) break;
                if (have < 0) throw LgErr(-9, "sweep log incomplete");  // every sweep was launched and read
                if (have > (int)nit) {
                    LGCHK(hipMemcpyAsync(h_log.p + nit, log.as<float>() + nit, (size_t)(have - (int)nit) * 4,
                                         hipMemcpyDeviceToHost, s));
                    LGCHK(hipEventRecord(ev, s));
                    pending = have;
                }
            }
        } catch (...) {
            (void)hipEventDestroy(ev);
            throw;
        }
        (void)hipEventDestroy(ev);
        T = nit;
        if (T < max_iter) {
            // the nodes below the check point ran past T: rerun from the start state with exactly T sweeps
            split();
            run_pass((int)T, s_lo, s_hi + 2 * ((int)T - 1));
        } else {
            // the cap: every node stops at max_iter sweeps by itself; finish the drain
            run_pass((int)max_iter, tau, tau_end);
        }
    }
    k_legacy_merge<<<nblk, TB, 0, s>>>(cc[0], cc[1], nx, ny, nz, G.hx, phi);
    LGCHK(hipGetLastError());
    LGCHK(hipEventRecord(eb, s));
    LGCHK(hipStreamSynchronize(s));
    float ms = 0.0f;
    LGCHK(hipEventElapsedTime(&ms, ea, eb));
    st.ms = ms;
    st.sweeps = T;
    st.dif_pct = dif;
    st.node_updates = nfree * T;
    return st;
}

int fail(const std::exception& e, int code)
{
    mq_set_last_error(e.what());
    return code;
}

void check_opts(const mq_legacy_opts* o)
{
    if (!o) throw LgErr(-1, "NULL options");
    if (!(o->scl >= 0.25f && o->scl <= 40.0f)) throw LgErr(-1, "scl must be in [0.25, 40] nodes per mm");
    if (!(o->precision_pct >= 0.0f)) throw LgErr(-1, "precision_pct must be >= 0");
}

void check_basis(int32_t basis)
{
    if (basis < MQ_LEGACY_RF || basis > MQ_LEGACY_AC) throw LgErr(-1, "basis must be MQ_LEGACY_RF, _DC or _AC");
}

// Rasterize + sweeps of one basis, left on the device.
mq_legacy_stats potential_dev(const mq_legacy_opts& o, int basis, float* d_phi, uint8_t* d_el, LgStream& S)
{
    const JGeo g = jgeo_of(o.scl);
    const size_t N = (size_t)g.nx * g.ny * g.nz;
    k_legacy_raster<<<(unsigned)((N + TB - 1) / TB), TB, 0, S.s>>>(g, basis, d_phi, d_el);
    LGCHK(hipGetLastError());
    int32_t c[3];
    check_point(g, c);
    return run_sweeps(g.nx, g.ny, g.nz, d_phi, d_el, c, o.precision_pct, o.max_iter, S.s, S.a, S.b);
}

void write_file(const std::string& path, const void* data, size_t bytes)
{
    const std::string tmp = path + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) throw LgErr(-4, "cannot write " + tmp);
    const size_t w = std::fwrite(data, 1, bytes, f);
    const int c = std::fclose(f);
    if (w != bytes || c != 0) throw LgErr(-4, "short write to " + tmp);
    if (std::rename(tmp.c_str(), path.c_str()) != 0) throw LgErr(-4, "cannot rename " + tmp);
}

std::vector<uint8_t> read_file(const std::string& path, size_t bytes)
{
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw LgErr(-4, "cannot open " + path);
    std::vector<uint8_t> v(bytes + 1);
    const size_t r = std::fread(v.data(), 1, bytes + 1, f);
    std::fclose(f);
    if (r != bytes) throw LgErr(-4, path + ": expected " + std::to_string(bytes) + " bytes");
    v.resize(bytes);
    return v;
}

}  // namespace

extern "C" {

void mq_legacy_default(mq_legacy_opts* o)
{
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->scl = 10.0f;
    o->precision_pct = 0.001f;
    o->max_iter = 6000;
}

int mq_legacy_dims(const mq_legacy_opts* o, int32_t* nx, int32_t* ny, int32_t* nz, int32_t check[3])
{
    try {
        check_opts(o);
        const JGeo g = jgeo_of(o->scl);
        if (nx) *nx = g.nx;
        if (ny) *ny = g.ny;
        if (nz) *nz = g.nz;
        if (check) check_point(g, check);
        return 0;
    } catch (const LgErr& e) { return fail(e, e.code); }
}

int mq_legacy_rasterize(const mq_legacy_opts* o, int32_t basis, float* phi, uint8_t* electrode)
{
    try {
        check_opts(o);
        check_basis(basis);
        if (!phi || !electrode) throw LgErr(-1, "NULL output");
        const JGeo g = jgeo_of(o->scl);
        const size_t N = (size_t)g.nx * g.ny * g.nz;
        LgStream S;
        DevBuf dp(N * 4), de(N);
        k_legacy_raster<<<(unsigned)((N + TB - 1) / TB), TB, 0, S.s>>>(g, basis, dp.as<float>(), de.as<uint8_t>());
        LGCHK(hipGetLastError());
        LGCHK(hipMemcpyAsync(phi, dp.p, N * 4, hipMemcpyDeviceToHost, S.s));
        LGCHK(hipMemcpyAsync(electrode, de.p, N, hipMemcpyDeviceToHost, S.s));
        LGCHK(hipStreamSynchronize(S.s));
        return 0;
    } catch (const LgErr& e) { return fail(e, e.code); } catch (const std::exception& e) { return fail(e, -9); }
}

int mq_legacy_sweeps(int32_t nx, int32_t ny, int32_t nz, float* phi, const uint8_t* electrode, const int32_t check[3],
                     float precision_pct, uint32_t max_iter, mq_legacy_stats* stats)
{
    try {
        if (!phi || !electrode || !check) throw LgErr(-1, "NULL argument");
        if (nx < 3 || ny < 3 || nz < 3) throw LgErr(-1, "grid needs at least 3 nodes per axis");
        if (!(precision_pct >= 0.0f)) throw LgErr(-1, "precision_pct must be >= 0");
        const size_t N = (size_t)nx * ny * nz;
        LgStream S;
        DevBuf dp(N * 4), de(N);
        LGCHK(hipMemcpyAsync(dp.p, phi, N * 4, hipMemcpyHostToDevice, S.s));
        LGCHK(hipMemcpyAsync(de.p, electrode, N, hipMemcpyHostToDevice, S.s));
        const mq_legacy_stats st = run_sweeps(nx, ny, nz, dp.as<float>(), de.as<uint8_t>(), check, precision_pct,
                                              max_iter, S.s, S.a, S.b);
        LGCHK(hipMemcpyAsync(phi, dp.p, N * 4, hipMemcpyDeviceToHost, S.s));
        LGCHK(hipStreamSynchronize(S.s));
        if (stats) *stats = st;
        return 0;
    } catch (const LgErr& e) { return fail(e, e.code); } catch (const std::exception& e) { return fail(e, -9); }
}

int mq_legacy_potential(const mq_legacy_opts* o, int32_t basis, float* phi, uint8_t* electrode,
                        mq_legacy_stats* stats)
{
    try {
        check_opts(o);
        check_basis(basis);
        if (!phi) throw LgErr(-1, "NULL output");
        const JGeo g = jgeo_of(o->scl);
        const size_t N = (size_t)g.nx * g.ny * g.nz;
        LgStream S;
        DevBuf dp(N * 4), de(N);
        const mq_legacy_stats st = potential_dev(*o, basis, dp.as<float>(), de.as<uint8_t>(), S);
        LGCHK(hipMemcpyAsync(phi, dp.p, N * 4, hipMemcpyDeviceToHost, S.s));
        if (electrode) LGCHK(hipMemcpyAsync(electrode, de.p, N, hipMemcpyDeviceToHost, S.s));
        LGCHK(hipStreamSynchronize(S.s));
        if (stats) *stats = st;
        return 0;
    } catch (const LgErr& e) { return fail(e, e.code); } catch (const std::exception& e) { return fail(e, -9); }
}

int mq_legacy_export(const mq_legacy_opts* o, const char* dir, mq_legacy_stats stats[3])
{
    try {
        check_opts(o);
        if (!dir) throw LgErr(-1, "NULL directory");
        mkdir(dir, 0755);
        const JGeo g = jgeo_of(o->scl);
        const size_t N = (size_t)g.nx * g.ny * g.nz;
        LgStream S;
        DevBuf dp(N * 4), de(N);
        std::vector<float> h(N);
        std::vector<uint8_t> m(N);
        const char* names[3] = {"RF", "DC", "AC"};
        for (int b = 0; b < 3; ++b) {
            const mq_legacy_stats st = potential_dev(*o, b, dp.as<float>(), de.as<uint8_t>(), S);
            LGCHK(hipMemcpyAsync(h.data(), dp.p, N * 4, hipMemcpyDeviceToHost, S.s));
            LGCHK(hipMemcpyAsync(m.data(), de.p, N, hipMemcpyDeviceToHost, S.s));
            LGCHK(hipStreamSynchronize(S.s));
            if (stats) stats[b] = st;
            const std::string base = std::string(dir) + "/eject_" + names[b];
            write_file(base + ".f32", h.data(), N * 4);
            if (b != MQ_LEGACY_DC) write_file(base + "_electrode.u8", m.data(), N);
        }
        return 0;
    } catch (const LgErr& e) { return fail(e, e.code); } catch (const std::exception& e) { return fail(e, -9); }
}

int mq_fields_from_theirs(const char* dir, mq_channel_fields* out)
{
    try {
        if (!dir || !out) throw LgErr(-1, "NULL argument");
        const int nx = 160, ny = 160, nz = 242;
        const double h = 1e-4;
        const size_t N = (size_t)nx * ny * nz;
        const std::string d(dir);
        const char* names[3] = {"RF", "DC", "AC"};
        std::vector<double> phi[3];
        for (int b = 0; b < 3; ++b) {
            const std::vector<uint8_t> raw = read_file(d + "/eject_" + names[b] + ".f32", N * 4);
            phi[b].resize(N);
            for (size_t n = 0; n < N; ++n) {
                float v;
                std::memcpy(&v, raw.data() + 4 * n, 4);
                phi[b][n] = (double)v;
            }
        }
        const std::vector<uint8_t> m_rf = read_file(d + "/eject_RF_electrode.u8", N);
        const std::vector<uint8_t> m_ac = read_file(d + "/eject_AC_electrode.u8", N);
        double* e_out[3] = {out->rf, out->dc[1], out->ac};
        for (int b = 0; b < 3; ++b)
            if (e_out[b] && mq_efield(phi[b].data(), nx, ny, nz, h, e_out[b]) != 0) return -3;
        for (int s : {0, 2, 3})
            if (out->dc[s]) std::memset(out->dc[s], 0, N * 24);
        double* p_out[6] = {phi[0].data(), phi[2].data(), nullptr, phi[1].data(), nullptr, nullptr};
        for (int b = 0; b < 6; ++b) {
            if (!out->phi[b]) continue;
            if (p_out[b]) std::memcpy(out->phi[b], p_out[b], N * 8);
            else std::memset(out->phi[b], 0, N * 8);
        }
        if (out->mask)
            for (size_t n = 0; n < N; ++n) out->mask[n] = (uint8_t)(m_rf[n] | m_ac[n]);
        for (int b = 0; b < 6; ++b) out->stats[b] = mq_solve_stats{0, 0.0, 1, 0.0};
        return 0;
    } catch (const LgErr& e) { return fail(e, e.code); } catch (const std::exception& e) { return fail(e, -9); }
}

}  // extern "C"
