// Basis-field grids of the channel geometry on the GPU: electrode
// rasterization, one Laplace solve per basis (FP64 conjugate gradients on the
// 7-point system), E = -grad(phi) by central differences and the mask.
//
// Reference: Fields::build / solve_basis / efield / Geom::classify
// (simulation/chsim/src/channel.rs:18-217) and the problem fieldsolve::solve
// defines (fieldsolve/src/lib.rs:104-166, mg.rs:962-965).
//
// Layout: node-major i + nx (j + ny k) for every grid; E as 3 doubles per
// node. All buffers live on the device for the whole build; the host gets E,
// the optional potentials and the mask at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/mq_fieldsolve.h"

void mq_set_last_error(const std::string& msg);  // engine.cpp
std::string mq_rust_debug_f64(double x);         // scan.cpp

namespace {

struct FsErr : std::runtime_error {
    int code;
    FsErr(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define FSCHK(expr)                                                                                      \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            throw FsErr(-3, std::string("HIP error ") + hipGetErrorString(e_) + " at " #expr);           \
    } while (0)

struct Dev {
    void* p = nullptr;
    explicit Dev(size_t bytes) { FSCHK(hipMalloc(&p, std::max<size_t>(bytes, 256))); }
    ~Dev() { if (p) (void)hipFree(p); }
    Dev(const Dev&) = delete;
    Dev& operator=(const Dev&) = delete;
    template <class T> T* as() const { return (T*)p; }
};

struct Stream {
    hipStream_t s = nullptr;
    Stream() { FSCHK(hipStreamCreate(&s)); }
    ~Stream() { if (s) (void)hipStreamDestroy(s); }
};

constexpr int TB = 256;     // threads per block of the node kernels
constexpr int RB = 1024;    // threads of the single-block reductions

// ------------------------------------------------------------ geometry

__host__ __device__ inline void layers_of(const mq_geom& g, double l[8])
{
    l[0] = g.q1_start;
    l[1] = l[0] + g.q1_len;
    l[2] = l[1] + g.gap;
    l[3] = l[2] + g.plate;
    l[4] = l[3] + g.gap;
    l[5] = l[4] + g.q2_len;
    l[6] = l[5] + g.gap;
    l[7] = l[6] + g.plate;
}

__host__ __device__ inline int nz_of(const mq_geom& g)
{
    double l[8];
    layers_of(g, l);
    return (int)(l[7] / g.h_mm + 1e-6);
}

// Rod centre of quadrant q for squeeze sq (the three tilt modes), plus the
// Q1 misalignment; true if (x, y) is strictly inside diameter d (+ growth).
__device__ bool rod_contains(const mq_geom& g, int q, double sq, double d, bool q1, double x, double y, double z,
                             double a, double c)
{
    double cx, cy;
    if (g.tilt_mode == 0) {
        cx = (q == 0 || q == 2) ? a - c : a + c;
        cy = (q < 2) ? a - c + sq : a + c - sq;
    } else if (g.tilt_mode == 2) {
        const double s = sq / sqrt(2.0);
        cx = (q == 0 || q == 2) ? a - c + s : a + c - s;
        cy = (q < 2) ? a - c + s : a + c - s;
    } else {
        cx = q == 0 ? a - c : q == 1 ? a + c : q == 2 ? a - c + sq : a + c - sq;
        cy = (q < 2) ? a - c : a + c;
    }
    if (q1) {
        const double zm = g.q1_start + g.q1_len / 2.0;
        cx += g.mis[q][0] + g.mis[q][2] * (z - zm);
        cy += g.mis[q][1] + g.mis[q][3] * (z - zm);
    }
    const double dx = x - cx, dy = y - cy;
    return sqrt(dx * dx + dy * dy) < d / 2.0 + g.grow * g.h_mm;
}

__device__ int classify(const mq_geom& g, int i, int j, int k)
{
    const double x = (double)i * g.h_mm, y = (double)j * g.h_mm, z = (double)k * g.h_mm;
    const double a = (double)(g.n_xy / 2) * g.h_mm;
    const double c = g.coef * g.de / 2.0 + g.de / 2.0;
    double l[8];
    layers_of(g, l);
    const double eps = 1e-9;
    const int q = (x < a && y < a) ? 0 : (x > a && y < a) ? 1 : (x < a && y > a) ? 2 : (x > a && y > a) ? 3 : -1;
    if (z <= l[0] + eps) return MQ_ELEC_NONE;
    if (z <= l[1] + eps)
        return (q >= 0 && rod_contains(g, q, g.alpha * (l[1] - z), g.de, true, x, y, z, a, c)) ? MQ_ELEC_Q1 + q : 0;
    if (z <= l[2] + eps)
        return (q >= 0 && rod_contains(g, q, 0.0, g.stem_d, false, x, y, z, a, c)) ? MQ_ELEC_STEM3 + q : 0;
    if (z <= l[3] + eps) {
        const double hr = g.hole_d / 2.0 - g.grow * g.h_mm;
        return ((x - a) * (x - a) + (y - a) * (y - a) < hr * hr) ? MQ_ELEC_NONE : MQ_ELEC_PLATE;
    }
    if (z <= l[4] + eps)
        return (q >= 0 && rod_contains(g, q, 0.0, g.stem_d, false, x, y, z, a, c)) ? MQ_ELEC_STEM5 + q : 0;
    if (z <= l[5] + eps) {
        const double sq = g.q2_flare_bug ? g.alpha * (l[1] - z) : 0.0;
        return (q >= 0 && rod_contains(g, q, sq, g.de, false, x, y, z, a, c)) ? MQ_ELEC_Q2 + q : 0;
    }
    if (z <= l[6] + eps)
        return (q >= 0 && rod_contains(g, q, 0.0, g.stem_d, false, x, y, z, a, c)) ? MQ_ELEC_STEM7 + q : 0;
    if (z <= l[7] + eps) return MQ_ELEC_SLAB;
    return MQ_ELEC_NONE;
}

// Potential of an electrode in one basis (channel.rs:110-119).
__device__ double basis_value(int basis, int code, int ac_mode)
{
    if (code == MQ_ELEC_NONE) return 0.0;
    const bool upstream = code < MQ_ELEC_PLATE;  // Q1, Stem3
    const bool rod = code != MQ_ELEC_PLATE && code != MQ_ELEC_SLAB;
    const int q = code < MQ_ELEC_STEM3 ? code - MQ_ELEC_Q1
                : code < MQ_ELEC_PLATE ? code - MQ_ELEC_STEM3
                : code < MQ_ELEC_Q2 ? code - MQ_ELEC_STEM5
                : code < MQ_ELEC_STEM7 ? code - MQ_ELEC_Q2 : code - MQ_ELEC_STEM7;
    switch (basis) {
        case MQ_BASIS_RF: return rod ? ((q == 0 || q == 3) ? -1.0 : 1.0) : 0.0;
        case MQ_BASIS_AC: return (!upstream || (ac_mode == 1 && q >= 2)) ? 0.0 : ((q == 0 || q == 2) ? -1.0 : 1.0);
        case MQ_BASIS_DC_Q1: return upstream ? 1.0 : 0.0;
        case MQ_BASIS_DC_PLATE: return code == MQ_ELEC_PLATE ? 1.0 : 0.0;
        case MQ_BASIS_DC_DOWN: return code >= MQ_ELEC_STEM5 ? 1.0 : 0.0;
        default: return 0.0;
    }
}

__global__ void k_classify(mq_geom g, int nx, int ny, int nz, uint8_t* codes)
{
    const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)nx * ny * nz;
    if (id >= n) return;
    const int i = (int)(id % nx), j = (int)((id / nx) % ny), k = (int)(id / ((size_t)nx * ny));
    codes[id] = (uint8_t)classify(g, i, j, k);
}

__global__ void k_basis(const uint8_t* codes, size_t n, int basis, int ac_mode, uint8_t* mask, double* fixed)
{
    const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int c = codes[id];
    mask[id] = c != MQ_ELEC_NONE;
    fixed[id] = basis_value(basis, c, ac_mode);
}

// ------------------------------------------------------------ Laplace solve
//
// Unknowns: the free nodes (not masked, not on a box face); every free node
// has all six neighbours. With A = 6 I - (sum over free neighbours), the
// system A u = b with b = sum over fixed neighbours of their values is
// symmetric positive definite; CG runs on it with its scalars kept on the
// device (no host round trip per iteration) and block sums reduced in a fixed
// order, so the iterate sequence is deterministic.

enum { SC_RR = 0, SC_ALPHA, SC_BETA, SC_RES, SC_N };

template <class T>
__device__ inline T block_sum(T v, T* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = TB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

__device__ inline double block_max(double v, double* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = TB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    return sh[0];
}

__device__ inline double nb_sum(const double* u, size_t id, size_t sy, size_t sz)
{
    return ((u[id - 1] + u[id + 1]) + (u[id - sy] + u[id + sy])) + (u[id - sz] + u[id + sz]);
}

__global__ void k_ls_setup(const uint8_t* mask, const double* fixed, double outer, int nx, int ny, int nz,
                           double* phi, uint8_t* fr, double* part_max)
{
    __shared__ double sh[TB];
    const size_t id = (size_t)blockIdx.x * TB + threadIdx.x;
    const size_t n = (size_t)nx * ny * nz;
    double mf = 0.0;
    if (id < n) {
        const int i = (int)(id % nx), j = (int)((id / nx) % ny), k = (int)(id / ((size_t)nx * ny));
        const bool face = i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1;
        if (mask[id]) {
            phi[id] = fixed[id];
            fr[id] = 0;
            mf = fabs(fixed[id]);
        } else if (face) {
            phi[id] = outer;
            fr[id] = 0;
        } else {
            phi[id] = 0.0;
            fr[id] = 1;
        }
    }
    const double m = block_max(mf, sh);
    if (threadIdx.x == 0) part_max[blockIdx.x] = m;
}

// r = true residual (sum of neighbours - 6 phi) on free nodes, p = r;
// block sums of r^2.
__global__ void k_ls_residual(const double* phi, const uint8_t* fr, int nx, int ny, size_t n, double* r, double* p,
                              double* part)
{
    __shared__ double sh[TB];
    const size_t id = (size_t)blockIdx.x * TB + threadIdx.x;
    double v = 0.0;
    if (id < n) {
        if (fr[id]) v = nb_sum(phi, id, (size_t)nx, (size_t)nx * ny) - 6.0 * phi[id];
        r[id] = v;
        p[id] = v;
    }
    const double s = block_sum(v * v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// q = A p on free nodes (p is zero elsewhere); block sums of p.q.
__global__ void k_ls_apply(const double* p, const uint8_t* fr, int nx, int ny, size_t n, double* q, double* part)
{
    __shared__ double sh[TB];
    const size_t id = (size_t)blockIdx.x * TB + threadIdx.x;
    double pq = 0.0;
    if (id < n) {
        double v = 0.0;
        if (fr[id]) v = 6.0 * p[id] - nb_sum(p, id, (size_t)nx, (size_t)nx * ny);
        q[id] = v;
        pq = p[id] * v;
    }
    const double s = block_sum(pq, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// phi += alpha p, r -= alpha q; block sums of r^2.
__global__ void k_ls_update(double* phi, const double* p, const double* q, double* r, const double* sc, size_t n,
                            double* part)
{
    __shared__ double sh[TB];
    const size_t id = (size_t)blockIdx.x * TB + threadIdx.x;
    const double alpha = sc[SC_ALPHA];
    double rr = 0.0;
    if (id < n) {
        const double pv = p[id];
        if (pv != 0.0) phi[id] = fma(alpha, pv, phi[id]);
        const double rv = fma(-alpha, q[id], r[id]);
        r[id] = rv;
        rr = rv * rv;
    }
    const double s = block_sum(rr, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_ls_direction(double* p, const double* r, const double* sc, size_t n)
{
    const size_t id = (size_t)blockIdx.x * TB + threadIdx.x;
    if (id < n) p[id] = fma(sc[SC_BETA], p[id], r[id]);
}

// Block maxima of |sum of neighbours - 6 phi| over free nodes.
__global__ void k_ls_truemax(const double* phi, const uint8_t* fr, int nx, int ny, size_t n, double* part)
{
    __shared__ double sh[TB];
    const size_t id = (size_t)blockIdx.x * TB + threadIdx.x;
    double v = 0.0;
    if (id < n && fr[id]) v = fabs(nb_sum(phi, id, (size_t)nx, (size_t)nx * ny) - 6.0 * phi[id]);
    const double m = block_max(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

enum { R_INIT = 0, R_ALPHA, R_BETA, R_MAX };

// One block, fixed summation order: thread t adds part[t], part[t + RB], ...
__global__ void __launch_bounds__(RB) k_ls_reduce(const double* part, int nb, double* sc, int mode)
{
    __shared__ double sh[RB];
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += RB) v = (mode == R_MAX) ? fmax(v, part[b]) : v + part[b];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = RB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            sh[threadIdx.x] = (mode == R_MAX) ? fmax(sh[threadIdx.x], sh[threadIdx.x + s])
                                              : sh[threadIdx.x] + sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double t = sh[0];
    if (mode == R_INIT) {
        sc[SC_RR] = t;
    } else if (mode == R_ALPHA) {
        sc[SC_ALPHA] = t > 0.0 ? sc[SC_RR] / t : 0.0;
    } else if (mode == R_BETA) {
        sc[SC_BETA] = sc[SC_RR] > 0.0 ? t / sc[SC_RR] : 0.0;
        sc[SC_RR] = t;
    } else {
        sc[SC_RES] = t;
    }
}

// E = -grad(phi), central differences, one-sided at the faces.
__global__ void k_efield(const double* phi, int nx, int ny, int nz, double h, double* e)
{
    const size_t id = (size_t)blockIdx.x * TB + threadIdx.x;
    const size_t n = (size_t)nx * ny * nz;
    if (id >= n) return;
    const int i = (int)(id % nx), j = (int)((id / nx) % ny), k = (int)(id / ((size_t)nx * ny));
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    const int im = i > 0 ? i - 1 : 0, ip = i + 1 < nx ? i + 1 : nx - 1;
    const int jm = j > 0 ? j - 1 : 0, jp = j + 1 < ny ? j + 1 : ny - 1;
    const int km = k > 0 ? k - 1 : 0, kp = k + 1 < nz ? k + 1 : nz - 1;
    auto at = [&](int a, int b, int c) { return phi[(size_t)a + sy * b + sz * c]; };
    e[3 * id + 0] = -(at(ip, j, k) - at(im, j, k)) / ((double)(ip - im) * h);
    e[3 * id + 1] = -(at(i, jp, k) - at(i, jm, k)) / ((double)(jp - jm) * h);
    e[3 * id + 2] = -(at(i, j, kp) - at(i, j, km)) / ((double)(kp - km) * h);
}

// ------------------------------------------------------------ host side

unsigned blocks_for(size_t n) { return (unsigned)((n + TB - 1) / TB); }

void check_dims(int nx, int ny, int nz)
{
    if (nx < 3 || ny < 3 || nz < 3) throw FsErr(-1, "grid needs at least 3 nodes per axis");
    if ((double)nx * ny * nz > 4.0e9) throw FsErr(-1, "grid too large");
}

void check_geom(const mq_geom* g)
{
    if (!g) throw FsErr(-1, "NULL geometry");
    if (!(g->h_mm > 0.0) || g->n_xy < 3) throw FsErr(-1, "geometry: h_mm must be > 0 and n_xy >= 3");
    if (!(g->de > 0.0) || !(g->stem_d > 0.0)) throw FsErr(-1, "geometry: rod and stem diameters must be > 0");
    if (g->tilt_mode < 0 || g->ac_mode < 0) throw FsErr(-1, "geometry: negative mode");
    const int nz = nz_of(*g);
    if (nz < 3) throw FsErr(-1, "geometry: fewer than 3 planes along z");
}

// Device-resident solver of one grid size; reused over the bases of a build.
struct Solver {
    int nx, ny, nz;
    size_t n;
    unsigned nb;
    Dev phi, r, p, q, fr, part, sc;
    hipStream_t st;
    Solver(int nx_, int ny_, int nz_, hipStream_t s)
        : nx(nx_), ny(ny_), nz(nz_), n((size_t)nx_ * ny_ * nz_), nb(blocks_for(n)), phi(n * 8), r(n * 8), p(n * 8),
          q(n * 8), fr(n), part((size_t)nb * 8), sc(SC_N * 8), st(s)
    {
    }

    double read_scalar(int k)
    {
        double v = 0.0;
        FSCHK(hipMemcpyAsync(&v, sc.as<double>() + k, 8, hipMemcpyDeviceToHost, st));
        FSCHK(hipStreamSynchronize(st));
        return v;
    }

    // mask/fixed on the device; leaves the solution in phi.
    mq_solve_stats solve(const uint8_t* d_mask, const double* d_fixed, double outer, double tol, uint32_t max_iter)
    {
        if (max_iter == 0) max_iter = 20000;
        hipEvent_t e0, e1;
        FSCHK(hipEventCreate(&e0));
        FSCHK(hipEventCreate(&e1));
        FSCHK(hipEventRecord(e0, st));
        double* P = phi.as<double>();
        double* part_ = part.as<double>();
        double* sc_ = sc.as<double>();
        k_ls_setup<<<nb, TB, 0, st>>>(d_mask, d_fixed, outer, nx, ny, nz, P, fr.as<uint8_t>(), part_);
        k_ls_reduce<<<1, RB, 0, st>>>(part_, (int)nb, sc_, R_MAX);
        FSCHK(hipGetLastError());
        double scale = std::max(read_scalar(SC_RES), std::fabs(outer));
        if (!(scale > 0.0)) scale = 1.0;
        k_ls_residual<<<nb, TB, 0, st>>>(P, fr.as<uint8_t>(), nx, ny, n, r.as<double>(), p.as<double>(), part_);
        k_ls_reduce<<<1, RB, 0, st>>>(part_, (int)nb, sc_, R_INIT);
        mq_solve_stats s{};
        const uint32_t check = 32;
        uint32_t it = 0;
        double res = INFINITY;
        for (;;) {
            if (it % check == 0) {
                k_ls_truemax<<<nb, TB, 0, st>>>(P, fr.as<uint8_t>(), nx, ny, n, part_);
                k_ls_reduce<<<1, RB, 0, st>>>(part_, (int)nb, sc_, R_MAX);
                FSCHK(hipGetLastError());
                res = read_scalar(SC_RES) / scale;
                if (res <= tol || it >= max_iter || !std::isfinite(res)) break;
            }
            k_ls_apply<<<nb, TB, 0, st>>>(p.as<double>(), fr.as<uint8_t>(), nx, ny, n, q.as<double>(), part_);
            k_ls_reduce<<<1, RB, 0, st>>>(part_, (int)nb, sc_, R_ALPHA);
            k_ls_update<<<nb, TB, 0, st>>>(P, p.as<double>(), q.as<double>(), r.as<double>(), sc_, n, part_);
            k_ls_reduce<<<1, RB, 0, st>>>(part_, (int)nb, sc_, R_BETA);
            k_ls_direction<<<nb, TB, 0, st>>>(p.as<double>(), r.as<double>(), sc_, n);
            ++it;
        }
        FSCHK(hipEventRecord(e1, st));
        FSCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        FSCHK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        s.iterations = it;
        s.residual = res;
        s.converged = res <= tol;
        s.ms = ms;
        return s;
    }
};

std::string geom_debug(const mq_geom& g)
{
    auto D = mq_rust_debug_f64;
    std::string s = "Geom { de: " + D(g.de) + ", coef: " + D(g.coef) + ", q1_start: " + D(g.q1_start) +
                    ", q1_len: " + D(g.q1_len) + ", alpha: " + D(g.alpha) + ", gap: " + D(g.gap) +
                    ", plate: " + D(g.plate) + ", hole_d: " + D(g.hole_d) + ", q2_len: " + D(g.q2_len) +
                    ", stem_d: " + D(g.stem_d) + ", q2_flare_bug: " + (g.q2_flare_bug ? "true" : "false") +
                    ", grow: " + D(g.grow) + ", tilt_mode: " + std::to_string(g.tilt_mode) +
                    ", ac_mode: " + std::to_string(g.ac_mode) + ", mis: [";
    for (int q = 0; q < 4; ++q) {
        s += q ? ", [" : "[";
        for (int k = 0; k < 4; ++k) s += (k ? ", " : "") + D(g.mis[q][k]);
        s += "]";
    }
    s += "], n_xy: " + std::to_string(g.n_xy) + ", h_mm: " + D(g.h_mm) + " }";
    return s;
}

const char* basis_name(int b)
{
    static const char* names[6] = {"Rf", "Ac", "DcQ1", "DcPlate", "DcDown", "DcBox"};
    return names[b];
}

std::string cache_file(const mq_geom& g, int basis)
{
    const std::string key = geom_debug(g) + basis_name(basis);
    uint64_t h = 1469598103934665603ull;
    for (unsigned char c : key) {
        h ^= (uint64_t)c;
        h *= 1099511628211ull;
    }
    char buf[32];
    std::snprintf(buf, sizeof buf, "ch_%016llx.bin", (unsigned long long)h);
    return buf;
}

bool read_cache(const std::string& path, size_t n, std::vector<double>& out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    f.seekg(0, std::ios::end);
    if ((size_t)f.tellg() != n * 8) return false;
    f.seekg(0);
    out.resize(n);
    static_assert(sizeof(double) == 8, "f64");
    f.read((char*)out.data(), (std::streamsize)(n * 8));  // little-endian host (x86-64)
    return (bool)f;
}

void write_cache(const std::string& dir, const std::string& path, const double* d, size_t n)
{
    ::mkdir(dir.c_str(), 0755);
    std::ofstream f(path, std::ios::binary);
    if (f) f.write((const char*)d, (std::streamsize)(n * 8));
}

int fail(const std::exception& e, int code)
{
    mq_set_last_error(e.what());
    return code;
}

}  // namespace

extern "C" {

void mq_geom_default(mq_geom* g)
{
    if (!g) return;
    std::memset(g, 0, sizeof *g);
    g->de = 6.35; g->coef = 0.323; g->q1_start = 3.0; g->q1_len = 12.0; g->alpha = 0.02; g->gap = 1.0;
    g->plate = 1.6; g->hole_d = 7.0; g->q2_len = 3.0; g->stem_d = 4.6; g->q2_flare_bug = 1; g->grow = 0.0;
    g->tilt_mode = 0; g->ac_mode = 0; g->n_xy = 160; g->h_mm = 0.1;
}

int mq_geom_dims(const mq_geom* g, int32_t* nx, int32_t* ny, int32_t* nz, double layers[8])
{
    try {
        check_geom(g);
        if (nx) *nx = g->n_xy;
        if (ny) *ny = g->n_xy;
        if (nz) *nz = nz_of(*g);
        if (layers) layers_of(*g, layers);
        return 0;
    } catch (const FsErr& e) { return fail(e, e.code); }
}

int mq_geom_fields_params(const mq_geom* g, mq_fields* f)
{
    try {
        check_geom(g);
        if (!f) throw FsErr(-1, "NULL fields");
        double l[8];
        layers_of(*g, l);
        f->nx = g->n_xy;
        f->ny = g->n_xy;
        f->nz = nz_of(*g);
        f->h = g->h_mm * 1e-3;
        const double a = (double)(g->n_xy / 2) * g->h_mm * 1e-3;
        f->axis_x = a;
        f->axis_y = a;
        f->z_inject = 1.0e-3;
        f->z_return = 0.5e-3;
        f->z_exit = l[4] * 1e-3;
        f->z_return_legacy = 1.0e-3;
        f->z_exit_legacy = 23.7e-3;
        f->crash_z0_legacy = l[0] * 1e-3;
        f->crash_z1_legacy = l[1] * 1e-3;
        return 0;
    } catch (const FsErr& e) { return fail(e, e.code); }
}

int mq_geom_classify(const mq_geom* g, uint8_t* codes)
{
    try {
        check_geom(g);
        if (!codes) throw FsErr(-1, "NULL output");
        const int nx = g->n_xy, ny = g->n_xy, nz = nz_of(*g);
        check_dims(nx, ny, nz);
        const size_t n = (size_t)nx * ny * nz;
        Stream s;
        Dev d(n);
        k_classify<<<blocks_for(n), TB, 0, s.s>>>(*g, nx, ny, nz, d.as<uint8_t>());
        FSCHK(hipGetLastError());
        FSCHK(hipMemcpyAsync(codes, d.p, n, hipMemcpyDeviceToHost, s.s));
        FSCHK(hipStreamSynchronize(s.s));
        return 0;
    } catch (const FsErr& e) { return fail(e, e.code); } catch (const std::exception& e) { return fail(e, -9); }
}

int mq_laplace_solve(int32_t nx, int32_t ny, int32_t nz, const uint8_t* mask, const double* fixed, double outer_bc,
                     double tol, uint32_t max_iter, double* phi, mq_solve_stats* stats)
{
    try {
        if (!mask || !fixed || !phi) throw FsErr(-1, "NULL argument");
        if (!(tol > 0.0)) throw FsErr(-1, "tol must be > 0");
        check_dims(nx, ny, nz);
        const size_t n = (size_t)nx * ny * nz;
        Stream s;
        Dev dm(n), dfx(n * 8);
        FSCHK(hipMemcpyAsync(dm.p, mask, n, hipMemcpyHostToDevice, s.s));
        FSCHK(hipMemcpyAsync(dfx.p, fixed, n * 8, hipMemcpyHostToDevice, s.s));
        Solver sv(nx, ny, nz, s.s);
        const mq_solve_stats st = sv.solve(dm.as<uint8_t>(), dfx.as<double>(), outer_bc, tol, max_iter);
        FSCHK(hipMemcpyAsync(phi, sv.phi.p, n * 8, hipMemcpyDeviceToHost, s.s));
        FSCHK(hipStreamSynchronize(s.s));
        if (stats) *stats = st;
        if (!st.converged) throw FsErr(-5, "Laplace solve did not reach the tolerance");
        return 0;
    } catch (const FsErr& e) { return fail(e, e.code); } catch (const std::exception& e) { return fail(e, -9); }
}

int mq_efield(const double* phi, int32_t nx, int32_t ny, int32_t nz, double h, double* e)
{
    try {
        if (!phi || !e) throw FsErr(-1, "NULL argument");
        if (!(h > 0.0)) throw FsErr(-1, "h must be > 0");
        if (nx < 2 || ny < 2 || nz < 2) throw FsErr(-1, "grid needs at least 2 nodes per axis");
        const size_t n = (size_t)nx * ny * nz;
        Stream s;
        Dev dp(n * 8), de(n * 24);
        FSCHK(hipMemcpyAsync(dp.p, phi, n * 8, hipMemcpyHostToDevice, s.s));
        k_efield<<<blocks_for(n), TB, 0, s.s>>>(dp.as<double>(), nx, ny, nz, h, de.as<double>());
        FSCHK(hipGetLastError());
        FSCHK(hipMemcpyAsync(e, de.p, n * 24, hipMemcpyDeviceToHost, s.s));
        FSCHK(hipStreamSynchronize(s.s));
        return 0;
    } catch (const FsErr& x) { return fail(x, x.code); } catch (const std::exception& x) { return fail(x, -9); }
}

int mq_fields_build(const mq_geom* g, double tol, const char* cache_dir, mq_channel_fields* out)
{
    try {
        check_geom(g);
        if (!out) throw FsErr(-1, "NULL output");
        if (!(tol > 0.0)) throw FsErr(-1, "tol must be > 0");
        const int nx = g->n_xy, ny = g->n_xy, nz = nz_of(*g);
        check_dims(nx, ny, nz);
        const size_t n = (size_t)nx * ny * nz;
        const double h = g->h_mm * 1e-3;
        Stream s;
        Dev codes(n), dm(n), dfx(n * 8), de(n * 24);
        k_classify<<<blocks_for(n), TB, 0, s.s>>>(*g, nx, ny, nz, codes.as<uint8_t>());
        FSCHK(hipGetLastError());
        Solver sv(nx, ny, nz, s.s);
        double* eout[6] = {out->rf, out->ac, out->dc[0], out->dc[1], out->dc[2], out->dc[3]};
        std::vector<double> cached;
        for (int b = 0; b < 6; ++b) {
            const double outer = b == MQ_BASIS_DC_BOX ? 1.0 : 0.0;
            k_basis<<<blocks_for(n), TB, 0, s.s>>>(codes.as<uint8_t>(), n, b, g->ac_mode, dm.as<uint8_t>(),
                                                  dfx.as<double>());
            FSCHK(hipGetLastError());
            std::string path;
            if (cache_dir) path = std::string(cache_dir) + "/" + cache_file(*g, b);
            if (cache_dir && read_cache(path, n, cached)) {
                FSCHK(hipMemcpyAsync(sv.phi.p, cached.data(), n * 8, hipMemcpyHostToDevice, s.s));
                out->stats[b] = mq_solve_stats{0, 0.0, 1, 0.0};
            } else {
                out->stats[b] = sv.solve(dm.as<uint8_t>(), dfx.as<double>(), outer, tol, 0);
                if (!out->stats[b].converged)
                    throw FsErr(-5, std::string("Laplace solve of basis ") + basis_name(b) + " did not converge");
                if (cache_dir) {
                    cached.resize(n);
                    FSCHK(hipMemcpyAsync(cached.data(), sv.phi.p, n * 8, hipMemcpyDeviceToHost, s.s));
                    FSCHK(hipStreamSynchronize(s.s));
                    write_cache(cache_dir, path, cached.data(), n);
                }
            }
            if (out->phi[b]) FSCHK(hipMemcpyAsync(out->phi[b], sv.phi.p, n * 8, hipMemcpyDeviceToHost, s.s));
            if (eout[b]) {
                k_efield<<<blocks_for(n), TB, 0, s.s>>>(sv.phi.as<double>(), nx, ny, nz, h, de.as<double>());
                FSCHK(hipGetLastError());
                FSCHK(hipMemcpyAsync(eout[b], de.p, n * 24, hipMemcpyDeviceToHost, s.s));
            }
            FSCHK(hipStreamSynchronize(s.s));
        }
        if (out->mask) {
            FSCHK(hipMemcpyAsync(out->mask, dm.p, n, hipMemcpyDeviceToHost, s.s));
            FSCHK(hipStreamSynchronize(s.s));
        }
        return 0;
    } catch (const FsErr& e) { return fail(e, e.code); } catch (const std::exception& e) { return fail(e, -9); }
}

int mq_field_cache_name(const mq_geom* g, int32_t basis, char* buf, size_t len)
{
    if (!g || !buf) { mq_set_last_error("NULL argument"); return -1; }
    if (basis < 0 || basis > 5) { mq_set_last_error("basis out of range"); return -1; }
    const std::string s = cache_file(*g, basis);
    if (s.size() + 1 > len) return (int)(s.size() + 1);
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return 0;
}

}  // extern "C"
