// engine.cpp -- host side of the MI355X batched ion-trajectory integrator and
// its C ABI (include/mq_engine.h).
//
// Mirrors chsim::channel::Engine (simulation/chsim/src/channel.rs:309-506) and
// the tiled launch loop (simulation/chsim/src/tiled.rs:87-218): the engine owns
// the device, one HIP stream, the uploaded basis grids, a cache of the last DC
// configuration (combined FP64 field + FP32 cell table) and pooled per-call
// buffers. Every call is synchronous; an engine is used from one thread.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mq_engine.h"
#include "host_util.h"
#include "kernels.h"
#include "mq_types.h"

using namespace mq;
using namespace mq::host;

namespace {

thread_local std::string g_err;

}  // namespace

// ---------------------------------------------------------- no-alloc scratch
// Device memory of mq_run_batch_ions_noalloc: the per-call buffers of a batch
// of at most MQ_NOALLOC_MAX_IONS ions and MQ_NOALLOC_MAX_JOBS jobs (ion state,
// live lists, re-sort bins, counters, job table, per-job accumulators) are
// carved from this array instead of being allocated. Every GPU holds its own
// copy, at its own address, until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns for the device (noalloc_ptr),
// never through the array's name. Calls on one GPU share the array, so calls
// that overlap on one GPU race.
namespace {

constexpr size_t NA_ALIGN = 256;
constexpr size_t na_round(size_t b) { return (b + NA_ALIGN - 1) / NA_ALIGN * NA_ALIGN; }
constexpr size_t NA_MAX_BINS = 1024u * 1024u;  // the re-sort's own limit
constexpr size_t NA_CTR_U32 = 2 * CTR_BLOCK;
constexpr size_t NA_IONS = MQ_NOALLOC_MAX_IONS, NA_JOBS = MQ_NOALLOC_MAX_JOBS;
constexpr size_t NA_BYTES = 7 * na_round(NA_IONS * 8)      // x y z phi vx vy vz
                            + 7 * na_round(NA_IONS * 4)    // fvx fvy fvz steps ncoll cd job
                            + na_round(NA_IONS)            // status
                            + 5 * na_round(NA_IONS * 4)    // three lists, bin_of, rank_of
                            + na_round(NA_CTR_U32 * 4)     // two counter blocks
                            + na_round(NA_MAX_BINS * 4)    // histogram / bin offsets
                            + na_round(1024 * 4)           // block sums of the scan
                            + na_round(NA_JOBS * sizeof(JobP))
                            + na_round(NA_JOBS * ACC * 8);  // accumulators

__device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mu;
void* g_na_ptr[NA_MAX_DEVICES] = {};

// Device address of `scratch` on `device` (the current device), looked up once
// per device.
void* noalloc_ptr(int device)
{
    if (device < 0 || device >= NA_MAX_DEVICES) throw Err(-1, "device index out of range for the no-alloc scratch");
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[device]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Err(-3, "hipGetSymbolAddress returned no address for the no-alloc scratch");
        g_na_ptr[device] = p;
    }
    return g_na_ptr[device];
}

// Sub-buffers of the scratch array, 256-byte aligned.
struct Carver {
    unsigned char* base;
    size_t off = 0;
    template <class T> T* take(size_t count)
    {
        const size_t b = na_round(count * sizeof(T));
        if (off + b > NA_BYTES) throw Err(-1, "no-alloc scratch overflow");
        T* p = reinterpret_cast<T*>(base + off);
        off += b;
        return p;
    }
};

}  // namespace

// Device buffers of one batch: pooled on the engine (grown on demand), or carved
// from the no-alloc scratch array.
struct CallBufs {
    JobP* jobs = nullptr;
    unsigned long long* acc = nullptr;
    uint32_t* list[3] = {nullptr, nullptr, nullptr};
    uint32_t* ctr = nullptr;
    uint32_t *bin_of = nullptr, *rank_of = nullptr, *hist = nullptr, *bsum = nullptr;
    IonState S{};
    bool noalloc = false;
};

// Steps per launch of the four-lane tail worker: each launch repacks the live
// ions into as few wavefronts as hold them.
constexpr uint32_t QUAD_CHUNK = 1024;

struct mq_engine {
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int nx = 0, ny = 0, nz = 0;
    double h = 0;
    mq_fields meta{};
    DevBuf rf, ac, dcb[4], mask;
    bool has_dc[4] = {false, false, false, false};
    // cache of the last DC configuration
    bool dc_valid = false;
    double dc_key[4] = {0, 0, 0, 0};
    DevBuf dcc;    // combined DC, FP64, 3N
    bool cells_valid = false;
    DevBuf cells;  // FP32 cell records
    // per-call pools
    DevBuf jobs, acc, list[3], ctr, bin_of, rank_of, hist, bsum;
    DevBuf sx, sy, sz, sphi, svx, svy, svz, sfx, sfy, sfz, ssteps, sncoll, scd, sjob, sstatus, srng;
    mq_tune tune{256, 8, 8, 256, 256, 1, 16, 0, 4, 0, 1, 1};
    // no-alloc entry point: this device's scratch array and host staging
    // reserved at creation for MQ_NOALLOC_MAX_JOBS jobs
    void* scratch_dev = nullptr;
    std::vector<JobP> jp_host;
    std::vector<unsigned long long> acc_host;
    // differentiable form: context of the last forward
    bool adj_valid = false;
    uint32_t adj_n = 0;
    AdjObj adj_obj{};
    BatchP adj_B{};
    DevBuf adj_mem, adj_cols, adj_part, adj_out;
    AdjBuf adj_buf{};
    AdjGrids adj_G{};
    // float cell records for the differentiable form, used when every value is
    // exactly a float: -1 not checked yet, 0 not exact, 1 exact (records built)
    int rfac32_state = -1, dcc32_state = -1;
    DevBuf adj_rec, exact_flag;
    // FP64 cell records of the plain path for the current combined DC
    DevBuf plain_rec;
    bool plain_rec_valid = false;
    std::vector<hipEvent_t> evs;
    mq_run_stats stats{};

    size_t nodes() const { return (size_t)nx * ny * nz; }

    void create(const mq_fields* f)
    {
        if (!f) throw Err(-1, "fields is NULL");
        if (f->nx < 2 || f->ny < 2 || f->nz < 2) throw Err(-1, "grid needs at least 2 nodes per axis");
        if (!(f->h > 0.0)) throw Err(-1, "grid spacing h must be positive");
        if (!f->rf || !f->ac) throw Err(-1, "RF and AC basis grids are required");
        if (3ull * (size_t)f->nx * f->ny * f->nz >= (1ull << 31)) throw Err(-1, "grid too large (3*N must be < 2^31)");
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        if (ndev < 1) throw Err(-2, "no HIP device");
        HIPCHK(hipGetDevice(&device));
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, device));
        cus = prop.multiProcessorCount;
        HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&ev0));
        HIPCHK(hipEventCreate(&ev1));
        nx = f->nx; ny = f->ny; nz = f->nz; h = f->h;
        meta = *f;
        meta.rf = meta.ac = nullptr;
        for (auto& p : meta.dc) p = nullptr;
        meta.mask = nullptr;
        const size_t n3 = 3 * nodes() * sizeof(double);
        rf.ensure(n3);
        HIPCHK(hipMemcpyAsync(rf.p, f->rf, n3, hipMemcpyHostToDevice, stream));
        ac.ensure(n3);
        HIPCHK(hipMemcpyAsync(ac.p, f->ac, n3, hipMemcpyHostToDevice, stream));
        for (int b = 0; b < 4; ++b) {
            if (!f->dc[b]) continue;
            has_dc[b] = true;
            dcb[b].ensure(n3);
            HIPCHK(hipMemcpyAsync(dcb[b].p, f->dc[b], n3, hipMemcpyHostToDevice, stream));
        }
        mask.ensure(nodes());
        if (f->mask) HIPCHK(hipMemcpyAsync(mask.p, f->mask, nodes(), hipMemcpyHostToDevice, stream));
        else HIPCHK(hipMemsetAsync(mask.p, 0, nodes(), stream));
        HIPCHK(hipStreamSynchronize(stream));
        scratch_dev = noalloc_ptr(device);
        jp_host.reserve(MQ_NOALLOC_MAX_JOBS);
        acc_host.reserve((size_t)MQ_NOALLOC_MAX_JOBS * ACC);
    }

    void destroy()
    {
        for (DevBuf* b : {&rf, &ac, &dcb[0], &dcb[1], &dcb[2], &dcb[3], &mask, &dcc, &cells, &jobs, &acc, &list[0],
                          &list[1], &list[2], &ctr, &bin_of, &rank_of, &hist, &bsum, &sx, &sy, &sz, &sphi, &svx, &svy, &svz, &sfx, &sfy, &sfz, &ssteps,
                          &sncoll, &scd, &sjob, &sstatus, &srng, &adj_mem, &adj_cols, &adj_part, &adj_out, &adj_rec,
                          &exact_flag})
            b->release();
        for (auto e : evs) (void)hipEventDestroy(e);
        evs.clear();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Combined DC field (Engine::dc_buffer, channel.rs:384-395) and, for the
    // batched path, its FP32 cell table (brick_buffer, tiled.rs:65-84).
    // noalloc: the tables must already have their full size (an earlier call on
    // this path built them); a new DC configuration is recomputed in place.
    void prepare_dc(const double dc[4], bool want_cells, bool noalloc = false)
    {
        const size_t nc = (size_t)(nx - 1) * (ny - 1) * (nz - 1);
        if (noalloc && (dcc.bytes < 3 * nodes() * sizeof(double) ||
                        (want_cells && cells.bytes < nc * CELL_STRIDE_F4 * sizeof(float4))))
            throw Err(-1, "the no-alloc entry point needs the engine's field tables: make one mq_run_batch call "
                          "on this path first");
        bool same = dc_valid;
        for (int b = 0; b < 4; ++b) same = same && dc_key[b] == dc[b];
        if (!same) {
            dcc.ensure(3 * nodes() * sizeof(double));
            const double* d[4];
            double v[4];
            for (int b = 0; b < 4; ++b) {
                d[b] = has_dc[b] ? dcb[b].as<double>() : nullptr;
                v[b] = dc[b];
                dc_key[b] = dc[b];
            }
            launch_combine_dc(dcc.as<double>(), d, v, (uint32_t)(3 * nodes()), stream);
            HIPCHK(hipGetLastError());
            dc_valid = true;
            dcc32_state = -1;
            plain_rec_valid = false;
            cells_valid = false;
        }
        if (want_cells && !cells_valid) {
            cells.ensure(nc * CELL_STRIDE_F4 * sizeof(float4));
            launch_build_cells(cells.as<float4>(), rf.as<double>(), dcc.as<double>(), ac.as<double>(),
                               mask.as<uint8_t>(), nx, ny, nz, stream);
            HIPCHK(hipGetLastError());
            cells_valid = true;
        }
    }

    IonState state_buffers(size_t n)
    {
        sx.ensure(n * 8); sy.ensure(n * 8); sz.ensure(n * 8); sphi.ensure(n * 8);
        svx.ensure(n * 8); svy.ensure(n * 8); svz.ensure(n * 8);
        sfx.ensure(n * 4); sfy.ensure(n * 4); sfz.ensure(n * 4);
        ssteps.ensure(n * 4); sncoll.ensure(n * 4); scd.ensure(n * 4); sjob.ensure(n * 4); sstatus.ensure(n);
        IonState S;
        S.x = sx.as<double>(); S.y = sy.as<double>(); S.z = sz.as<double>(); S.phi = sphi.as<double>();
        S.vx = svx.as<double>(); S.vy = svy.as<double>(); S.vz = svz.as<double>();
        S.fvx = sfx.as<float>(); S.fvy = sfy.as<float>(); S.fvz = sfz.as<float>();
        S.steps = ssteps.as<uint32_t>(); S.ncoll = sncoll.as<uint32_t>(); S.cd = scd.as<uint32_t>();
        S.job = sjob.as<uint32_t>(); S.status = sstatus.as<uint8_t>();
        return S;
    }

    // Device buffers of one batch of n ions and njobs jobs on `path`; compat adds
    // the xoshiro128** states of the reference's streams (S.rng, 4 words per ion).
    CallBufs call_bufs(size_t n, size_t njobs, bool noalloc, int path, bool compat)
    {
        CallBufs c;
        c.noalloc = noalloc;
        const bool batched = path == MQ_PATH_BATCHED;
        const bool resort = batched && tune.resort != 0;
        const uint32_t nbins = resort ? sort_geom().nbins : 0u;
        if (nbins > NA_MAX_BINS) throw Err(-1, "grid too large for the re-sort bins");
        const int nlists = batched ? (resort ? 3 : 2) : 0;
        if (!noalloc) {
            jobs.ensure(njobs * sizeof(JobP));
            c.jobs = jobs.as<JobP>();
            acc.ensure(njobs * ACC * sizeof(unsigned long long));
            c.acc = acc.as<unsigned long long>();
            c.S = state_buffers(std::max<size_t>(n, 1));
            if (compat) {
                srng.ensure(std::max<size_t>(n, 1) * 16);
                c.S.rng = srng.as<uint32_t>();
            }
            ctr.ensure(NA_CTR_U32 * 4);
            c.ctr = ctr.as<uint32_t>();
            for (int k = 0; k < nlists; ++k) {
                list[k].ensure(n * 4);
                c.list[k] = list[k].as<uint32_t>();
            }
            if (resort) {
                bin_of.ensure(n * 4);
                rank_of.ensure(n * 4);
                hist.ensure((size_t)nbins * 4);
                bsum.ensure(1024 * 4);
                c.bin_of = bin_of.as<uint32_t>(); c.rank_of = rank_of.as<uint32_t>();
                c.hist = hist.as<uint32_t>(); c.bsum = bsum.as<uint32_t>();
            }
            return c;
        }
        if (n > MQ_NOALLOC_MAX_IONS || njobs > MQ_NOALLOC_MAX_JOBS) throw Err(-1, "no-alloc batch beyond its maximum");
        if (compat) throw Err(-1, "the no-alloc entry point runs the Philox streams only (rng = 0)");
        Carver cv{static_cast<unsigned char*>(scratch_dev)};
        const size_t m = std::max<size_t>(n, 1);
        IonState& S = c.S;
        S.x = cv.take<double>(m); S.y = cv.take<double>(m); S.z = cv.take<double>(m); S.phi = cv.take<double>(m);
        S.vx = cv.take<double>(m); S.vy = cv.take<double>(m); S.vz = cv.take<double>(m);
        S.fvx = cv.take<float>(m); S.fvy = cv.take<float>(m); S.fvz = cv.take<float>(m);
        S.steps = cv.take<uint32_t>(m); S.ncoll = cv.take<uint32_t>(m); S.cd = cv.take<uint32_t>(m);
        S.job = cv.take<uint32_t>(m); S.status = cv.take<uint8_t>(m);
        S.rng = nullptr;
        for (int k = 0; k < nlists; ++k) c.list[k] = cv.take<uint32_t>(m);
        if (resort) {
            c.bin_of = cv.take<uint32_t>(m); c.rank_of = cv.take<uint32_t>(m);
            c.hist = cv.take<uint32_t>(nbins); c.bsum = cv.take<uint32_t>(1024);
        }
        c.ctr = cv.take<uint32_t>(NA_CTR_U32);
        c.jobs = cv.take<JobP>(std::max<size_t>(njobs, 1));
        c.acc = cv.take<unsigned long long>(std::max<size_t>(njobs, 1) * ACC);
        return c;
    }

    // The reference's injection and xoshiro128** seeds (MQ_RNG_COMPAT), on the
    // host as Engine::inject does it, uploaded into the ion state.
    void inject_compat(const IonState& S, const mq_job* jobs_h, size_t njobs, const mq_physics& ph, uint32_t n)
    {
        std::vector<double> x(n), y(n), z(n), vx(n), vy(n), vz(n), phi(n);
        std::vector<uint32_t> cd(n), rng(4 * (size_t)n), job(n), zero(n, 0u);
        std::vector<float> fv(n);
        std::vector<uint8_t> st(n, 0);
        uint64_t off = 0;
        for (size_t k = 0; k < njobs; ++k) {
            const uint64_t m = jobs_h[k].n;
            inject_reference(meta, jobs_h[k], ph, m, &x[off], &y[off], &z[off], &vx[off], &vy[off], &vz[off],
                             &phi[off], &cd[off], &rng[4 * off]);
            for (uint64_t i = 0; i < m; ++i) job[off + i] = (uint32_t)k;
            off += m;
        }
        if (!S.rng) throw Err(-1, "internal error: no random-stream state buffer for the reference's streams");
        auto up = [&](void* dst, const void* src, size_t bytes) {
            HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
        };
        up(S.x, x.data(), n * 8); up(S.y, y.data(), n * 8); up(S.z, z.data(), n * 8); up(S.phi, phi.data(), n * 8);
        up(S.vx, vx.data(), n * 8); up(S.vy, vy.data(), n * 8); up(S.vz, vz.data(), n * 8);
        up(S.cd, cd.data(), n * 4); up(S.rng, rng.data(), (size_t)n * 16); up(S.job, job.data(), n * 4);
        up(S.steps, zero.data(), n * 4); up(S.ncoll, zero.data(), n * 4); up(S.status, st.data(), n);
        float* fdst[3] = {S.fvx, S.fvy, S.fvz};
        const std::vector<double>* vsrc[3] = {&vx, &vy, &vz};
        for (int d = 0; d < 3; ++d) {
            for (uint32_t i = 0; i < n; ++i) fv[i] = (float)(*vsrc[d])[i];
            up(fdst[d], fv.data(), n * 4);
            HIPCHK(hipStreamSynchronize(stream));  // fv is reused
        }
        HIPCHK(hipStreamSynchronize(stream));
    }

    BatchP batch_params(const mq_physics& ph, uint32_t njobs) const
    {
        BatchP B{};
        B.dt = ph.dt;
        B.h = h;
        B.inv_h = 1.0 / h;
        B.lim0 = (double)(nx - 1) * h;
        B.lim1 = (double)(ny - 1) * h;
        B.lim2 = (double)(nz - 1) * h;
        B.z_ret = ph.legacy_criteria ? meta.z_return_legacy : meta.z_return;
        B.z_exit = ph.legacy_criteria ? meta.z_exit_legacy : meta.z_exit;
        B.cz0 = ph.legacy_criteria ? meta.crash_z0_legacy : -1e30;
        B.cz1 = ph.legacy_criteria ? meta.crash_z1_legacy : 1e30;
        B.nx = nx; B.ny = ny; B.nz = nz;
        B.nearest = ph.legacy_fields ? 1 : 0;
        B.legacy_coll = ph.legacy_collisions ? 1 : 0;
        B.compat = ph.rng == MQ_RNG_COMPAT ? 1 : 0;
        B.max_steps = max_steps_of(ph);
        B.njobs = njobs;
        return B;
    }

    // One batch: inject, run to completion on `path`, summarize. Returns the
    // device ion state (valid until the next call). noalloc: the per-call
    // buffers come from the scratch array, and nothing is allocated on the device
    // (the host staging vectors stay within the capacity reserved at creation).
    IonState run(const mq_job* jobs_h, size_t njobs, const double dc[4], const mq_physics& ph, int path,
                 std::vector<unsigned long long>& acc_h, size_t& n_total, bool noalloc = false)
    {
        auto t_wall = std::chrono::steady_clock::now();
        adj_valid = false;
        if (!jobs_h || njobs == 0) throw Err(-1, "batch has no jobs");
        if (!(ph.dt > 0.0)) throw Err(-1, "dt must be positive");
        if (path != MQ_PATH_BATCHED && path != MQ_PATH_PLAIN) throw Err(-1, "unknown path");
        if (ph.rng != MQ_RNG_PHILOX && ph.rng != MQ_RNG_COMPAT) throw Err(-1, "unknown random stream (rng)");
        if (noalloc) {
            if (ph.rng != MQ_RNG_PHILOX)
                throw Err(-1, "the no-alloc entry point runs the Philox streams only (rng = 0): the reference's "
                              "streams are injected on the host");
            if (njobs > MQ_NOALLOC_MAX_JOBS)
                throw Err(-1, "no-alloc batch has " + std::to_string(njobs) + " jobs, more than its maximum of " +
                                  std::to_string(MQ_NOALLOC_MAX_JOBS));
        }
        std::vector<JobP>& jp = jp_host;
        jp.resize(njobs);
        uint64_t off = 0;
        for (size_t k = 0; k < njobs; ++k) {
            jp[k] = job_params(jobs_h[k], ph, off);
            off += jobs_h[k].n;
        }
        if (off >= 0xFFFFFF00ull) throw Err(-1, "too many ions in one batch (limit 2^32 - 256)");
        if (noalloc && off > MQ_NOALLOC_MAX_IONS)
            throw Err(-1, "no-alloc batch has " + std::to_string(off) + " ions, more than its maximum of " +
                              std::to_string(MQ_NOALLOC_MAX_IONS));
        const uint32_t n = (uint32_t)off;
        n_total = n;
        prepare_dc(dc, path == MQ_PATH_BATCHED, noalloc);
        const CallBufs cb = call_bufs(n, njobs, noalloc, path, ph.rng == MQ_RNG_COMPAT);
        HIPCHK(hipMemcpyAsync(cb.jobs, jp.data(), njobs * sizeof(JobP), hipMemcpyHostToDevice, stream));
        launch_fill_u64(cb.acc, 0ull, (uint32_t)(njobs * ACC), stream);
        // every kernel of the batch, injection included, reads this one set of pointers
        const IonState& S = cb.S;
        InjP ip{};
        ip.axis_x = meta.axis_x; ip.axis_y = meta.axis_y; ip.z_inject = meta.z_inject;
        ip.temp_k = ph.temp_k; ip.r_inject_mm = ph.r_inject_mm; ip.inj_sigma_mm = ph.inj_sigma_mm; ip.dt = ph.dt;
        ip.model = ph.injection == MQ_INJ_GAUSSIAN ? 2 : (ph.legacy_injection ? 1 : 0);
        if (ph.rng == MQ_RNG_COMPAT) {
            if (n) inject_compat(S, jobs_h, njobs, ph, n);
        } else {
            launch_inject(S, cb.jobs, (uint32_t)njobs, ip, n, stream);
        }
        HIPCHK(hipGetLastError());
        const BatchP B = batch_params(ph, (uint32_t)njobs);
        stats = mq_run_stats{};
        stats.n_ions = n;
        float kms = 0.0f;
        if (n) {
            if (path == MQ_PATH_PLAIN) {
                const double2* rec = plain_records(noalloc);
                stats.plain_recs = rec ? 1u : 0u;
                run_plain(cb, B, n, kms, rec);
            } else {
                run_batched(cb, B, n, kms);
            }
        }
        launch_summarize(S, cb.acc, n, stream);
        HIPCHK(hipGetLastError());
        acc_h.resize(njobs * ACC);
        HIPCHK(hipMemcpyAsync(acc_h.data(), cb.acc, njobs * ACC * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              stream));
        HIPCHK(hipStreamSynchronize(stream));
        uint64_t steps = 0, unfinished = 0;
        for (size_t k = 0; k < njobs; ++k) {
            steps += acc_h[k * ACC + 6];
            unfinished += acc_h[k * ACC + ACC_UNFINISHED];
        }
        if (unfinished)
            throw Err(-5, "internal error: " + std::to_string(unfinished) + " ions never finished (status 0)");
        stats.ion_steps = steps;
        stats.kernel_ms = kms;
        stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall).count();
        return S;
    }

    // The plain path's FP64 cell records (72 doubles per cell, 576 B): built for
    // the current combined DC when tune.plain_recs asks for them and device
    // memory allows (at most half of what is free); with noalloc only in a buffer
    // that an earlier call allocated. nullptr: the kernel gathers from the node
    // tables instead, with the same results.
    const double2* plain_records(bool noalloc)
    {
        if (!tune.plain_recs) return nullptr;
        const uint64_t nc = (uint64_t)(nx - 1) * (ny - 1) * (nz - 1);
        const size_t need = nc * PLAIN_REC_D2 * sizeof(double2);
        if (plain_rec.bytes < need) {
            if (noalloc) return nullptr;
            size_t free_b = 0, total_b = 0;
            HIPCHK(hipMemGetInfo(&free_b, &total_b));
            if (need > free_b / 2) return nullptr;
            plain_rec.ensure(need);
            plain_rec_valid = false;
        }
        if (!plain_rec_valid) {
            launch_build_plain_recs(plain_rec.as<double2>(), rf.as<double>(), dcc.as<double>(), ac.as<double>(), nx,
                                    ny, nz, stream);
            HIPCHK(hipGetLastError());
            plain_rec_valid = true;
        }
        return plain_rec.as<double2>();
    }

    void run_plain(const CallBufs& cb, const BatchP& B, uint32_t n, float& kms, const double2* rec)
    {
        uint32_t nsteps = (uint32_t)std::max(1000.0, std::floor(2e-4 / B.dt));
        nsteps = (nsteps + 3u) & ~3u;
        uint32_t* live = cb.ctr;
        kms = 0.0f;
        for (;;) {
            launch_fill_u32(live, 0u, 1, stream);
            HIPCHK(hipEventRecord(ev0, stream));
            launch_plain(cb.S, cb.jobs, B, rf.as<double>(), dcc.as<double>(), ac.as<double>(), rec,
                         mask.as<uint8_t>(), n, nsteps, live, stream);
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(ev1, stream));
            stats.launches += 1;
            uint32_t alive = 0;
            HIPCHK(hipMemcpyAsync(&alive, live, 4, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
            kms += ms;
            if (alive == 0) break;
        }
    }

    // ---- differentiable form
    double colsum(double* const* cols, uint32_t ncols, uint32_t n, double* host)
    {
        adj_cols.ensure(16 * sizeof(double*));
        adj_part.ensure((size_t)ncols * std::max<uint32_t>(1, (n + 1023) / 1024) * sizeof(double));
        adj_out.ensure(16 * sizeof(double));
        HIPCHK(hipMemcpyAsync(adj_cols.p, cols, ncols * sizeof(double*), hipMemcpyHostToDevice, stream));
        launch_colsum(adj_cols.as<const double*>(), ncols, n, adj_part.as<double>(), adj_out.as<double>(), stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(host, adj_out.p, ncols * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return host[0];
    }

    double forward(const mq_job& job, const double dc[4], const mq_physics& ph, const mq_objective& o)
    {
        auto t_wall = std::chrono::steady_clock::now();
        adj_valid = false;
        if (!(ph.dt > 0.0)) throw Err(-1, "dt must be positive");
        if (job.n == 0 || job.n >= (1ull << 31)) throw Err(-1, "the job needs between 1 and 2^31 ions");
        if (o.nsteps == 0) throw Err(-1, "objective needs nsteps > 0");
        if (!(o.scale > 0.0)) throw Err(-1, "objective scale must be positive");
        if ((ph.pressure_pa > 0.0 || job.pressure_pa > 0.0) && ph.legacy_collisions)
            throw Err(-1, "the differentiable form supports the thermal collision model only (legacy_collisions = 0)");
        if (ph.rng != MQ_RNG_PHILOX) throw Err(-1, "the differentiable form injects with the Philox stream (rng = 0)");
        const uint32_t n = (uint32_t)job.n;
        JobP jp = job_params(job, ph, 0);
        const int nseg = jp.nsig_dt > 0.0 ? 6 : 3;
        // Checkpoints: 1 keeps the whole trajectory (the forward writes every step's position, and with gas its
        // post-kick velocity; the backward walks them without a recompute); 0 does so when that fits a quarter of
        // the free device memory, else checkpoints every ceil(sqrt(nsteps)) steps; K > 1 checkpoints every K.
        bool full = o.ckpt_every == 1;
        if (o.ckpt_every == 0) {
            size_t free_b = 0, total_b = 0;
            HIPCHK(hipMemGetInfo(&free_b, &total_b));
            const double traj = (double)nseg * (double)o.nsteps * (double)n * sizeof(double);
            full = traj <= 0.25 * (double)(free_b + adj_mem.bytes);
        }
        const uint32_t K = full ? o.nsteps
                                : (o.ckpt_every ? o.ckpt_every : (uint32_t)std::ceil(std::sqrt((double)o.nsteps)));
        const uint32_t slots = (o.nsteps + K - 1) / K;
        jobs.ensure(sizeof(JobP));
        HIPCHK(hipMemcpyAsync(jobs.p, &jp, sizeof(JobP), hipMemcpyHostToDevice, stream));
        prepare_dc(dc, false);
        IonState S = state_buffers(n);
        InjP ip{};
        ip.axis_x = meta.axis_x; ip.axis_y = meta.axis_y; ip.z_inject = meta.z_inject;
        ip.temp_k = ph.temp_k; ip.r_inject_mm = ph.r_inject_mm; ip.inj_sigma_mm = ph.inj_sigma_mm; ip.dt = ph.dt;
        ip.model = ph.injection == MQ_INJ_GAUSSIAN ? 2 : (ph.legacy_injection ? 1 : 0);
        launch_inject(S, jobs.as<JobP>(), 1, ip, n, stream);
        HIPCHK(hipGetLastError());
        // buffers: 6 slots of checkpoints, 3 x K segment (6 x K with gas: the post-kick velocities too),
        // 3 final, J, 5 + 6 gradients
        const size_t per_ion = (size_t)6 * slots + (size_t)nseg * K + 3 + 1 + 5 + 6;
        adj_mem.ensure(per_ion * n * sizeof(double));
        double* p = adj_mem.as<double>();
        AdjBuf A{};
        for (int k = 0; k < 6; ++k) { A.ck[k] = p; p += (size_t)slots * n; }
        for (int k = 0; k < nseg; ++k) { A.seg[k] = p; p += (size_t)K * n; }
        for (int k = 0; k < 3; ++k) { A.fin[k] = p; p += n; }
        A.Ji = p; p += n;
        for (int k = 0; k < 5; ++k) { A.gp[k] = p; p += n; }
        for (int k = 0; k < 6; ++k) { A.gs[k] = p; p += n; }
        AdjObj O{};
        O.z_exit = o.z_exit; O.scale = o.scale; O.r_leak = o.r_leak; O.w_leak = o.w_leak;
        O.axis_x = meta.axis_x; O.axis_y = meta.axis_y; O.nsteps = o.nsteps; O.ckpt = K;
        O.gas = nseg == 6 ? 1u : 0u;
        O.full = full ? 1u : 0u;
        BatchP B = batch_params(ph, 1);
        const AdjGrids G = adj_grids();
        stats = mq_run_stats{};
        stats.n_ions = n;
        HIPCHK(hipEventRecord(ev0, stream));
        launch_adj_forward(S, jobs.as<JobP>(), B, G, O, A, n, stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(ev1, stream));
        double sum[1];
        double* cols[1] = {A.Ji};
        colsum(cols, 1, n, sum);
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        stats.kernel_ms = ms;
        stats.launches = 1;
        stats.ion_steps = (uint64_t)n * o.nsteps;
        stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall).count();
        adj_buf = A; adj_obj = O; adj_B = B; adj_G = G; adj_n = n; adj_valid = true;
        return sum[0] / (double)n;
    }

    void backward(double dJ, double grad[5], double* grad_state)
    {
        auto t_wall = std::chrono::steady_clock::now();
        if (!adj_valid) throw Err(-1, "mq_backward needs a preceding mq_forward on this engine");
        const uint32_t n = adj_n;
        IonState S = state_buffers(n);
        HIPCHK(hipEventRecord(ev0, stream));
        launch_adj_backward(S, jobs.as<JobP>(), adj_B, adj_G, adj_obj, adj_buf, n, dJ / (double)n, stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(ev1, stream));
        double g[5];
        double* cols[5] = {adj_buf.gp[0], adj_buf.gp[1], adj_buf.gp[2], adj_buf.gp[3], adj_buf.gp[4]};
        colsum(cols, 5, n, g);
        for (int k = 0; k < 5; ++k) grad[k] = g[k];
        if (grad_state)
            for (int k = 0; k < 6; ++k)
                HIPCHK(hipMemcpyAsync(grad_state + (size_t)k * n, adj_buf.gs[k], (size_t)n * sizeof(double),
                                      hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        stats = mq_run_stats{};
        stats.n_ions = n;
        stats.kernel_ms = ms;
        stats.launches = 1;
        stats.ion_steps = (adj_obj.full ? 1ull : 2ull) * n * adj_obj.nsteps;  // (recompute +) backward
        stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall).count();
    }

    SortGeom sort_geom() const
    {
        SortGeom G{};
        G.sx = G.sy = G.sz = 2;  // 4 x 4 x 4 cells per bin
        G.nbx = (uint32_t)((nx - 2) >> G.sx) + 1;
        G.nby = (uint32_t)((ny - 2) >> G.sy) + 1;
        G.nbz = (uint32_t)((nz - 2) >> G.sz) + 1;
        G.nbins = G.nbx * G.nby * G.nbz;
        return G;
    }

    // Persistent-worker launches (run_tail, tiled.rs:159-218): each launch advances
    // every listed ion by up to `chunk` steps and appends the unfinished ones to the
    // next list. With re-sort on, the list is sorted by cell before every launch
    // and consumed as one segment per XCD, and the chunk stays fixed (it is the
    // re-sort interval); otherwise the chunk doubles up to its cap. `sync_every`
    // launches are enqueued back to back (every kernel reads its list length from
    // device memory, so launches past the end of the work are empty) and the host
    // checks the live count once per block; with `graph` on, the block is
    // captured once as a hipGraph and replayed.
    // bins_known: bin_of, rank_of and hist hold the bins, ranks within their bin
    // and counts per bin of list[cur] (the last launch wrote them); pass_bins:
    // each launch counts and ranks the ions it carries, so that the next re-sort
    // neither gathers their positions nor ranks them again (off in graph mode,
    // where one captured block must serve every replay).
    void enqueue_block(const CallBufs& cb, const BatchP& B, uint32_t n, const SortGeom& G, int& cur, uint32_t& chunk,
                       uint32_t cap, uint32_t period, uint32_t grid, uint32_t block, int iters, bool timed,
                       std::vector<hipEvent_t>* evs, bool resort, bool& bins_known, bool pass_bins,
                       uint32_t lanes = 64, bool staged = false, uint32_t quad_grid = 0)
    {
        const bool quad = quad_grid != 0;
        uint32_t* c0 = cb.ctr;
        uint32_t* c1 = c0 + CTR_BLOCK;
        for (int it = 0; it < iters; ++it) {
            uint32_t* cin = cur ? c1 : c0;
            uint32_t* cout = cur ? c0 : c1;
            const uint32_t* lin = cb.list[cur];
            // The injected list is not sorted: its ions all start in the few bins
            // of the injection spot, where sorting gains nothing and the bin
            // counts serialise on a handful of atomics (0.5 ms on the mass scan).
            // The first launch takes it as it is and writes the bins for the next.
            if (resort && (bins_known || !pass_bins)) {
                launch_sort(cb.S, B, lin, cin, cb.list[2], cb.bin_of, cb.rank_of, cb.hist, cb.bsum, G, n, bins_known,
                            stream);
                lin = cb.list[2];
            }
            WorkerArgs a{};
            a.S = cb.S;
            a.jobs = cb.jobs;
            a.B = B;
            a.cells = cells.as<float4>();
            a.list_in = lin;
            a.list_out = cb.list[cur ^ 1];
            a.ctr_in = cin;
            a.ctr_out = cout;
            a.chunk = chunk;
            a.period = period;
            // one queue per XCD, each the home of at least one workgroup
            a.nqueues = resort ? std::min<uint32_t>(MAX_QUEUES, grid) : 1u;
            a.coll_batch = (uint32_t)std::max(1, tune.coll_batch);
            a.bin_out = resort && pass_bins ? cb.bin_of : nullptr;
            // the worker also counts and ranks the ions it carries, so the next re-sort needs no ranking pass
            a.rank_out = a.bin_out ? cb.rank_of : nullptr;
            a.hist_out = a.bin_out ? cb.hist : nullptr;
            if (a.hist_out) launch_fill_hist(cb.hist, G.nbins, stream);
            a.sg = G;
            a.lanes = lanes;
            bins_known = a.bin_out != nullptr;
            launch_prepare(a, stream);
            if (timed) HIPCHK(hipEventRecord((*evs)[2 * it], stream));
            // staged tiles: 256-thread workgroups, two per CU (63 KB of LDS each)
            if (staged) {
                launch_staged(a, (uint32_t)cus * 2u, stream);
            } else if (quad) {
                // four lanes per ion: the same wavefronts in 256-thread workgroups, a refill
                // period of at most 8 steps (the candidate blocks a quad draws cover 8)
                if (a.nqueues != 1 || a.bin_out) throw Err(-1, "the four-lane tail takes one queue and no re-sort");
                a.period = std::min(period, 8u);
                launch_tail4(a, quad_grid, stream);
            } else {
                launch_batched(a, grid, block, stream);
            }
            if (timed) HIPCHK(hipEventRecord((*evs)[2 * it + 1], stream));
            cur ^= 1;
            if (!resort && !quad) chunk = std::min(cap, chunk * 2u);
        }
    }

    // True when every value of the n doubles at a is exactly a float.
    bool exact_f32(const double* a, uint64_t n)
    {
        exact_flag.ensure(4);
        launch_fill_u32(exact_flag.as<uint32_t>(), 1u, 1, stream);
        launch_exact_f32(a, n, exact_flag.as<uint32_t>(), stream);
        uint32_t f = 0;
        HIPCHK(hipMemcpyAsync(&f, exact_flag.p, 4, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return f != 0;
    }

    // Tables of the differentiable form for the current combined DC: per-cell
    // float records of the corner values when RF, AC and the combined DC are all
    // exactly floats (the kernels convert back to double exactly, so the results
    // are the same bit for bit), else the double node tables.
    AdjGrids adj_grids()
    {
        const uint64_t n3 = 3 * nodes();
        if (rfac32_state < 0) rfac32_state = exact_f32(rf.as<double>(), n3) && exact_f32(ac.as<double>(), n3) ? 1 : 0;
        if (rfac32_state == 1 && dcc32_state < 0) {
            dcc32_state = exact_f32(dcc.as<double>(), n3) ? 1 : 0;
            if (dcc32_state) {
                const uint64_t nc = (uint64_t)(nx - 1) * (ny - 1) * (nz - 1);
                adj_rec.ensure(nc * ADJ_REC_F4 * sizeof(float4));
                launch_build_adj_recs(adj_rec.as<float4>(), rf.as<double>(), dcc.as<double>(), ac.as<double>(), nx, ny,
                                      nz, stream);
            }
        }
        AdjGrids G{rf.as<double>(), dcc.as<double>(), ac.as<double>(), nullptr};
        if (rfac32_state == 1 && dcc32_state == 1) G.rec = adj_rec.as<float4>();
        HIPCHK(hipGetLastError());
        return G;
    }

    // noalloc: no graph capture and no per-launch events (the engine's two events
    // time the whole run); results per ion are the same either way.
    void run_batched(const CallBufs& cb, const BatchP& B, uint32_t n, float& kms)
    {
        const bool resort = tune.resort != 0;
        const bool staged = tune.resort == 2 && !B.compat;  // the staged worker draws from Philox streams only
        uint32_t* c0 = cb.ctr;
        const SortGeom G = resort ? sort_geom() : SortGeom{};
        launch_iota(cb.list[0], n, stream);
        launch_fill_u32(c0 + CTR_COUNT, n, 1, stream);
        const uint32_t block = (uint32_t)tune.block;
        const uint32_t grid = std::max<uint32_t>(1, (uint32_t)((uint64_t)cus * tune.waves_per_cu * 64 / block));
        uint32_t chunk = ((uint32_t)tune.chunk + 3u) & ~3u;
        uint32_t cap = std::max(chunk, ((uint32_t)tune.chunk_cap + 3u) & ~3u);
        const uint32_t period = std::max(4u, ((uint32_t)tune.refill_period + 3u) & ~3u);
        const int iters = std::max(1, tune.sync_every);
        kms = 0.0f;
        int cur = 0;
        if (tune.graph && !cb.noalloc) {
            // capture one block of `iters` launches; replay it until the list is empty
            hipGraph_t g = nullptr;
            hipGraphExec_t ge = nullptr;
            int cur_cap = cur;
            uint32_t chunk_cap = chunk;
            bool known = false;
            HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            try {
                enqueue_block(cb, B, n, G, cur_cap, chunk_cap, cap, period, grid, block, iters, false, nullptr, resort,
                              known, false, 64, staged);
            } catch (...) {
                (void)hipStreamEndCapture(stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIPCHK(hipStreamEndCapture(stream, &g));
            HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            // the block leaves its output in list[cur_cap]; with an even number of
            // launches the ping-pong ends where it started and the block can repeat
            if (cur_cap != cur) {
                (void)hipGraphExecDestroy(ge);
                (void)hipGraphDestroy(g);
                throw Err(-1, "graph mode needs an even sync_every");
            }
            uint32_t* cfinal = cur ? c0 + CTR_BLOCK : c0;
            for (;;) {
                HIPCHK(hipEventRecord(ev0, stream));
                HIPCHK(hipGraphLaunch(ge, stream));
                HIPCHK(hipEventRecord(ev1, stream));
                stats.launches += (uint32_t)iters;
                uint32_t alive = 0;
                HIPCHK(hipMemcpyAsync(&alive, cfinal + CTR_COUNT, 4, hipMemcpyDeviceToHost, stream));
                HIPCHK(hipStreamSynchronize(stream));
                float ms = 0.0f;
                HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
                kms += ms;
                if (alive == 0) break;
            }
            (void)hipGraphExecDestroy(ge);
            (void)hipGraphDestroy(g);
            return;
        }
        const bool timed = !cb.noalloc;
        if (timed && (int)evs.size() < 2 * iters) {
            for (auto e : evs) (void)hipEventDestroy(e);
            evs.assign(2 * iters, nullptr);
            for (auto& e : evs) HIPCHK(hipEventCreate(&e));
        }
        if (!timed) HIPCHK(hipEventRecord(ev0, stream));
        // Tail: once no more ions are live than the launch has lanes, a re-sort no
        // longer puts several ions of one region into a wavefront, and the launches
        // are bound by the step latency of the longest-lived ions. The remaining
        // launches then skip the re-sort and double the chunk up to the reference's
        // 200000 (tiled.rs:26-27, 214), so the tail takes a few launches, not one per
        // re-sort interval.
        const uint64_t lanes = (uint64_t)grid * block;
        const uint64_t waves = lanes / 64;
        bool tail = false;
        bool bins_known = false;  // list[0] is the injected list: no bins yet
        uint32_t alive_last = n;
        for (;;) {
            // In the tail the few live ions are spread over every wavefront (tail_lanes ions per wave at most,
            // or as few as fill all waves when tail_lanes is 0), so that a step's record reloads are not
            // serialised over 64 lanes of divergent addresses.
            uint32_t act = 64;
            // Once at most 16 ions per wavefront are left (Philox streams), each takes four lanes
            // (k_tail4): a wavefront's step then runs about half the instructions. Its wavefronts are
            // packed with 16 ions each and only as many are launched as the ions fill, so that once they
            // fill no more wavefronts than there are SIMDs each busy SIMD steps one; a fixed chunk of
            // QUAD_CHUNK steps repacks the shrinking population at every launch.
            const bool quad = tail && tune.tail_quads && !B.compat && alive_last <= 16 * waves;
            uint32_t quad_grid = 0, qchunk = QUAD_CHUNK;
            if (quad) {
                act = tune.tail_lanes > 0 ? std::min<uint32_t>(16u, (uint32_t)tune.tail_lanes) : 16u;
                const uint64_t wv = (alive_last + act - 1) / act;  // wavefronts the live ions fill
                quad_grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(waves / 4, (wv + 3) / 4));
            } else if (tail) {
                act = tune.tail_lanes > 0 ? (uint32_t)tune.tail_lanes
                                          : (uint32_t)std::min<uint64_t>(64, (alive_last + waves - 1) / waves);
                act = std::max(1u, std::min(64u, act));
            }
            // the tail's few launches are checked one at a time, so that each picks its worker
            // and shape from its own live count
            const int it = tail ? 1 : iters;
            enqueue_block(cb, B, n, G, cur, quad ? qchunk : chunk, cap, period, grid, block, it, timed, &evs,
                          resort && !tail, bins_known, true, act, staged && !tail, quad_grid);
            HIPCHK(hipGetLastError());
            stats.launches += (uint32_t)it;
            if (quad_grid) stats.quad_launches += (uint32_t)it;
            uint32_t alive = 0;
            uint32_t* cfinal = cur ? c0 + CTR_BLOCK : c0;
            HIPCHK(hipMemcpyAsync(&alive, cfinal + CTR_COUNT, 4, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            if (timed) {
                for (int k = 0; k < it; ++k) {
                    float ms = 0.0f;
                    HIPCHK(hipEventElapsedTime(&ms, evs[2 * k], evs[2 * k + 1]));
                    kms += ms;
                }
            }
            if (alive == 0) break;
            alive_last = alive;
            if (resort && !tail && alive <= lanes) {
                tail = true;
                cap = std::max(cap, 200000u);
            }
        }
        if (!timed) {
            HIPCHK(hipEventRecord(ev1, stream));
            HIPCHK(hipEventSynchronize(ev1));
            HIPCHK(hipEventElapsedTime(&kms, ev0, ev1));
        }
    }
};

// ------------------------------------------------------------------ engine view
// What the live population (live.cpp) needs from an engine: its stream, CUs,
// tuning, geometry and the FP32 cell table of the DC configuration dc.
namespace mq {
namespace host {

EngineView engine_view(mq_engine* e, const double dc[4], const mq_physics& ph, uint32_t njobs)
{
    e->prepare_dc(dc, true);
    HIPCHK(hipGetLastError());
    EngineView v;
    v.stream = e->stream;
    v.cus = e->cus;
    v.tune = e->tune;
    v.meta = e->meta;
    v.cells = e->cells.as<float4>();
    v.B = e->batch_params(ph, njobs);
    return v;
}

}  // namespace host
}  // namespace mq

// ------------------------------------------------------------------ autotuner

namespace {

std::string tune_key(const mq_engine& e, const mq_physics& ph)
{
    char b[256];
    std::snprintf(b, sizeof b, "grid=%dx%dx%d h=%.9g legacy=%d%d%d%d%d dt=%.9g", e.nx, e.ny, e.nz, e.h, ph.legacy_force,
                  ph.legacy_collisions, ph.legacy_fields, ph.legacy_injection, ph.legacy_criteria, ph.dt);
    return b;
}

bool read_tune_cache(const char* path, const std::string& key, mq_tune& t)
{
    if (!path) return false;
    FILE* f = std::fopen(path, "r");
    if (!f) return false;
    char line[1024];
    bool found = false;
    while (std::fgets(line, sizeof line, f)) {
        std::string l(line);
        const size_t tab = l.find('\t');
        if (tab == std::string::npos || l.substr(0, tab) != key) continue;
        mq_tune x{};
        x.tail_lanes = 64;  // entries written before the tail_lanes field
        x.tail_quads = 1;   // entries written before the tail_quads field (results are the same either way)
        if (std::sscanf(l.c_str() + tab + 1, "%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d", &x.block, &x.waves_per_cu,
                        &x.refill_period, &x.chunk, &x.chunk_cap, &x.resort, &x.sync_every, &x.graph,
                        &x.coll_batch, &x.tail_lanes, &x.tail_quads) >= 9) {
            t = x;  // the last entry for a key wins
            found = true;
        }
    }
    std::fclose(f);
    return found;
}

}  // namespace

// Picks the fastest launch configuration for this grid and physics by timing
// candidate tunings on an evenly spaced subset of the jobs (about `sample_ions`
// ions), and appends the winner to the TSV cache `cache_path` (key: grid and
// physics switches), where later calls find it without measuring. Results per
// ion do not depend on the tuning.
static void autotune(mq_engine& e, const mq_job* jobs, size_t njobs, const double dc[4], const mq_physics& ph,
                     uint64_t sample_ions, const char* cache_path, int force, mq_tune& best, double& best_rate)
{
    const std::string key = tune_key(e, ph);
    best_rate = 0.0;
    if (!force && read_tune_cache(cache_path, key, best)) {
        best.plain_recs = e.tune.plain_recs;  // the plain path's setting is not part of the entries
        e.tune = best;
        return;
    }
    uint64_t total = 0;
    for (size_t k = 0; k < njobs; ++k) total += jobs[k].n;
    if (total == 0) throw Err(-1, "autotune needs ions");
    const size_t take = std::max<size_t>(1, std::min<size_t>(njobs, (size_t)std::ceil((double)njobs *
                                                                                      (double)sample_ions / (double)total)));
    std::vector<mq_job> sub;
    for (size_t k = 0; k < take; ++k) sub.push_back(jobs[(size_t)((double)k * (double)njobs / (double)take)]);
    // Coordinate search: starting from the defaults (re-sort on, 256-step interval, 8 waves per CU, 256-thread
    // groups, refill every 8 steps, collision batch 4, no graph), each knob in turn is swept over its values with
    // the others held at the best so far; two passes. A run whose rate the best beats by less than 1 % keeps the
    // earlier choice, so noise does not walk the tuning around.
    mq_tune base = e.tune;
    base.block = 256; base.waves_per_cu = 8; base.refill_period = 8; base.chunk = 256; base.chunk_cap = 256;
    base.resort = 1; base.sync_every = 16; base.graph = 0; base.coll_batch = 4; base.tail_lanes = 0;
    base.tail_quads = 1;
    struct Knob {
        const char* name;
        std::vector<int> values;
        void (*set)(mq_tune&, int);
    };
    const std::vector<Knob> knobs = {
        {"resort", {1, 0, 2}, [](mq_tune& t, int v) {
             t.resort = v;
             t.chunk = v ? 256 : 20000;
             t.chunk_cap = v ? 256 : 200000;
         }},
        {"chunk", {128, 256, 512}, [](mq_tune& t, int v) {
             if (t.resort) { t.chunk = v; t.chunk_cap = v; }
         }},
        {"waves_per_cu", {4, 8, 12}, [](mq_tune& t, int v) { t.waves_per_cu = v; }},
        {"block", {64, 128, 256}, [](mq_tune& t, int v) { t.block = v; }},
        {"refill_period", {4, 8, 16}, [](mq_tune& t, int v) { t.refill_period = v; }},
        {"coll_batch", {1, 4, 8}, [](mq_tune& t, int v) { t.coll_batch = v; }},
        {"graph", {0, 1}, [](mq_tune& t, int v) { t.graph = v; }},
        {"tail_lanes", {0, 64}, [](mq_tune& t, int v) { t.tail_lanes = v; }},
        {"tail_quads", {1, 0}, [](mq_tune& t, int v) { t.tail_quads = v; }},
    };
    const mq_tune saved = e.tune;
    std::vector<unsigned long long> acc;
    size_t n = 0;
    std::vector<std::pair<mq_tune, double>> measured;
    auto same = [](const mq_tune& a, const mq_tune& b) { return std::memcmp(&a, &b, sizeof a) == 0; };
    auto rate_of = [&](const mq_tune& t) {
        for (auto& m : measured)
            if (same(m.first, t)) return m.second;
        e.tune = t;
        e.run(sub.data(), sub.size(), dc, ph, MQ_PATH_BATCHED, acc, n);
        const double r = (double)e.stats.ion_steps / (e.stats.wall_ms * 1e-3);
        measured.push_back({t, r});
        return r;
    };
    try {
        best = base;
        best_rate = rate_of(base);
        for (int pass = 0; pass < 2; ++pass) {
            for (const Knob& k : knobs) {
                for (int v : k.values) {
                    mq_tune t = best;
                    k.set(t, v);
                    const double r = rate_of(t);
                    if (r > 1.01 * best_rate) { best_rate = r; best = t; }
                }
            }
        }
    } catch (...) {
        e.tune = saved;
        throw;
    }
    e.tune = best;
    if (cache_path) {
        FILE* f = std::fopen(cache_path, "a");
        if (f) {
            std::fprintf(f, "%s\t%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d\t%.6g\n", key.c_str(), best.block,
                         best.waves_per_cu, best.refill_period, best.chunk, best.chunk_cap, best.resort,
                         best.sync_every, best.graph, best.coll_batch, best.tail_lanes, best.tail_quads, best_rate);
            std::fclose(f);
        }
    }
}

// ------------------------------------------------------------------ C ABI

void mq_set_last_error(const std::string& msg) { g_err = msg; }

#define MQ_GUARD(...)                                       \
    try {                                                   \
        g_err.clear();                                      \
        __VA_ARGS__;                                        \
        return 0;                                           \
    } catch (const Err& e) {                                \
        g_err = e.what();                                   \
        return e.code;                                      \
    } catch (const std::exception& e) {                     \
        g_err = e.what();                                   \
        return -9;                                          \
    }

extern "C" {

const char* mq_last_error(void) { return g_err.c_str(); }

int mq_engine_create(const mq_fields* fields, mq_engine** out)
{
    mq_engine* e = nullptr;
    try {
        g_err.clear();
        if (!out) throw Err(-1, "out is NULL");
        e = new mq_engine();
        e->create(fields);
        *out = e;
        return 0;
    } catch (const Err& x) {
        g_err = x.what();
        if (e) { e->destroy(); delete e; }
        return x.code;
    } catch (const std::exception& x) {
        g_err = x.what();
        if (e) { e->destroy(); delete e; }
        return -9;
    }
}

void mq_engine_destroy(mq_engine* eng)
{
    if (!eng) return;
    eng->destroy();
    delete eng;
}

static void fill_outcomes(const std::vector<unsigned long long>& a, size_t njobs, double dt, mq_outcome* out)
{
    for (size_t k = 0; k < njobs; ++k) {
        const unsigned long long* p = &a[k * ACC];
        mq_outcome o{};
        o.n = p[0];
        o.crashed = p[1]; o.returned = p[2]; o.transmitted = p[3]; o.left = p[4]; o.timeout = p[5];
        o.mean_coll = o.n ? (double)p[7] / (double)o.n : 0.0;
        o.mean_time = o.n ? (double)p[6] * dt / (double)o.n : 0.0;
        out[k] = o;
    }
}

int mq_run_batch(mq_engine* eng, const mq_job* jobs, size_t njobs, const double dc[4], const mq_physics* ph,
                 int32_t path, mq_outcome* out)
{
    MQ_GUARD({
        if (!eng || !ph || !dc || !out) throw Err(-1, "NULL argument");
        std::vector<unsigned long long> acc;
        size_t n = 0;
        eng->run(jobs, njobs, dc, *ph, path, acc, n);
        fill_outcomes(acc, njobs, ph->dt, out);
    })
}

// Final states of the last run into the caller's arrays. The FP32 velocities
// of the batched path are copied into the caller's double arrays and widened in
// place (from the last element down: element i is read before the doubles at
// or above i, which cover floats 2i and 2i + 1, are written), so no host buffer
// is needed.
static void copy_out(mq_engine* eng, const IonState& S, size_t n, int32_t path, const mq_ion_soa* out,
                     uint32_t* out_job)
{
    hipStream_t st = eng->stream;
    auto d2h = [&](void* dst, const void* src, size_t bytes) {
        if (dst) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    };
    d2h(out->x, S.x, n * 8); d2h(out->y, S.y, n * 8); d2h(out->z, S.z, n * 8);
    d2h(out->steps, S.steps, n * 4); d2h(out->ncoll, S.ncoll, n * 4); d2h(out->status, S.status, n);
    d2h(out->phi_ac0, S.phi, n * 8); d2h(out->cd, S.cd, n * 4);
    if (out_job) d2h(out_job, S.job, n * 4);
    if (path == MQ_PATH_PLAIN) {
        d2h(out->vx, S.vx, n * 8); d2h(out->vy, S.vy, n * 8); d2h(out->vz, S.vz, n * 8);
        HIPCHK(hipStreamSynchronize(st));
        return;
    }
    double* dsts[3] = {out->vx, out->vy, out->vz};
    float* srcs[3] = {S.fvx, S.fvy, S.fvz};
    for (int d = 0; d < 3; ++d) d2h(dsts[d], srcs[d], n * 4);
    HIPCHK(hipStreamSynchronize(st));
    for (int d = 0; d < 3; ++d) {
        if (!dsts[d]) continue;
        unsigned char* b = reinterpret_cast<unsigned char*>(dsts[d]);
        for (size_t i = n; i-- > 0;) {
            float f;
            std::memcpy(&f, b + 4 * i, 4);
            const double v = (double)f;
            std::memcpy(b + 8 * i, &v, 8);
        }
    }
}

int mq_run_batch_ions(mq_engine* eng, const mq_job* jobs, size_t njobs, const double dc[4], const mq_physics* ph,
                      int32_t path, const mq_ion_soa* out, uint32_t* out_job)
{
    MQ_GUARD({
        if (!eng || !ph || !dc || !out) throw Err(-1, "NULL argument");
        std::vector<unsigned long long> acc;
        size_t n = 0;
        IonState S = eng->run(jobs, njobs, dc, *ph, path, acc, n);
        if (n == 0) return 0;
        copy_out(eng, S, n, path, out, out_job);
    })
}

int mq_run_batch_ions_noalloc(mq_engine* eng, const mq_job* jobs, size_t njobs, const double dc[4],
                              const mq_physics* ph, int32_t path, const mq_ion_soa* out, uint32_t* out_job)
{
    MQ_GUARD({
        if (!eng || !ph || !dc || !out) throw Err(-1, "NULL argument");
        size_t n = 0;
        IonState S = eng->run(jobs, njobs, dc, *ph, path, eng->acc_host, n, true);
        if (n == 0) return 0;
        copy_out(eng, S, n, path, out, out_job);
    })
}

int mq_summarize(const uint8_t* status, const uint32_t* ncoll, const uint32_t* steps, const uint32_t* job, size_t n,
                 size_t njobs, double dt, mq_outcome* out)
{
    MQ_GUARD({
        if (!out || (n && (!status || !ncoll || !steps || !job))) throw Err(-1, "NULL argument");
        std::vector<unsigned long long> a(njobs * ACC, 0ull);
        for (size_t i = 0; i < n; ++i) {
            if (job[i] >= njobs) throw Err(-1, "job index out of range");
            unsigned long long* p = &a[(size_t)job[i] * ACC];
            p[0] += 1;
            uint8_t s = status[i];
            p[(s >= 1 && s <= 5) ? s : 5] += 1;  // summarize counts any other status as a timeout
            p[6] += steps[i];
            p[7] += ncoll[i];
        }
        fill_outcomes(a, njobs, dt, out);
    })
}

int mq_last_stats(const mq_engine* eng, mq_run_stats* out)
{
    MQ_GUARD({
        if (!eng || !out) throw Err(-1, "NULL argument");
        *out = eng->stats;
    })
}

int mq_gather_probe(mq_engine* eng, uint64_t window_cells, uint32_t iters, double* bytes_per_s)
{
    MQ_GUARD({
        if (!eng || !bytes_per_s) throw Err(-1, "NULL argument");
        if (!eng->cells_valid) throw Err(-1, "no cell table yet: run a batched batch first");
        const uint64_t nc = (uint64_t)(eng->nx - 1) * (eng->ny - 1) * (eng->nz - 1);
        const uint64_t window = window_cells ? std::min<uint64_t>(window_cells, nc) : nc;
        if (iters == 0 || iters > 4096) throw Err(-1, "iters must be in [1, 4096]");
        const uint32_t block = 256;
        const uint32_t grid = std::max<uint32_t>(1, (uint32_t)((uint64_t)eng->cus * eng->tune.waves_per_cu * 64 / block));
        DevBuf out;
        out.ensure((size_t)grid * block * sizeof(float));
        launch_gather_probe(eng->cells.as<float4>(), window, iters, out.as<float>(), grid, block, eng->stream);  // warm
        HIPCHK(hipEventRecord(eng->ev0, eng->stream));
        launch_gather_probe(eng->cells.as<float4>(), window, iters, out.as<float>(), grid, block, eng->stream);
        HIPCHK(hipEventRecord(eng->ev1, eng->stream));
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventSynchronize(eng->ev1));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, eng->ev0, eng->ev1));
        out.release();
        *bytes_per_s = (double)grid * block * iters * CELL_F4 * 16.0 / (ms * 1e-3);
    })
}

int mq_field_probe(mq_engine* eng, const double dc[4], int32_t nearest, double a_rf, double s_dc, double a_ac,
                   size_t n, const double* x, const double* y, const double* z, float* ex, float* ey, float* ez,
                   uint8_t* crash)
{
    MQ_GUARD({
        if (!eng || !dc || (n && (!x || !y || !z || !ex || !ey || !ez || !crash))) throw Err(-1, "NULL argument");
        if (n >= (1ull << 31)) throw Err(-1, "too many probe points");
        if (n == 0) return 0;
        eng->prepare_dc(dc, true);
        mq_physics ph{};
        ph.dt = 1.0;
        ph.legacy_fields = nearest ? 1 : 0;
        const BatchP B = eng->batch_params(ph, 1);
        DevBuf pos, fout, cout;
        pos.ensure(3 * n * sizeof(double));
        fout.ensure(3 * n * sizeof(float));
        cout.ensure(n);
        double* p = pos.as<double>();
        float* f = fout.as<float>();
        hipStream_t st = eng->stream;
        try {
            HIPCHK(hipMemcpyAsync(p, x, n * 8, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(p + n, y, n * 8, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(p + 2 * n, z, n * 8, hipMemcpyHostToDevice, st));
            launch_field_probe(eng->cells.as<float4>(), B, (uint32_t)n, p, p + n, p + 2 * n, (float)a_rf, (float)s_dc,
                               (float)a_ac, f, f + n, f + 2 * n, cout.as<uint8_t>(), st);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(ex, f, n * 4, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(ey, f + n, n * 4, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(ez, f + 2 * n, n * 4, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(crash, cout.p, n, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
        } catch (...) {
            pos.release(); fout.release(); cout.release();
            throw;
        }
        pos.release(); fout.release(); cout.release();
    })
}

int mq_set_tune(mq_engine* eng, const mq_tune* t)
{
    MQ_GUARD({
        if (!eng || !t) throw Err(-1, "NULL argument");
        mq_tune n = eng->tune;
        if (t->block) {
            // k_batched is compiled for at most 256 threads per workgroup (__launch_bounds__)
            if (t->block < 64 || t->block > 256 || t->block % 64) throw Err(-1, "block must be a multiple of 64 in [64, 256]");
            n.block = t->block;
        }
        if (t->waves_per_cu) {
            if (t->waves_per_cu < 1 || t->waves_per_cu > 32) throw Err(-1, "waves_per_cu must be in [1, 32]");
            n.waves_per_cu = t->waves_per_cu;
        }
        if (t->refill_period) {
            if (t->refill_period < 4 || t->refill_period > 4096) throw Err(-1, "refill_period must be in [4, 4096]");
            n.refill_period = t->refill_period;
        }
        if (t->chunk) {
            if (t->chunk < 4) throw Err(-1, "chunk must be >= 4");
            n.chunk = t->chunk;
        }
        if (t->chunk_cap) n.chunk_cap = t->chunk_cap;
        if (t->resort < 0 || t->resort > 2) throw Err(-1, "resort must be 0, 1 or 2");
        n.resort = t->resort;
        if (t->sync_every) {
            if (t->sync_every < 1 || t->sync_every > 1024) throw Err(-1, "sync_every must be in [1, 1024]");
            n.sync_every = t->sync_every;
        }
        if (t->graph < 0 || t->graph > 1) throw Err(-1, "graph must be 0 or 1");
        if (t->coll_batch) {
            if (t->coll_batch < 1 || t->coll_batch > 64) throw Err(-1, "coll_batch must be in [1, 64]");
            n.coll_batch = t->coll_batch;
        }
        if (t->tail_lanes < 0 || t->tail_lanes > 64) throw Err(-1, "tail_lanes must be in [0, 64]");
        n.tail_lanes = t->tail_lanes;
        if (t->tail_quads < 0 || t->tail_quads > 1) throw Err(-1, "tail_quads must be 0 or 1");
        n.tail_quads = t->tail_quads;
        if (t->plain_recs < 0 || t->plain_recs > 1) throw Err(-1, "plain_recs must be 0 or 1");
        n.plain_recs = t->plain_recs;
        n.graph = t->graph;
        if (n.graph && (n.sync_every % 2)) throw Err(-1, "graph mode needs an even sync_every");
        eng->tune = n;
    })
}

int mq_autotune(mq_engine* eng, const mq_job* jobs, size_t njobs, const double dc[4], const mq_physics* ph,
                uint64_t sample_ions, const char* cache_path, int32_t force, mq_tune* out, double* ion_steps_per_s)
{
    MQ_GUARD({
        if (!eng || !jobs || !dc || !ph) throw Err(-1, "NULL argument");
        mq_tune best{};
        double rate = 0.0;
        autotune(*eng, jobs, njobs, dc, *ph, sample_ions ? sample_ions : 262144, cache_path, force, best, rate);
        if (out) *out = best;
        if (ion_steps_per_s) *ion_steps_per_s = rate;
    })
}

int mq_forward(mq_engine* eng, const mq_job* job, const double dc[4], const mq_physics* ph, const mq_objective* obj,
               double* J)
{
    MQ_GUARD({
        if (!eng || !job || !dc || !ph || !obj || !J) throw Err(-1, "NULL argument");
        *J = eng->forward(*job, dc, *ph, *obj);
    })
}

int mq_backward(mq_engine* eng, double dJ_up, double grad[5], double* grad_state)
{
    MQ_GUARD({
        if (!eng || !grad) throw Err(-1, "NULL argument");
        eng->backward(dJ_up, grad, grad_state);
    })
}

int mq_forward_per_ion(mq_engine* eng, double* J_ion)
{
    MQ_GUARD({
        if (!eng || !J_ion) throw Err(-1, "NULL argument");
        if (!eng->adj_valid) throw Err(-1, "needs a preceding mq_forward on this engine");
        HIPCHK(hipMemcpy(J_ion, eng->adj_buf.Ji, (size_t)eng->adj_n * sizeof(double), hipMemcpyDeviceToHost));
    })
}

int mq_forward_final(mq_engine* eng, double* x, double* y, double* z)
{
    MQ_GUARD({
        if (!eng || !x || !y || !z) throw Err(-1, "NULL argument");
        if (!eng->adj_valid) throw Err(-1, "needs a preceding mq_forward on this engine");
        double* out[3] = {x, y, z};
        for (int d = 0; d < 3; ++d)
            HIPCHK(hipMemcpy(out[d], eng->adj_buf.fin[d], (size_t)eng->adj_n * sizeof(double), hipMemcpyDeviceToHost));
    })
}

int mq_backward_per_ion(mq_engine* eng, double* grad_ion)
{
    MQ_GUARD({
        if (!eng || !grad_ion) throw Err(-1, "NULL argument");
        if (!eng->adj_valid) throw Err(-1, "needs a preceding mq_forward and mq_backward on this engine");
        for (int k = 0; k < 5; ++k)
            HIPCHK(hipMemcpy(grad_ion + (size_t)k * eng->adj_n, eng->adj_buf.gp[k], (size_t)eng->adj_n * sizeof(double),
                             hipMemcpyDeviceToHost));
    })
}

int mq_get_tune(const mq_engine* eng, mq_tune* out)
{
    MQ_GUARD({
        if (!eng || !out) throw Err(-1, "NULL argument");
        *out = eng->tune;
    })
}

}  // extern "C"
