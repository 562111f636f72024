// tandem.cpp -- host side of the tandem multi-domain batch (include/mq_tandem.h).
//
// Owns the two grids (cube: splitter and combiner; channel: the nine ports),
// the four combined DC fields and FP32 cell tables they give (splitter,
// combiner, driven port, idle port), pooled ion state and the per-port
// counters. The launch loops follow the engine's: FP64 relaunch every
// max(2e-4/dt, 1000) steps (channel.rs:478-493), FP32 persistent workers with
// chunk doubling (tiled.rs:159-218).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mq_tandem.h"
#include "host_util.h"
#include "kernels.h"
#include "mq_types.h"

using namespace mq;
using namespace mq::host;

void mq_set_last_error(const std::string& msg);  // engine.cpp

namespace mq {
void launch_tandem_init(const IonState& S, uint32_t n, hipStream_t st);
// rec: per-cell FP64 records of the four field sets (launch_build_plain_recs), or
// nullptr for the node tables.
void launch_tandem_plain(const IonState& S, const JobP* jobs, const TandemP& T, const TGrid& cube, const TGrid& port,
                         const double* const dc[4], const double2* const rec[4], unsigned long long* ctr, uint32_t n,
                         uint32_t nsteps, uint32_t* live, hipStream_t st);
void launch_tandem_batched(const IonState& S, const JobP* jobs, const TandemP& T, const TGrid& cube,
                           const TGrid& port, const float4* const cells[4], unsigned long long* ctr,
                           const uint32_t* list_in, uint32_t* list_out, uint32_t* ctr_in, uint32_t* ctr_out,
                           uint32_t chunk, uint32_t period, uint32_t nqueues, uint32_t grid, uint32_t block,
                           hipStream_t st);
void launch_tandem_bin(const IonState& S, const TandemP& T, const TGrid& cube, const TGrid& port, const uint32_t* list,
                       const uint32_t* ctr, uint32_t* bin_of, uint32_t* rank_of, uint32_t* hist, uint32_t nbc,
                       uint32_t nbp, uint32_t nmax, hipStream_t st);
void launch_tandem_totals(const IonState& S, uint32_t n, unsigned long long* tot, hipStream_t st);
}  // namespace mq

namespace {

// One uploaded grid: FP64 bases, mask, geometry.
struct Grid {
    mq_fields meta{};
    DevBuf rf, ac, dc[4], mask;
    bool has_ac = false, has_dc[4] = {false, false, false, false};
    size_t nodes() const { return (size_t)meta.nx * meta.ny * meta.nz; }

    void upload(const mq_fields* f, bool need_ac, const char* what, hipStream_t st)
    {
        if (!f) throw Err(-1, std::string(what) + " grid is NULL");
        if (f->nx < 3 || f->ny < 3 || f->nz < 3) throw Err(-1, std::string(what) + " grid needs at least 3 nodes per axis");
        if (!(f->h > 0.0)) throw Err(-1, std::string(what) + " grid spacing must be positive");
        if (!f->rf) throw Err(-1, std::string(what) + " grid needs the RF basis");
        if (need_ac && !f->ac) throw Err(-1, std::string(what) + " grid needs the AC basis");
        if (3ull * (size_t)f->nx * f->ny * f->nz >= (1ull << 31)) throw Err(-1, std::string(what) + " grid too large");
        meta = *f;
        const size_t n3 = 3 * nodes() * sizeof(double);
        rf.ensure(n3);
        HIPCHK(hipMemcpyAsync(rf.p, f->rf, n3, hipMemcpyHostToDevice, st));
        if (f->ac) {
            has_ac = true;
            ac.ensure(n3);
            HIPCHK(hipMemcpyAsync(ac.p, f->ac, n3, hipMemcpyHostToDevice, st));
        }
        for (int b = 0; b < 4; ++b) {
            if (!f->dc[b]) continue;
            has_dc[b] = true;
            dc[b].ensure(n3);
            HIPCHK(hipMemcpyAsync(dc[b].p, f->dc[b], n3, hipMemcpyHostToDevice, st));
        }
        const size_t mwords = (nodes() + 3) / 4;  // the fill writes whole words
        mask.ensure(mwords * 4);
        if (f->mask) HIPCHK(hipMemcpyAsync(mask.p, f->mask, nodes(), hipMemcpyHostToDevice, st));
        else launch_fill_u32(mask.as<uint32_t>(), 0u, (uint32_t)mwords, st);
        meta.rf = meta.ac = nullptr;
        for (auto& p : meta.dc) p = nullptr;
        meta.mask = nullptr;
    }

    TGrid tgrid() const
    {
        TGrid g{};
        g.rf = rf.as<double>();
        g.ac = has_ac ? ac.as<double>() : nullptr;
        g.mask = mask.as<uint8_t>();
        g.inv_h = 1.0 / meta.h;
        g.lim0 = (double)(meta.nx - 1) * meta.h;
        g.lim1 = (double)(meta.ny - 1) * meta.h;
        g.lim2 = (double)(meta.nz - 1) * meta.h;
        g.nx = meta.nx; g.ny = meta.ny; g.nz = meta.nz;
        return g;
    }
};

// Combined DC field of one table and its FP32 cell records, cached by voltages.
struct Table {
    bool valid = false, cells_valid = false, prec_valid = false;
    double key[4] = {0, 0, 0, 0};
    DevBuf dcc, cells;
    DevBuf prec;  // FP64 cell records of the plain path
};

}  // namespace

struct mq_tandem {
    int device = 0, cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    Grid cube, port;
    mq_tandem_geom geom{};
    Table tab[4];  // splitter, combiner, driven port, idle port
    DevBuf jobs, ctr, tot, list[3], lctr, bin_of, rank_of, hist, bsum;
    DevBuf sx, sy, sz, sphi, svx, svy, svz, sfx, sfy, sfz, ssteps, sncoll, scd, sjob, sstatus, sregion, slast;
    mq_run_stats stats{};
    bool plain_recs = true;  // mq_tandem_set_plain_recs

    void create(const mq_fields* c, const mq_fields* p, const mq_tandem_geom* g)
    {
        if (!g) throw Err(-1, "geometry is NULL");
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
        cube.upload(c, false, "cube", stream);
        port.upload(p, true, "port", stream);
        geom = *g;
        const TGrid cg = cube.tgrid(), pg = port.tgrid();
        if (!(g->ap_r > 0.0)) throw Err(-1, "mouth aperture radius must be positive");
        if (!(g->seam_z > 0.0 && g->seam_z < cg.lim2)) throw Err(-1, "need 0 < seam_z < cube height");
        if (!(g->entrance[0] - g->ap_r > 0.0 && g->entrance[0] + g->ap_r < cg.lim0 && g->entrance[1] - g->ap_r > 0.0 &&
              g->entrance[1] + g->ap_r < cg.lim2))
            throw Err(-1, "the entrance aperture does not fit on the cube's y face");
        if (!(g->l_ext > 0.0)) throw Err(-1, "entrance tube length l_ext must be positive");
        for (int k = 0; k < MQ_TANDEM_PORTS; ++k) {
            const double mx = g->mouth[k][0], my = g->mouth[k][1];
            if (!(mx - g->ap_r > 0.0 && my - g->ap_r > 0.0 && mx + g->ap_r < cg.lim0 && my + g->ap_r < cg.lim1))
                throw Err(-1, "a mouth aperture does not fit inside the cube");
        }
        const mq_fields& pm = port.meta;
        if (!(pm.z_return > 0.0 && pm.z_return < pm.z_exit && pm.z_exit < pg.lim2))
            throw Err(-1, "port grid needs 0 < z_return < z_exit < its height");
        if (!(pm.axis_x - g->ap_r > 0.0 && pm.axis_y - g->ap_r > 0.0 && pm.axis_x + g->ap_r < pg.lim0 &&
              pm.axis_y + g->ap_r < pg.lim1))
            throw Err(-1, "the mouth aperture does not fit inside the port grid");
        HIPCHK(hipStreamSynchronize(stream));
    }

    void destroy()
    {
        for (DevBuf* b : {&cube.rf, &cube.ac, &cube.dc[0], &cube.dc[1], &cube.dc[2], &cube.dc[3], &cube.mask, &port.rf,
                          &port.ac, &port.dc[0], &port.dc[1], &port.dc[2], &port.dc[3], &port.mask, &tab[0].dcc,
                          &tab[0].cells, &tab[1].dcc, &tab[1].cells, &tab[2].dcc, &tab[2].cells, &tab[3].dcc,
                          &tab[3].cells, &jobs, &ctr, &tot, &list[0], &list[1], &list[2], &lctr, &bin_of, &rank_of, &hist,
                          &bsum, &sx, &sy, &sz, &sphi, &svx, &svy,
                          &svz, &sfx, &sfy, &sfz, &ssteps, &sncoll, &scd, &sjob, &sstatus, &sregion, &slast})
            b->release();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void prepare_table(int t, const double v[4], bool want_cells)
    {
        Grid& G = t < 2 ? cube : port;
        Table& T = tab[t];
        bool same = T.valid;
        for (int b = 0; b < 4; ++b) same = same && T.key[b] == v[b];
        if (!same) {
            T.dcc.ensure(3 * G.nodes() * sizeof(double));
            const double* d[4];
            double vv[4];
            for (int b = 0; b < 4; ++b) {
                d[b] = G.has_dc[b] ? G.dc[b].as<double>() : nullptr;
                vv[b] = v[b];
                T.key[b] = v[b];
            }
            launch_combine_dc(T.dcc.as<double>(), d, vv, (uint32_t)(3 * G.nodes()), stream);
            HIPCHK(hipGetLastError());
            T.valid = true;
            T.cells_valid = false;
            T.prec_valid = false;
        }
        if (want_cells && !T.cells_valid) {
            const size_t nc = (size_t)(G.meta.nx - 1) * (G.meta.ny - 1) * (G.meta.nz - 1);
            T.cells.ensure(nc * CELL_STRIDE_F4 * sizeof(float4));
            // the cube has no excitation: its AC slot is never weighted (a_ac = 0)
            const double* acp = G.has_ac ? G.ac.as<double>() : G.rf.as<double>();
            launch_build_cells(T.cells.as<float4>(), G.rf.as<double>(), T.dcc.as<double>(), acp,
                               G.mask.as<uint8_t>(), G.meta.nx, G.meta.ny, G.meta.nz, stream);
            HIPCHK(hipGetLastError());
            T.cells_valid = true;
        }
    }

    IonState state_buffers(size_t n)
    {
        sx.ensure(n * 8); sy.ensure(n * 8); sz.ensure(n * 8); sphi.ensure(n * 8);
        svx.ensure(n * 8); svy.ensure(n * 8); svz.ensure(n * 8);
        sfx.ensure(n * 4); sfy.ensure(n * 4); sfz.ensure(n * 4);
        ssteps.ensure(n * 4); sncoll.ensure(n * 4); scd.ensure(n * 4); sjob.ensure(n * 4);
        sstatus.ensure(n); sregion.ensure(n); slast.ensure(n);
        IonState S;
        S.x = sx.as<double>(); S.y = sy.as<double>(); S.z = sz.as<double>(); S.phi = sphi.as<double>();
        S.vx = svx.as<double>(); S.vy = svy.as<double>(); S.vz = svz.as<double>();
        S.fvx = sfx.as<float>(); S.fvy = sfy.as<float>(); S.fvz = sfz.as<float>();
        S.steps = ssteps.as<uint32_t>(); S.ncoll = sncoll.as<uint32_t>(); S.cd = scd.as<uint32_t>();
        S.job = sjob.as<uint32_t>(); S.status = sstatus.as<uint8_t>();
        S.region = sregion.as<uint8_t>(); S.last = slast.as<uint8_t>();
        return S;
    }

    TandemP tandem_params(const mq_tandem_drive& d, const mq_physics& ph) const
    {
        TandemP T{};
        T.dt = ph.dt;
        T.seam = geom.seam_z;
        T.ap_r2 = geom.ap_r * geom.ap_r;
        T.ent_x = geom.entrance[0];
        T.ent_z = geom.entrance[1];
        T.l_ext = geom.l_ext;
        T.z_ret = port.meta.z_return;
        T.z_exit = port.meta.z_exit;
        T.ax_p = port.meta.axis_x;
        T.ay_p = port.meta.axis_y;
        for (int k = 0; k < 9; ++k) {
            T.mx[k] = geom.mouth[k][0];
            T.my[k] = geom.mouth[k][1];
            T.w_ac[k] = d.tones[k] * ph.dt;
        }
        T.v_rf = d.v_rf;
        T.w_rf = d.f_rf * ph.dt;
        T.v_ac = d.v_ac;
        T.max_steps = max_steps_of(ph);
        T.nearest = ph.legacy_fields ? 1u : 0u;
        return T;
    }

    IonState run(const mq_job* jobs_h, size_t njobs, const mq_tandem_drive* d, const mq_physics* php, int path,
                 std::vector<unsigned long long>& ctr_h, std::vector<unsigned long long>& tot_h, size_t& n_total)
    {
        auto t_wall = std::chrono::steady_clock::now();
        if (!jobs_h || njobs == 0) throw Err(-1, "batch has no jobs");
        if (njobs > MQ_TANDEM_MAX_JOBS) throw Err(-1, "a tandem batch takes at most MQ_TANDEM_MAX_JOBS jobs");
        if (!d || !php) throw Err(-1, "NULL drive or physics");
        const mq_physics& ph = *php;
        if (!(ph.dt > 0.0)) throw Err(-1, "dt must be positive");
        if (ph.legacy_collisions || ph.legacy_criteria)
            throw Err(-1, "the tandem batch runs the thermal bath with geometric planes (no legacy collisions/criteria)");
        if (path != MQ_PATH_BATCHED && path != MQ_PATH_PLAIN) throw Err(-1, "unknown path");
        if (ph.rng != MQ_RNG_PHILOX) throw Err(-1, "the tandem batch uses the Philox streams (rng = 0)");
        for (int k = 0; k < 9; ++k)
            if (!(d->tones[k] >= 0.0) || !std::isfinite(d->tones[k])) throw Err(-1, "tones must be finite and >= 0");
        std::vector<JobP> jp(njobs);
        uint64_t off = 0;
        for (size_t k = 0; k < njobs; ++k) {
            if (jobs_h[k].dc_scale != 1.0) throw Err(-1, "dc_scale must be 1 in a tandem batch");
            jp[k] = job_params(jobs_h[k], ph, off);
            off += jobs_h[k].n;
        }
        if (off >= 0xFFFFFF00ull) throw Err(-1, "too many ions in one batch");
        const uint32_t n = (uint32_t)off;
        n_total = n;
        jobs.ensure(njobs * sizeof(JobP));
        HIPCHK(hipMemcpyAsync(jobs.p, jp.data(), njobs * sizeof(JobP), hipMemcpyHostToDevice, stream));
        const uint32_t nctr = 9 * TC_PORT;
        ctr.ensure(nctr * 8);
        tot.ensure((TT_N + 1) * 8);
        launch_fill_u64(ctr.as<unsigned long long>(), 0ull, nctr, stream);
        launch_fill_u64(tot.as<unsigned long long>(), 0ull, TT_N + 1, stream);
        const bool cells = path == MQ_PATH_BATCHED;
        double idle[4];
        for (int b = 0; b < 4; ++b) idle[b] = d->dc_port[b] + d->bias_idle[b];
        prepare_table(0, d->dc_split, cells);
        prepare_table(1, d->dc_combine, cells);
        prepare_table(2, d->dc_port, cells);
        prepare_table(3, idle, cells);
        IonState S = state_buffers(std::max<size_t>(n, 1));
        InjP ip{};
        ip.axis_x = cube.meta.axis_x; ip.axis_y = cube.meta.axis_y; ip.z_inject = cube.meta.z_inject;
        ip.temp_k = ph.temp_k; ip.r_inject_mm = ph.r_inject_mm; ip.inj_sigma_mm = ph.inj_sigma_mm; ip.dt = ph.dt;
        ip.model = ph.injection == MQ_INJ_GAUSSIAN ? 2 : (ph.legacy_injection ? 1 : 0);
        launch_inject(S, jobs.as<JobP>(), (uint32_t)njobs, ip, n, stream);
        launch_tandem_init(S, n, stream);
        HIPCHK(hipGetLastError());
        const TandemP T = tandem_params(*d, ph);
        stats = mq_run_stats{};
        stats.n_ions = n;
        float kms = 0.0f;
        if (n) {
            if (path == MQ_PATH_PLAIN) run_plain(S, T, n, kms);
            else run_batched(S, T, n, kms);
        }
        launch_tandem_totals(S, n, tot.as<unsigned long long>(), stream);
        HIPCHK(hipGetLastError());
        ctr_h.resize(nctr);
        tot_h.resize(TT_N + 1);
        HIPCHK(hipMemcpyAsync(ctr_h.data(), ctr.p, nctr * 8, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(tot_h.data(), tot.p, (TT_N + 1) * 8, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (tot_h[TT_N])
            throw Err(-5, "internal error: " + std::to_string(tot_h[TT_N]) + " ions never finished (status 0)");
        stats.ion_steps = tot_h[7];
        stats.kernel_ms = kms;
        stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall).count();
        return S;
    }

    // FP64 cell records of the four field sets for the plain path (as the single-channel engine's), when
    // plain_recs is on and all four fit in half of the free device memory; false: node tables.
    bool plain_records(const double2* rec[4])
    {
        if (!plain_recs) return false;
        size_t need[4], total = 0;
        for (int t = 0; t < 4; ++t) {
            const Grid& G = t < 2 ? cube : port;
            const uint64_t nc = (uint64_t)(G.meta.nx - 1) * (G.meta.ny - 1) * (G.meta.nz - 1);
            if (nc >= (1ull << 30)) return false;  // the kernel's cache key is 4 x cell + set
            need[t] = nc * PLAIN_REC_D2 * sizeof(double2);
            if (tab[t].prec.bytes < need[t]) total += need[t];
        }
        if (total) {
            size_t free_b = 0, total_b = 0;
            HIPCHK(hipMemGetInfo(&free_b, &total_b));
            if (total > free_b / 2) return false;
        }
        for (int t = 0; t < 4; ++t) {
            const Grid& G = t < 2 ? cube : port;
            Table& T = tab[t];
            if (T.prec.bytes < need[t]) {
                T.prec.ensure(need[t]);
                T.prec_valid = false;
            }
            if (!T.prec_valid) {
                // the cube has no excitation: its AC slot is never used
                const double* acp = G.has_ac ? G.ac.as<double>() : G.rf.as<double>();
                launch_build_plain_recs(T.prec.as<double2>(), G.rf.as<double>(), T.dcc.as<double>(), acp, G.meta.nx,
                                        G.meta.ny, G.meta.nz, stream);
                HIPCHK(hipGetLastError());
                T.prec_valid = true;
            }
            rec[t] = T.prec.as<double2>();
        }
        return true;
    }

    void run_plain(const IonState& S, const TandemP& T, uint32_t n, float& kms)
    {
        const double2* rec[4] = {nullptr, nullptr, nullptr, nullptr};
        const bool use_rec = plain_records(rec);
        stats.plain_recs = use_rec ? 1u : 0u;
        uint32_t nsteps = (uint32_t)std::max(1000.0, std::floor(2e-4 / T.dt));
        nsteps = (nsteps + 3u) & ~3u;
        lctr.ensure(256);
        uint32_t* live = lctr.as<uint32_t>();
        const double* dc[4] = {tab[0].dcc.as<double>(), tab[1].dcc.as<double>(), tab[2].dcc.as<double>(),
                               tab[3].dcc.as<double>()};
        const TGrid cg = cube.tgrid(), pg = port.tgrid();
        for (;;) {
            launch_fill_u32(live, 0u, 1, stream);
            HIPCHK(hipEventRecord(ev0, stream));
            launch_tandem_plain(S, jobs.as<JobP>(), T, cg, pg, dc, use_rec ? rec : nullptr,
                                ctr.as<unsigned long long>(), n, nsteps, live, stream);
            HIPCHK(hipEventRecord(ev1, stream));
            HIPCHK(hipGetLastError());
            uint32_t alive = 0;
            HIPCHK(hipMemcpyAsync(&alive, live, 4, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
            kms += ms;
            stats.launches += 1;
            if (alive == 0) break;
        }
    }

    // Persistent-worker launches as in the single-channel engine: before every
    // launch the live list is sorted by (cell table, 4 x 4 x 4-cell bin) and cut into
    // one segment per XCD, and each launch advances every ion by up to `chunk`
    // steps. Once no more ions are live than the launch has lanes, the re-sort is
    // dropped and the chunk doubles (tiled.rs:26-27, 214).
    void run_batched(const IonState& S, const TandemP& T, uint32_t n, float& kms)
    {
        for (int k = 0; k < 3; ++k) list[k].ensure((size_t)n * 4);
        lctr.ensure(2 * CTR_BLOCK * 4);
        const TGrid cg = cube.tgrid(), pg = port.tgrid();
        auto nbins = [](const TGrid& g) {
            return (uint32_t)(((g.nx - 2) >> 2) + 1) * (uint32_t)(((g.ny - 2) >> 2) + 1) * (uint32_t)(((g.nz - 2) >> 2) + 1);
        };
        const uint32_t nbc = nbins(cg), nbp = nbins(pg), nb = 2 * nbc + 2 * nbp;
        if (nb > 1024u * 1024u) throw Err(-1, "tandem grids too large for the re-sort bins");
        bin_of.ensure((size_t)n * 4);
        rank_of.ensure((size_t)n * 4);
        hist.ensure((size_t)nb * 4);
        bsum.ensure(1024 * 4);
        uint32_t* c0 = lctr.as<uint32_t>();
        uint32_t* c1 = c0 + CTR_BLOCK;
        launch_iota(list[0].as<uint32_t>(), n, stream);
        launch_fill_u32(c0, 0u, 2 * CTR_BLOCK, stream);
        launch_fill_u32(c0 + CTR_COUNT, n, 1, stream);
        const uint32_t block = 256;
        const uint32_t grid = std::max<uint32_t>(1, (uint32_t)((uint64_t)cus * 8 * 64 / block));
        const uint64_t lanes = (uint64_t)grid * block;
        const float4* cells[4] = {tab[0].cells.as<float4>(), tab[1].cells.as<float4>(), tab[2].cells.as<float4>(),
                                  tab[3].cells.as<float4>()};
        uint32_t chunk = 1024;   // measured on the 1M-ion batch: 256 and 1024 within 1.5 %, 4096 7 % slower
        bool tail = false;
        int cur = 0;
        for (;;) {
            uint32_t* cin = cur ? c1 : c0;
            uint32_t* cout = cur ? c0 : c1;
            const uint32_t* lin = list[cur].as<uint32_t>();
            if (!tail) {
                launch_fill_hist(hist.as<uint32_t>(), nb, stream);
                launch_tandem_bin(S, T, cg, pg, lin, cin, bin_of.as<uint32_t>(), rank_of.as<uint32_t>(),
                                  hist.as<uint32_t>(), nbc, nbp, n, stream);
                launch_sort_place(lin, cin, list[2].as<uint32_t>(), bin_of.as<uint32_t>(), rank_of.as<uint32_t>(),
                                  hist.as<uint32_t>(), bsum.as<uint32_t>(), nb, n, stream);
                lin = list[2].as<uint32_t>();
            }
            WorkerArgs pa{};
            pa.ctr_in = cin;
            pa.ctr_out = cout;
            launch_prepare(pa, stream);
            HIPCHK(hipEventRecord(ev0, stream));
            launch_tandem_batched(S, jobs.as<JobP>(), T, cg, pg, cells, ctr.as<unsigned long long>(), lin,
                                  list[cur ^ 1].as<uint32_t>(), cin, cout, chunk, 8u, tail ? 1u : std::min<uint32_t>(MAX_QUEUES, grid), grid,
                                  block, stream);
            HIPCHK(hipEventRecord(ev1, stream));
            HIPCHK(hipGetLastError());
            uint32_t alive = 0;
            HIPCHK(hipMemcpyAsync(&alive, cout + CTR_COUNT, 4, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
            kms += ms;
            stats.launches += 1;
            cur ^= 1;
            if (alive == 0) break;
            if (!tail && alive <= lanes) tail = true;
            if (tail) chunk = std::min<uint32_t>(chunk * 2u, 1u << 20);
        }
    }
};

namespace {

void fill_counts(const std::vector<unsigned long long>& c, const std::vector<unsigned long long>& t,
                 mq_port_counts* counts, mq_tandem_totals* totals)
{
    if (counts) {
        for (int p = 0; p < MQ_TANDEM_PORTS; ++p) {
            const unsigned long long* a = &c[(size_t)p * TC_PORT];
            mq_port_counts& o = counts[p];
            o.entered = a[TC_ENTERED]; o.returned = a[TC_RETURNED]; o.forwarded = a[TC_FORWARDED];
            o.revisits = a[TC_REVISITS]; o.from_other = a[TC_FROM_OTHER]; o.reverse_entries = a[TC_REVERSE];
            o.lost = a[TC_LOST];
            for (int k = 0; k < MQ_TANDEM_PORTS; ++k) o.retry_from[k] = a[TC_RETRY + k];
            for (int k = 0; k < MQ_TANDEM_MAX_JOBS; ++k) o.by_job[k] = a[TC_BYJOB + k];
        }
    }
    if (totals) {
        totals->n = t[0]; totals->split = t[1]; totals->channel = t[2]; totals->combine = t[3];
        totals->lost = t[4]; totals->back = t[5]; totals->detected = t[6]; totals->ion_steps = t[7];
    }
}

int fail(const Err& e)
{
    mq_set_last_error(e.what());
    return e.code;
}

int fail(const std::exception& e)
{
    mq_set_last_error(e.what());
    return -9;
}

}  // namespace

static_assert(MQ_TANDEM_PORTS == 9 && MQ_TANDEM_MAX_JOBS == TC_MAX_JOBS, "counter layout");

extern "C" {

int mq_tandem_create(const mq_fields* cube, const mq_fields* port, const mq_tandem_geom* geom, mq_tandem** out)
{
    mq_tandem* t = nullptr;
    try {
        mq_set_last_error("");
        if (!out) throw Err(-1, "out is NULL");
        t = new mq_tandem();
        t->create(cube, port, geom);
        *out = t;
        return 0;
    } catch (const Err& e) {
        if (t) { t->destroy(); delete t; }
        return fail(e);
    } catch (const std::exception& e) {
        if (t) { t->destroy(); delete t; }
        return fail(e);
    }
}

void mq_tandem_destroy(mq_tandem* t)
{
    if (!t) return;
    t->destroy();
    delete t;
}

int mq_tandem_run(mq_tandem* t, const mq_job* jobs, size_t njobs, const mq_tandem_drive* d, const mq_physics* ph,
                  int32_t path, mq_port_counts* counts, mq_tandem_totals* totals)
{
    return mq_tandem_run_ions(t, jobs, njobs, d, ph, path, counts, totals, nullptr, nullptr, nullptr, nullptr);
}

int mq_tandem_run_ions(mq_tandem* t, const mq_job* jobs, size_t njobs, const mq_tandem_drive* d,
                       const mq_physics* ph, int32_t path, mq_port_counts* counts, mq_tandem_totals* totals,
                       const mq_ion_soa* s, uint8_t* region, uint8_t* last, uint32_t* job)
{
    try {
        mq_set_last_error("");
        if (!t) throw Err(-1, "tandem is NULL");
        std::vector<unsigned long long> c, tt;
        size_t n = 0;
        IonState S = t->run(jobs, njobs, d, ph, path, c, tt, n);
        fill_counts(c, tt, counts, totals);
        if (n == 0) return 0;
        hipStream_t st = t->stream;
        auto d2h = [&](void* dst, const void* src, size_t bytes) {
            if (dst) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
        };
        d2h(region, S.region, n);
        d2h(last, S.last, n);
        d2h(job, S.job, n * 4);
        if (s) {
            d2h(s->x, S.x, n * 8); d2h(s->y, S.y, n * 8); d2h(s->z, S.z, n * 8);
            d2h(s->steps, S.steps, n * 4); d2h(s->ncoll, S.ncoll, n * 4); d2h(s->status, S.status, n);
            d2h(s->phi_ac0, S.phi, n * 8); d2h(s->cd, S.cd, n * 4);
            if (path == MQ_PATH_PLAIN) {
                d2h(s->vx, S.vx, n * 8); d2h(s->vy, S.vy, n * 8); d2h(s->vz, S.vz, n * 8);
            } else {
                std::vector<float> f(n);
                double* dsts[3] = {s->vx, s->vy, s->vz};
                float* srcs[3] = {S.fvx, S.fvy, S.fvz};
                for (int k = 0; k < 3; ++k) {
                    if (!dsts[k]) continue;
                    HIPCHK(hipMemcpyAsync(f.data(), srcs[k], n * 4, hipMemcpyDeviceToHost, st));
                    HIPCHK(hipStreamSynchronize(st));
                    for (size_t i = 0; i < n; ++i) dsts[k][i] = (double)f[i];
                }
            }
        }
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    } catch (const Err& e) {
        return fail(e);
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int mq_tandem_last_stats(const mq_tandem* t, mq_run_stats* out)
{
    if (!t || !out) { mq_set_last_error("NULL argument"); return -1; }
    *out = t->stats;
    return 0;
}

int mq_tandem_set_plain_recs(mq_tandem* t, int32_t on)
{
    if (!t) { mq_set_last_error("tandem is NULL"); return -1; }
    if (on < 0 || on > 1) { mq_set_last_error("plain_recs must be 0 or 1"); return -1; }
    t->plain_recs = on != 0;
    mq_set_last_error("");
    return 0;
}

}  // extern "C"
