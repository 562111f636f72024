// kernels.h -- host launchers of the HIP kernels (one stream per call).
#pragma once

#include <hip/hip_runtime.h>

#include "mq_types.h"

namespace mq {

void launch_inject(const IonState& S, const JobP* jobs, uint32_t njobs, const InjP& P, uint32_t n, hipStream_t st);
void launch_combine_dc(double* out, const double* const d[4], const double v[4], uint32_t n3, hipStream_t st);
// FP64 plain step. rec != nullptr: per-cell FP64 records of the 8 corner node
// values of the three bases (PLAIN_REC_D2 double2 per cell, [basis][corner]
// [component], built by launch_build_plain_recs from the same tables), cached
// per lane; the results are those of the node tables bit for bit.
constexpr int PLAIN_REC_D2 = 36;
void launch_plain(const IonState& S, const JobP* jobs, const BatchP& B, const double* rf, const double* dc,
                  const double* ac, const double2* rec, const uint8_t* mask, uint32_t n, uint32_t nsteps,
                  uint32_t* live, hipStream_t st);
void launch_build_plain_recs(double2* rec, const double* rf, const double* dc, const double* ac, int nx, int ny,
                             int nz, hipStream_t st);
void launch_summarize(const IonState& S, unsigned long long* acc, uint32_t n, hipStream_t st);
void launch_fill_u64(unsigned long long* p, unsigned long long v, uint32_t n, hipStream_t st);
void launch_fill_u32(uint32_t* p, uint32_t v, uint32_t n, hipStream_t st);

// Live population (mq_live_*): one new ion: its slot, job (species) and random-stream number.
struct LiveNew {
    uint32_t slot, job, sid, pad;
    double pha;  // AC phase of the ion's tone at its birth [cycles]
};
void launch_live_inject(const IonState& S, const JobP* jobs, const InjP& P, const LiveNew* rec, uint32_t k, double phr,
                        uint32_t* list, uint32_t* count, hipStream_t st);
void launch_live_collect(const IonState& S, const uint32_t* list, const uint32_t* count, uint32_t nmax, uint32_t* out,
                         uint32_t* nout, hipStream_t st);
void launch_live_rephase(const IonState& S, const uint32_t* list, const uint32_t* count, uint32_t nmax, double pha,
                         double w_ac, uint32_t job_end, hipStream_t st);

// Batched path.
void launch_build_cells(float4* cells, const double* rf, const double* dc, const double* ac, const uint8_t* mask,
                        int nx, int ny, int nz, hipStream_t st);
void launch_iota(uint32_t* list, uint32_t n, hipStream_t st);

// Counter block of one live list: queue heads at [q * CTR_Q], q < MAX_QUEUES,
// and the list length at [CTR_COUNT]; CTR_BLOCK u32 per block.
constexpr uint32_t MAX_QUEUES = 8;
constexpr uint32_t CTR_Q = 32;
constexpr uint32_t CTR_COUNT = MAX_QUEUES * CTR_Q;
constexpr uint32_t CTR_BLOCK = CTR_COUNT + 32;

// Coarse cell bins of the re-sort: bin = ((k >> sz) nby + (j >> sy)) nbx + (i >> sx).
struct SortGeom {
    uint32_t sx, sy, sz, nbx, nby, nbz, nbins;
};

struct WorkerArgs {
    IonState S;
    const JobP* jobs;
    BatchP B;
    const float4* cells;
    const uint32_t* list_in;
    uint32_t* list_out;
    uint32_t* ctr_in;   // heads + count of list_in
    uint32_t* ctr_out;  // count of list_out (ions carried to the next launch)
    uint32_t chunk;     // max steps per ion in this launch (multiple of 4)
    uint32_t period;    // steps between refills (multiple of 4)
    uint32_t nqueues;   // 1, or MAX_QUEUES contiguous segments (one per XCD)
    uint32_t coll_batch;  // pending collisions per wavefront that trigger the collision code
    uint32_t* bin_out;    // optional: re-sort bin of each carried ion, at its list_out position
    uint32_t* rank_out;   // with bin_out: the ion's rank among the carried ions of its bin (at the same position)
    uint32_t* hist_out;   // with rank_out: carried ions per bin (zeroed before the launch)
    SortGeom sg;          // bins of bin_out
    uint32_t lanes;       // lanes per wavefront that take ions (1..64; fewer spread a small tail over more waves)
};

// ranked: the worker launch that wrote the list has already put each entry's
// bin and rank in bin_of / rank_of and the counts per bin in hist, so only
// the offsets and the scatter remain; otherwise k_bin gathers the positions.
void launch_sort(const IonState& S, const BatchP& B, const uint32_t* list, const uint32_t* ctr, uint32_t* out,
                 uint32_t* bin_of, uint32_t* rank_of, uint32_t* hist, uint32_t* bsum, const SortGeom& G,
                 uint32_t nmax, bool ranked, hipStream_t st);
// Pieces of launch_sort for callers with their own binning kernel: clear the
// histogram; then, after bin_of / rank_of / hist are filled, turn counts into
// offsets and place every entry.
void launch_fill_hist(uint32_t* hist, uint32_t nb, hipStream_t st);
void launch_sort_place(const uint32_t* list, const uint32_t* ctr, uint32_t* out, const uint32_t* bin_of,
                       const uint32_t* rank_of, uint32_t* hist, uint32_t* bsum, uint32_t nb, uint32_t nmax, hipStream_t st);
void launch_prepare(const WorkerArgs& a, hipStream_t st);
void launch_gather_probe(const float4* cells, uint64_t window, uint32_t iters, float* out, uint32_t grid,
                         uint32_t block, hipStream_t st);
void launch_batched(const WorkerArgs& a, uint32_t grid, uint32_t block, hipStream_t st);
// The tail worker: four lanes per ion, 16 ions per wavefront (a.lanes quads of each
// wavefront take ions); Philox streams, one queue (a.nqueues 1), refill period at most 8, no re-sort bins.
void launch_tail4(const WorkerArgs& a, uint32_t grid, hipStream_t st);
// The staged-tile worker (Philox streams): 256-thread workgroups, each taking 256
// list entries per chunk; the counter ctr_in[0] hands out the chunks.
void launch_staged(const WorkerArgs& a, uint32_t grid, hipStream_t st);
void launch_field_probe(const float4* cells, const BatchP& B, uint32_t n, const double* x, const double* y,
                        const double* z, float a_rf, float s_dc, float a_ac, float* ex, float* ey, float* ez,
                        uint8_t* crash, hipStream_t st);
// Differentiable form (kernels_adjoint.hip).
// The node tables of the differentiable form (double), and per-cell float
// records of the corner values when every value of all three tables is exactly
// a float (rec != nullptr): ADJ_REC_F4 float4 per cell, 72 floats.
constexpr int ADJ_REC_F4 = 18;
struct AdjGrids {
    const double *rf, *dc, *ac;
    const float4* rec;
};
struct AdjObj {
    double z_exit, scale, r_leak, w_leak, axis_x, axis_y;
    uint32_t nsteps, ckpt;
    uint32_t gas;  // 1: thermal collisions with frozen draws (seg[3..5] allocated)
    uint32_t full; // 1: one segment of nsteps (ckpt = nsteps) that the forward itself fills, so the backward
                   // walks the stored trajectory without recomputing it
};
struct AdjBuf {
    double* ck[6];   // checkpoints (x, y, z, vx, vy, vz) [slot * n + ion]
    double* seg[6];  // positions of one segment [k * n + ion] (recomputed, or with `full` the whole trajectory
                     // written by the forward); with gas also the post-kick velocities
    double* fin[3];  // final position
    double* Ji;      // objective per ion
    double* gp[5];   // per-ion dJ/d(V_rf, f_rf, V_ac, f_ac, s_dc)
    double* gs[6];   // per-ion dJ/d(initial x, y, z, vx, vy, vz)
};
void launch_adj_forward(const IonState& S, const JobP* jobs, const BatchP& B, const AdjGrids& G, const AdjObj& O,
                        const AdjBuf& A, uint32_t n, hipStream_t st);
void launch_adj_backward(const IonState& S, const JobP* jobs, const BatchP& B, const AdjGrids& G, const AdjObj& O,
                         const AdjBuf& A, uint32_t n, double scale, hipStream_t st);
void launch_colsum(const double* const* cols_dev, uint32_t ncols, uint32_t n, double* part, double* out, hipStream_t st);
void launch_exact_f32(const double* a, uint64_t n, uint32_t* flag, hipStream_t st);
void launch_build_adj_recs(float4* rec, const double* rf, const double* dc, const double* ac, int nx, int ny, int nz,
                           hipStream_t st);

}  // namespace mq
