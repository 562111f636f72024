// fp32_cells.h -- FP32 cell-record access shared by the batched and tandem
// workers: record load into VGPRs, polynomial evaluation, hardware
// transcendental functions in revolutions.
#pragma once

#include <hip/hip_runtime.h>

#include "mq_types.h"

namespace mq {

__device__ __forceinline__ float sin_rev(float p) { return __builtin_amdgcn_sinf(p); }
__device__ __forceinline__ float cos_rev(float p) { return __builtin_amdgcn_cosf(p); }
__device__ __forceinline__ float log_f(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945f; }

__device__ __forceinline__ void load_cell(const float4* __restrict__ cells, int32_t c, float4 R[CELL_F4])
{
    const float4* src = cells + (size_t)c * CELL_STRIDE_F4;
#pragma unroll
    for (int q = 0; q < CELL_F4; ++q) R[q] = src[q];
}

// Rank of live-list entry i within its re-sort bin: hist[bin] counts the bin's
// entries and rank_of[i] is this entry's place among them. Lanes of a wavefront
// that share a bin take their ranks with one atomic: the bins are peeled from
// the lowest pending lane (freshly injected ions all start in a few bins, and
// one atomic per ion serialised on those counters), then the lowest lane of
// each bin adds its bin's count, all bins in one instruction, so the wavefront
// waits for a single round trip of returning atomics.
// All active lanes of the wavefront must call this together.
__device__ __forceinline__ void bin_rank(uint32_t* __restrict__ hist, uint32_t bin, uint32_t i,
                                         uint32_t* __restrict__ rank_of)
{
    const uint32_t lane = threadIdx.x & 63;
    uint64_t todo = __ballot(1);
    int first = (int)lane;  // lowest lane of this lane's bin
    uint32_t below = 0, count = 0;
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const uint32_t lb = (uint32_t)__shfl((int)bin, leader);
        const uint64_t same = __ballot(bin == lb) & todo;
        if ((same >> lane) & 1ull) {
            first = leader;
            below = (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        }
        if (lane == (uint32_t)leader) count = (uint32_t)__popcll(same);
        todo &= ~same;
    }
    uint32_t base = 0;
    if (count) base = atomicAdd(&hist[bin], count);
    base = (uint32_t)__shfl((int)base, first);
    rank_of[i] = base + below;
}

// Record layout (component-major): float 24 d + 8 q + k holds coefficient k of
// component d (x, y, z) of basis q (RF, combined DC, AC), so component d is the
// six float4 R[6 d .. 6 d + 5] and basis q within it the pair R[6 d + 2 q],
// R[6 d + 2 q + 1]; the corner conductor bits are R[18].x.
__device__ __forceinline__ float poly8(const float4 lo, const float4 hi, float tx, float ty, float tz, float pxy,
                                       float pyz, float pxz, float pxyz)
{
    float f = lo.x;
    f = fmaf(tx, lo.y, f);
    f = fmaf(ty, lo.z, f);
    f = fmaf(tz, lo.w, f);
    f = fmaf(pxy, hi.x, f);
    f = fmaf(pyz, hi.y, f);
    f = fmaf(pxz, hi.z, f);
    f = fmaf(pxyz, hi.w, f);
    return f;
}

// Component d of basis q of the record at (tx, ty, tz).
__device__ __forceinline__ float poly(const float4 R[CELL_F4], int q, int d, float tx, float ty, float tz, float pxy,
                                      float pyz, float pxz, float pxyz)
{
    return poly8(R[6 * d + 2 * q], R[6 * d + 2 * q + 1], tx, ty, tz, pxy, pyz, pxz, pxyz);
}

}  // namespace mq
