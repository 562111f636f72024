// fp64_sample.h -- FP64 field sampling shared by the plain and tandem kernels:
// cell location with clamping (locate, channel.metal:55-61) and the
// trilinear lerp chain of interp (channel.metal:63-68).
#pragma once

#include <hip/hip_runtime.h>

#include "mq_math.h"

namespace mq {

__device__ __forceinline__ double norm3(double a, double b, double c) { return sqrt((a * a + b * b) + c * c); }

// ------------------------------------------------------ FP64 sampling
__device__ __forceinline__ void loc_d(double g, int n, int& i, double& t, int nearest)
{
    double gc = g;
    if (!(gc >= 0.0)) gc = 0.0; /* also maps NaN to 0 */
    if (gc > (double)(n - 1)) gc = (double)(n - 1);
    int ii = (int)gc;
    if (ii > n - 2) ii = n - 2;
    i = ii;
    t = nearest ? 0.0 : gc - (double)ii;
}

// Trilinear interpolation with the corner order and lerp chain of `interp`
// (channel.metal:63-68): x first, then y, then z.
__device__ __forceinline__ void sample_d(const double* __restrict__ F, uint32_t b, uint32_t sy, uint32_t sz,
                                         double tx, double ty, double tz, double out[3])
{
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double c00 = lerp_d(F[3 * b + d], F[3 * (b + 1) + d], tx);
        double c10 = lerp_d(F[3 * (b + sy) + d], F[3 * (b + 1 + sy) + d], tx);
        double c01 = lerp_d(F[3 * (b + sz) + d], F[3 * (b + 1 + sz) + d], tx);
        double c11 = lerp_d(F[3 * (b + sy + sz) + d], F[3 * (b + 1 + sy + sz) + d], tx);
        out[d] = lerp_d(lerp_d(c00, c10, ty), lerp_d(c01, c11, ty), tz);
    }
}

// The three basis fields at (tx, ty, tz) in a cell, in the corner order and
// lerp chain of sample_d, from a per-cell FP64 record of the 8 corner node
// values, [basis][corner][component] (72 doubles).
__device__ __forceinline__ void sample_rec(const double v[72], double tx, double ty, double tz, double er[3],
                                           double ed[3], double ea[3])
{
    double* out[3] = {er, ed, ea};
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double* f = v + 24 * q + d;
            const double c00 = lerp_d(f[0], f[3], tx), c10 = lerp_d(f[6], f[9], tx);
            const double c01 = lerp_d(f[12], f[15], tx), c11 = lerp_d(f[18], f[21], tx);
            out[q][d] = lerp_d(lerp_d(c00, c10, ty), lerp_d(c01, c11, ty), tz);
        }
}

}  // namespace mq
