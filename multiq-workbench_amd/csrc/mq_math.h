// mq_math.h -- counter-based RNG and elementary functions shared by the HIP
// kernels and the host-side code of the engine.
//
// FP64 routines are written with explicit fma() and compiled with
// -ffp-contract=off, so host and device give identical bits; sqrt and division
// are IEEE correctly rounded on both. FP32 routines use the gfx950 hardware
// transcendental instructions (v_sin_f32 / v_cos_f32 take revolutions).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MQ_HD __host__ __device__ __forceinline__

namespace mq {

enum : uint32_t { D_CAND = 0, D_COLL = 1, D_INJ = 2 };

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 10 rounds of the
// multiply-xor bijection with multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key
// increments 0x9E3779B9 / 0xBB67AE85.
struct u32x4 { uint32_t a, b, c, d; };

MQ_HD u32x4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return u32x4{c0, c1, c2, c3};
}

MQ_HD uint32_t pick(const u32x4& w, uint32_t i)
{
    return i == 0 ? w.a : (i == 1 ? w.b : (i == 2 ? w.c : w.d));
}

// xoshiro128** (Blackman and Vigna, "Scrambled linear pseudorandom number
// generators", 2018): the per-ion stream of the reference's kernel
// (next_u32, channel.metal:38-42). State s[0..3] = uint4 (x, y, z, w).
MQ_HD uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
MQ_HD uint32_t xo_next(uint32_t& s0, uint32_t& s1, uint32_t& s2, uint32_t& s3)
{
    const uint32_t r = rotl32(s1 * 5u, 7) * 9u;
    const uint32_t t = s1 << 9;
    s2 ^= s0;
    s3 ^= s1;
    s1 ^= s2;
    s0 ^= s3;
    s2 ^= t;
    s3 = rotl32(s3, 11);
    return r;
}
// rand01 (channel.metal:43): top 24 bits, (w >> 8) 2^-24 in [0, 1) (exact in FP32 and FP64).
MQ_HD float xo_u01f(uint32_t w) { return (float)(w >> 8) * (1.0f / 16777216.0f); }
MQ_HD double xo_u01(uint32_t w) { return (double)(w >> 8) * (1.0 / 16777216.0); }

// Uniform in (0,1) from one 32-bit word: (w + 1/2) 2^-32 (exact in FP64).
MQ_HD double u01(uint32_t w) { return ((double)w + 0.5) * 0x1p-32; }
// FP32 uniform in (0,1): ((w >> 9) + 1/2) 2^-23 (exact in FP32).
MQ_HD float u01f(uint32_t w) { return ((float)(w >> 9) + 0.5f) * 0x1p-23f; }

// ---------------------------------------------------------------- FP64
// Taylor coefficients (-1)^k (pi/2)^(2k+1)/(2k+1)! and (-1)^k (pi/2)^(2k)/(2k)!
// rounded to nearest double; |r| <= 1/2 quarter turn.
#define MQ_S0 0x1.921fb54442d18p+0
#define MQ_S1 -0x1.4abbce625be53p-1
#define MQ_S2 0x1.466bc6775aae2p-4
#define MQ_S3 -0x1.32d2cce62bd86p-8
#define MQ_S4 0x1.50783487ee782p-13
#define MQ_S5 -0x1.e3074fde8871fp-19
#define MQ_S6 0x1.e8f434d018d63p-25
#define MQ_S7 -0x1.6fadb9f155744p-31
#define MQ_S8 0x1.aaec32af93359p-38
#define MQ_S9 -0x1.8a404211f9547p-45
#define MQ_C0 0x1.0000000000000p+0
#define MQ_C1 -0x1.3bd3cc9be45dep+0
#define MQ_C2 0x1.03c1f081b5ac4p-2
#define MQ_C3 -0x1.55d3c7e3cbffap-6
#define MQ_C4 0x1.e1f506891babbp-11
#define MQ_C5 -0x1.a6d1f2a204a8cp-16
#define MQ_C6 0x1.f9d38a3763cc3p-22
#define MQ_C7 -0x1.b6e24f44b128fp-28
#define MQ_C8 0x1.20c62c2f2d7f5p-34
#define MQ_C9 -0x1.2a0c591af8314p-41

MQ_HD void quarter(double p, double& r, int& k)
{
    double q = p * 4.0;
    double kk = rint(q);
    r = q - kk;
    k = (int)((int64_t)kk & 3);
}

MQ_HD double sin_q(double r, double r2)
{
    double ps = MQ_S9;
    ps = fma(ps, r2, MQ_S8); ps = fma(ps, r2, MQ_S7); ps = fma(ps, r2, MQ_S6);
    ps = fma(ps, r2, MQ_S5); ps = fma(ps, r2, MQ_S4); ps = fma(ps, r2, MQ_S3);
    ps = fma(ps, r2, MQ_S2); ps = fma(ps, r2, MQ_S1); ps = fma(ps, r2, MQ_S0);
    return ps * r;
}

MQ_HD double cos_q(double r2)
{
    double pc = MQ_C9;
    pc = fma(pc, r2, MQ_C8); pc = fma(pc, r2, MQ_C7); pc = fma(pc, r2, MQ_C6);
    pc = fma(pc, r2, MQ_C5); pc = fma(pc, r2, MQ_C4); pc = fma(pc, r2, MQ_C3);
    pc = fma(pc, r2, MQ_C2); pc = fma(pc, r2, MQ_C1); pc = fma(pc, r2, MQ_C0);
    return pc;
}

// sin(2 pi p) and cos(2 pi p).
MQ_HD void sincos2pi(double p, double& s, double& c)
{
    double r; int k;
    quarter(p, r, k);
    double r2 = r * r;
    double sn = sin_q(r, r2), cs = cos_q(r2);
    switch (k) {
    case 0: s = sn; c = cs; break;
    case 1: s = cs; c = -sn; break;
    case 2: s = -sn; c = -cs; break;
    default: s = -cs; c = sn; break;
    }
}

MQ_HD double sin2pi(double p)
{
    double r; int k;
    quarter(p, r, k);
    double r2 = r * r;
    double v = (k & 1) ? cos_q(r2) : sin_q(r, r2);
    return (k & 2) ? -v : v;
}

// The same values (bit for bit) with both polynomials evaluated and the
// quadrant k applied by selects and sign flips: (s, c) = (sn, cs), (cs, -sn),
// (-sn, -cs), (-cs, sn). Lanes of a wavefront in different quadrants then take
// no divergent branches; where the quadrant is mostly uniform across a
// wavefront the forms above, which evaluate one polynomial, are cheaper.
MQ_HD void sincos2pi_sel(double p, double& s, double& c)
{
    double r; int k;
    quarter(p, r, k);
    const double r2 = r * r;
    const double sn = sin_q(r, r2), cs = cos_q(r2);
    const bool odd = (k & 1) != 0;
    const double s0 = odd ? cs : sn, c0 = odd ? sn : cs;
    s = (k & 2) ? -s0 : s0;
    c = ((k + 1) & 2) ? -c0 : c0;
}

MQ_HD double sin2pi_sel(double p)
{
    double r; int k;
    quarter(p, r, k);
    const double r2 = r * r;
    const double sn = sin_q(r, r2), cs = cos_q(r2);
    const double v = (k & 1) ? cs : sn;
    return (k & 2) ? -v : v;
}

// Natural log of a positive normal double: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh(s), s = (m-1)/(m+1), series 2 s (1 + s^2/3 + ... + s^24/25).
MQ_HD double log_d(double x)
{
    uint64_t b = (uint64_t)__builtin_bit_cast(uint64_t, x);
    int e = (int)((b >> 52) & 0x7ff) - 1022;
    b = (b & 0x000fffffffffffffULL) | 0x3fe0000000000000ULL;
    double m = __builtin_bit_cast(double, b);
    if (m < 0x1.6a09e667f3bcdp-1) { m = m * 2.0; e -= 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double P = 0x1.47ae147ae147bp-5;
    P = fma(P, z, 0x1.642c8590b2164p-5); P = fma(P, z, 0x1.8618618618618p-5);
    P = fma(P, z, 0x1.af286bca1af28p-5); P = fma(P, z, 0x1.e1e1e1e1e1e1ep-5);
    P = fma(P, z, 0x1.1111111111111p-4); P = fma(P, z, 0x1.3b13b13b13b14p-4);
    P = fma(P, z, 0x1.745d1745d1746p-4); P = fma(P, z, 0x1.c71c71c71c71cp-4);
    P = fma(P, z, 0x1.2492492492492p-3); P = fma(P, z, 0x1.999999999999ap-3);
    P = fma(P, z, 0x1.5555555555555p-2);
    double t = s * (z * P);
    double lm = 2.0 * (s + t);
    double de = (double)e;
    return de * 0x1.62e4200000000p-1 + (de * 0x1.fdf473de6af28p-22 + lm);
}

MQ_HD double frac_d(double p) { return p - floor(p); }
MQ_HD double lerp_d(double a, double b, double t) { return fma(t, b - a, a); }

// Legacy collision countdown from a floored step count (clamped to [1, 2^31-1]).
MQ_HD uint32_t countdown(double x)
{
    if (!(x < 2147483647.0)) return 2147483647u;
    if (x < 1.0) return 1u;
    return (uint32_t)x;
}

}  // namespace mq
