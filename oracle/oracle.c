/*
 * oracle.c -- CPU FP64 oracle of the batched ion-trajectory integrator.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the HIP path and
 * the timed CPU baseline. Build with -O2 -ffp-contract=off (oracle/Makefile):
 * every fused multiply-add below is an explicit fma() call, every other
 * product and sum is rounded separately, so the GPU FP64 path, which follows
 * the same operation order, is bit-identical.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#define AMU 1.6605390666e-27
#define QE 1.602176634e-19
#define KB 1.380649e-23

enum { D_CAND = 0, D_COLL = 1, D_INJ = 2 };

/* ------------------------------------------------------------------ Philox */

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
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
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void blk(const uint32_t key[2], uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t w[4])
{
    uint32_t ctr[4] = {a, b, c, d};
    or_philox4x32_10(ctr, key, w);
}

/* Uniform in (0,1): (w + 1/2) 2^-32, exact in FP64. */
static inline double u01(uint32_t w) { return ((double)w + 0.5) * 0x1p-32; }

/* -------------------------------------------------- sin/cos/log (own code) */
/* sin(pi r/2) = r * sum S_k r^2k, cos(pi r/2) = sum C_k r^2k for |r| <= 1/2:
 * Taylor coefficients (-1)^k (pi/2)^(2k+1)/(2k+1)! and (-1)^k (pi/2)^(2k)/(2k)!
 * rounded to nearest double (truncation error < 1e-19 on |r| <= 1/2). */
static const double S_[10] = {
    0x1.921fb54442d18p+0, -0x1.4abbce625be53p-1, 0x1.466bc6775aae2p-4, -0x1.32d2cce62bd86p-8,
    0x1.50783487ee782p-13, -0x1.e3074fde8871fp-19, 0x1.e8f434d018d63p-25, -0x1.6fadb9f155744p-31,
    0x1.aaec32af93359p-38, -0x1.8a404211f9547p-45};
static const double C_[10] = {
    0x1.0000000000000p+0, -0x1.3bd3cc9be45dep+0, 0x1.03c1f081b5ac4p-2, -0x1.55d3c7e3cbffap-6,
    0x1.e1f506891babbp-11, -0x1.a6d1f2a204a8cp-16, 0x1.f9d38a3763cc3p-22, -0x1.b6e24f44b128fp-28,
    0x1.20c62c2f2d7f5p-34, -0x1.2a0c591af8314p-41};
/* 1/(2k+1), k = 1..12, for log(m) = 2 atanh(s) = 2 s (1 + s^2/3 + s^4/5 + ...). */
static const double L_[12] = {
    0x1.5555555555555p-2, 0x1.999999999999ap-3, 0x1.2492492492492p-3, 0x1.c71c71c71c71cp-4,
    0x1.745d1745d1746p-4, 0x1.3b13b13b13b14p-4, 0x1.1111111111111p-4, 0x1.e1e1e1e1e1e1ep-5,
    0x1.af286bca1af28p-5, 0x1.8618618618618p-5, 0x1.642c8590b2164p-5, 0x1.47ae147ae147bp-5};
static const double LN2_HI = 0x1.62e4200000000p-1; /* ln 2 with the low 32 mantissa bits cleared */
static const double LN2_LO = 0x1.fdf473de6af28p-22; /* ln 2 - LN2_HI */
static const double SQRT_HALF = 0x1.6a09e667f3bcdp-1;

void or_sincos2pi(double p, double* s, double* c)
{
    double q = p * 4.0;          /* quarter turns, exact */
    double k = rint(q);
    double r = q - k;            /* exact, |r| <= 1/2 */
    double r2 = r * r;
    double ps = S_[9], pc = C_[9];
    for (int i = 8; i >= 0; --i) {
        ps = fma(ps, r2, S_[i]);
        pc = fma(pc, r2, C_[i]);
    }
    double sn = ps * r, cs = pc;
    switch ((int)((int64_t)k & 3)) {
    case 0: *s = sn; *c = cs; break;
    case 1: *s = cs; *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
    }
}

static inline double sin2pi(double p)
{
    double s, c;
    or_sincos2pi(p, &s, &c);
    return s;
}

double or_log(double x)
{
    uint64_t b;
    memcpy(&b, &x, 8);
    int e = (int)((b >> 52) & 0x7ff) - 1022;
    b = (b & 0x000fffffffffffffULL) | 0x3fe0000000000000ULL;
    double m;
    memcpy(&m, &b, 8);            /* x = m 2^e, m in [1/2, 1) */
    if (m < SQRT_HALF) { m = m * 2.0; e -= 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double P = L_[11];
    for (int i = 10; i >= 0; --i) P = fma(P, z, L_[i]);
    double t = s * (z * P);
    double lm = 2.0 * (s + t);
    double de = (double)e;
    return de * LN2_HI + (de * LN2_LO + lm);
}

static inline double frac(double p) { return p - floor(p); }
static inline double lerp(double a, double b, double t) { return fma(t, b - a, a); }
static inline double norm3(const double v[3]) { return sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]); }

static inline uint32_t cd_from(double x)
{
    if (!(x < 2147483647.0)) return 2147483647u;
    if (x < 1.0) return 1u;
    return (uint32_t)x;
}

/* ---------------------------------------------------------- field sampling */

static inline void loc(double g, int n, int* i, double* t, int nearest)
{
    double gc = g;
    if (!(gc >= 0.0)) gc = 0.0; /* also maps NaN to 0 */
    if (gc > (double)(n - 1)) gc = (double)(n - 1);
    int ii = (int)gc;
    if (ii > n - 2) ii = n - 2;
    *i = ii;
    *t = nearest ? 0.0 : gc - (double)ii;
}

void or_sample(const double* F, int nx, int ny, int nz, const double g[3], int nearest, double out[3])
{
    int i, j, k;
    double tx, ty, tz;
    loc(g[0], nx, &i, &tx, nearest);
    loc(g[1], ny, &j, &ty, nearest);
    loc(g[2], nz, &k, &tz, nearest);
    size_t sy = (size_t)nx, sz = (size_t)nx * (size_t)ny;
    size_t b = (size_t)i + sy * ((size_t)j + (size_t)ny * (size_t)k);
    for (int d = 0; d < 3; ++d) {
        double c00 = lerp(F[3 * b + d], F[3 * (b + 1) + d], tx);
        double c10 = lerp(F[3 * (b + sy) + d], F[3 * (b + 1 + sy) + d], tx);
        double c01 = lerp(F[3 * (b + sz) + d], F[3 * (b + 1 + sz) + d], tx);
        double c11 = lerp(F[3 * (b + sy + sz) + d], F[3 * (b + 1 + sy + sz) + d], tx);
        out[d] = lerp(lerp(c00, c10, ty), lerp(c01, c11, ty), tz);
    }
}

/* ------------------------------------------------------ job/batch params */

typedef struct {
    double kq, v_rf, w_rf, v_ac, w_ac, s_dc;
    double nsig_dt, s_gas, v_bound, mr, one_m_mr, floor_v, mfp_mean, m;
    uint32_t key[2];
} or_jp;

typedef struct {
    const mq_fields* f;
    const double* dcc; /* combined DC, 3N */
    double dt, inv_h, lim[3];
    double z_ret, z_exit, cz0, cz1;
    uint64_t max_steps;
    int nearest, legacy_coll, compat;
} or_ctx;

/* channel.rs:376-381 */
static double legacy_mfp_mean(const mq_species* sp, double pressure_pa)
{
    double p_torr = pressure_pa / 133.322 + 1e-10;
    double np = 2.68e5 * (p_torr / 760.0);
    double mm = sp->mass;
    double ccs_he = 2.81e-9 * mm * mm * mm - 3.55e-5 * mm * mm + 0.232 * mm + 41.91;
    double ccs = ccs_he * 10.0 * sqrt(sp->z);
    return 1.0 / (np * ccs);
}

/* channel.rs:412-426 */
static void job_params(const mq_job* j, const mq_physics* ph, or_jp* J)
{
    double m = j->sp.mass * AMU, mg = 28.0 * AMU;
    double qm = j->sp.z * QE / m;
    double p = j->pressure_pa > 0.0 ? j->pressure_pa : ph->pressure_pa;
    double n_gas = p / (KB * ph->temp_k);
    double s_gas = sqrt(KB * ph->temp_k / mg);
    J->m = m;
    J->kq = (ph->legacy_force ? 0.5 : 1.0) * qm * ph->dt;
    J->v_rf = j->drive.v_rf;
    J->w_rf = j->drive.f_rf * ph->dt;
    J->v_ac = j->drive.v_ac;
    J->w_ac = j->drive.f_ac * ph->dt;
    J->s_dc = j->dc_scale;
    J->nsig_dt = n_gas * j->sp.ccs_n2 * 1e-20 * ph->dt;
    J->s_gas = s_gas;
    J->v_bound = 6.0 * s_gas;
    J->mr = mg / (m + mg);
    J->one_m_mr = 1.0 - J->mr;
    J->floor_v = ph->legacy_collisions ? sqrt(2.0 * 0.035 * 1.6e-19 / (j->sp.mass * 1.66e-27)) : -1.0;
    J->mfp_mean = legacy_mfp_mean(&j->sp, p);
    J->key[0] = (uint32_t)j->seed;
    J->key[1] = (uint32_t)(j->seed >> 32);
}

static uint64_t max_steps_of(const mq_physics* ph)
{
    if (ph->max_steps) return ph->max_steps;
    return (uint64_t)llround(ph->t_max / ph->dt);
}

static void ctx_init(or_ctx* C, const mq_fields* f, const double* dcc, const mq_physics* ph)
{
    C->f = f;
    C->dcc = dcc;
    C->dt = ph->dt;
    C->inv_h = 1.0 / f->h;
    C->lim[0] = (double)(f->nx - 1) * f->h;
    C->lim[1] = (double)(f->ny - 1) * f->h;
    C->lim[2] = (double)(f->nz - 1) * f->h;
    C->z_ret = ph->legacy_criteria ? f->z_return_legacy : f->z_return;
    C->z_exit = ph->legacy_criteria ? f->z_exit_legacy : f->z_exit;
    C->cz0 = ph->legacy_criteria ? f->crash_z0_legacy : -1e30;
    C->cz1 = ph->legacy_criteria ? f->crash_z1_legacy : 1e30;
    C->max_steps = max_steps_of(ph);
    C->nearest = ph->legacy_fields != 0;
    C->legacy_coll = ph->legacy_collisions != 0;
    C->compat = ph->rng == MQ_RNG_COMPAT;
}

/* --------------------------------------------------------------- inject */

void or_inject(const mq_fields* f, const mq_job* job, const mq_physics* ph, uint32_t ion, or_ion* o)
{
    or_jp J;
    job_params(job, ph, &J);
    uint32_t w[12];
    blk(J.key, ion, 0, 0, D_INJ, w);
    blk(J.key, ion, 0, 1, D_INJ, w + 4);
    blk(J.key, ion, 0, 2, D_INJ, w + 8);
    double u[12];
    for (int i = 0; i < 12; ++i) u[i] = u01(w[i]);
    double m = J.m, ax = f->axis_x, ay = f->axis_y, z0 = f->z_inject;
    double ufirst, sn, cs;
    if (ph->injection == MQ_INJ_GAUSSIAN) {
        double sig = ph->inj_sigma_mm * 1e-3, s = sqrt(KB * ph->temp_k / m);
        double r1 = sqrt(-2.0 * or_log(u[0]));
        or_sincos2pi(u[1], &sn, &cs);
        o->x[0] = ax + sig * (r1 * cs);
        o->x[1] = ay + sig * (r1 * sn);
        o->x[2] = z0;
        double r2 = sqrt(-2.0 * or_log(u[2]));
        or_sincos2pi(u[3], &sn, &cs);
        o->v[0] = s * (r2 * cs);
        o->v[1] = s * (r2 * sn);
        double r3 = sqrt(-2.0 * or_log(u[4]));
        or_sincos2pi(u[5], &sn, &cs);
        o->v[2] = s * (r3 * cs);
        o->phi_ac = u[6];
        ufirst = u[7];
    } else if (ph->legacy_injection) {
        /* channel.rs:350-355 */
        double kin = 0.1 * (0.75 + 0.75 * u[0]) * job->sp.z;
        double vin = sqrt(2.0 * kin * QE / m);
        or_sincos2pi((-10.0 + 20.0 * u[1]) / 360.0, &sn, &cs);
        double vx = 0.707 * vin * sn;
        o->x[0] = ax; o->x[1] = ay; o->x[2] = z0;
        o->v[0] = vx; o->v[1] = vx; o->v[2] = vin * cs;
        o->phi_ac = u[2];
        ufirst = u[3];
    } else {
        /* channel.rs:357-361, draw order r, theta, vz, vx(2), vy(2) */
        double s = sqrt(KB * ph->temp_k / m);
        double r = ph->r_inject_mm * 1e-3 * sqrt(u[0]);
        or_sincos2pi(u[1], &sn, &cs);
        double vz = s * sqrt(-2.0 * or_log(u[2]));
        o->x[0] = ax + r * cs;
        o->x[1] = ay + r * sn;
        o->x[2] = z0;
        double c3, s3;
        or_sincos2pi(u[4], &s3, &c3);
        o->v[0] = s * (sqrt(-2.0 * or_log(u[3])) * c3);
        or_sincos2pi(u[6], &s3, &c3);
        o->v[1] = s * (sqrt(-2.0 * or_log(u[5])) * c3);
        o->v[2] = vz;
        o->phi_ac = u[7];
        ufirst = u[8];
    }
    double vmag = norm3(o->v);
    double first_mfp = -or_log(fmax(ufirst, 1e-7)) * J.mfp_mean / 10.0;
    o->cd = cd_from(floor(first_mfp / vmag / ph->dt) + 1.0);
    o->steps = 0;
    o->ncoll = 0;
    o->status = MQ_ALIVE;
}

/* ------------------------------------------------ reference streams (COMPAT) */

uint64_t or_xorshift64(uint64_t* s)
{
    uint64_t x = *s;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    *s = x;
    return x;
}

uint32_t or_xoshiro128ss(uint32_t s[4])
{
    const uint32_t r = ((s[1] * 5u) << 7 | (s[1] * 5u) >> 25) * 9u;
    const uint32_t t = s[1] << 9;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = (s[3] << 11) | (s[3] >> 21);
    return r;
}

static double xs_u(uint64_t* s) { return (double)(or_xorshift64(s) >> 11) / 9007199254740992.0; }
static double xo_u(uint32_t s[4]) { return (double)(or_xoshiro128ss(s) >> 8) / 16777216.0; }

/* channel.rs:342-374: draws per ion in order (model draws, AC phase, first
 * free path, four seed words), stored as the reference's ChIon. */
void or_inject_reference(const mq_fields* f, const mq_job* job, const mq_physics* ph, or_ion* out)
{
    const double PI_ = 3.14159265358979323846;
    uint64_t st = job->seed * 0x9E3779B97F4A7C15ull ^ 0xD1B54A32D192ED03ull;
    const double m = job->sp.mass * AMU;
    const double mfp_mean = legacy_mfp_mean(&job->sp, job->pressure_pa > 0.0 ? job->pressure_pa : ph->pressure_pa);
    for (uint64_t i = 0; i < job->n; ++i) {
        double pos[3], vel[3];
        if (ph->injection == MQ_INJ_GAUSSIAN) {
            double sig = ph->inj_sigma_mm * 1e-3, s = sqrt(KB * ph->temp_k / m);
            double r = sqrt(-2.0 * log(fmax(xs_u(&st), 1e-300))), th = 2.0 * PI_ * xs_u(&st);
            pos[0] = f->axis_x + sig * (r * cos(th));
            pos[1] = f->axis_y + sig * (r * sin(th));
            pos[2] = f->z_inject;
            r = sqrt(-2.0 * log(fmax(xs_u(&st), 1e-300)));
            th = 2.0 * PI_ * xs_u(&st);
            vel[0] = s * (r * cos(th));
            vel[1] = s * (r * sin(th));
            r = sqrt(-2.0 * log(fmax(xs_u(&st), 1e-300)));
            th = 2.0 * PI_ * xs_u(&st);
            vel[2] = s * (r * cos(th));
        } else if (ph->legacy_injection) {
            double kin = 0.1 * (0.75 + 0.75 * xs_u(&st)) * job->sp.z;
            double vin = sqrt(2.0 * kin * QE / m);
            double ang = (-10.0 + 20.0 * xs_u(&st)) * (PI_ / 180.0);
            double vxy = 0.707 * vin * sin(ang);
            pos[0] = f->axis_x; pos[1] = f->axis_y; pos[2] = f->z_inject;
            vel[0] = vxy; vel[1] = vxy; vel[2] = vin * cos(ang);
        } else {
            double s = sqrt(KB * ph->temp_k / m);
            double r = ph->r_inject_mm * 1e-3 * sqrt(xs_u(&st));
            double th = 2.0 * PI_ * xs_u(&st);
            double vz = s * sqrt(-2.0 * log(fmax(xs_u(&st), 1e-300)));
            pos[0] = f->axis_x + r * cos(th);
            pos[1] = f->axis_y + r * sin(th);
            pos[2] = f->z_inject;
            double u1 = xs_u(&st), u2 = xs_u(&st);
            vel[0] = s * (sqrt(-2.0 * log(fmax(u1, 1e-300))) * cos(2.0 * PI_ * u2));
            double u3 = xs_u(&st), u4 = xs_u(&st);
            vel[1] = s * (sqrt(-2.0 * log(fmax(u3, 1e-300))) * cos(2.0 * PI_ * u4));
            vel[2] = vz;
        }
        double phase = 2.0 * PI_ * xs_u(&st);
        double vmag = sqrt(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
        double first_mfp = -log(fmax(xs_u(&st), 1e-7)) * mfp_mean / 10.0;
        double steps0 = floor(first_mfp / vmag / ph->dt) + 1.0;
        or_ion* o = &out[i];
        for (int d = 0; d < 3; ++d) {
            double q = pos[d] * MQ_POS_SCALE;
            int32_t p = !(q == q) ? 0 : q >= 2147483647.0 ? 2147483647 : q <= -2147483648.0 ? INT32_MIN : (int32_t)q;
            o->x[d] = (double)p / MQ_POS_SCALE;
            o->v[d] = (double)(float)vel[d];
        }
        o->phi_ac = (double)(float)(phase / (2.0 * PI_));
        float s0 = (float)steps0;
        o->cd = !(s0 >= 1.0f) ? 1u : (s0 >= 4294967040.0f ? 4294967040u : (uint32_t)s0);
        o->rng[0] = (uint32_t)or_xorshift64(&st) | 1u;
        o->rng[1] = (uint32_t)or_xorshift64(&st);
        o->rng[2] = (uint32_t)or_xorshift64(&st);
        o->rng[3] = (uint32_t)or_xorshift64(&st);
        o->steps = 0;
        o->ncoll = 0;
        o->status = MQ_ALIVE;
    }
}

/* ---------------------------------------------------------------- advance */

static void advance(const or_ctx* C, const or_jp* J, or_ion* s, uint32_t ion)
{
    const mq_fields* f = C->f;
    const double dt = C->dt;
    double x[3] = {s->x[0], s->x[1], s->x[2]}, v[3] = {s->v[0], s->v[1], s->v[2]};
    uint32_t n = s->steps, ncoll = s->ncoll, cd = s->cd;
    int status = 0;
    while (status == 0) {
        double g[3] = {x[0] * C->inv_h, x[1] * C->inv_h, x[2] * C->inv_h};
        double er[3], ed[3], ea[3];
        or_sample(f->rf, f->nx, f->ny, f->nz, g, C->nearest, er);
        or_sample(C->dcc, f->nx, f->ny, f->nz, g, C->nearest, ed);
        or_sample(f->ac, f->nx, f->ny, f->nz, g, C->nearest, ea);
        double a_rf = J->v_rf * sin2pi(frac(J->w_rf * (double)n));
        double a_ac = J->v_ac * sin2pi(frac(fma(J->w_ac, (double)n, s->phi_ac)));
        for (int d = 0; d < 3; ++d) {
            double E = (a_rf * er[d] + J->s_dc * ed[d]) + a_ac * ea[d];
            v[d] = fma(J->kq, E, v[d]);
            x[d] = fma(v[d], dt, x[d]);
        }
        const uint32_t st = n;
        n += 1;
        /* collisions (channel.metal:93-113) */
        if (C->compat) {
            /* the reference's xoshiro128** draws: rand01, gauss3, rand01, iso_dir */
            if (!C->legacy_coll) {
                double bound = norm3(v) + J->v_bound;
                if (xo_u(s->rng) < J->nsig_dt * bound) {
                    double u1 = fmax(xo_u(s->rng), 1e-7), u2 = xo_u(s->rng);
                    double u3 = fmax(xo_u(s->rng), 1e-7), u4 = xo_u(s->rng);
                    double r1 = sqrt(-2.0 * or_log(u1)), r2 = sqrt(-2.0 * or_log(u3));
                    double s1, c1, s2, c2;
                    or_sincos2pi(u2, &s1, &c1);
                    or_sincos2pi(u4, &s2, &c2);
                    double gv[3] = {v[0] - J->s_gas * (r1 * c1), v[1] - J->s_gas * (r1 * s1), v[2] - J->s_gas * (r2 * c2)};
                    double gl = norm3(gv);
                    if (xo_u(s->rng) * bound < gl) {
                        double c = 2.0 * xo_u(s->rng) - 1.0, sp, cp;
                        or_sincos2pi(xo_u(s->rng), &sp, &cp);
                        double sn = sqrt(fmax(0.0, 1.0 - c * c));
                        double nn[3] = {sn * cp, sn * sp, c};
                        double mg = J->mr * gl;
                        for (int d = 0; d < 3; ++d) v[d] = (v[d] - J->mr * gv[d]) + mg * nn[d];
                        ncoll += 1;
                    }
                }
            } else {
                cd -= 1;
                if (cd == 0) {
                    double vm = norm3(v);
                    if (vm > J->floor_v) {
                        double c = 2.0 * xo_u(s->rng) - 1.0, sp, cp;
                        or_sincos2pi(xo_u(s->rng), &sp, &cp);
                        double sn = sqrt(fmax(0.0, 1.0 - c * c));
                        double nn[3] = {sn * cp, sn * sp, c};
                        double mv = J->mr * vm;
                        for (int d = 0; d < 3; ++d) v[d] = v[d] * J->one_m_mr + mv * nn[d];
                        ncoll += 1;
                    }
                    double mfp = -or_log(fmax(xo_u(s->rng), 1e-7)) * J->mfp_mean;
                    cd = cd_from(floor(mfp / fmax(norm3(v), 1e-3) / dt));
                }
            }
        } else if (!C->legacy_coll) {
            if (J->nsig_dt > 0.0) {
                double bound = norm3(v) + J->v_bound;
                uint32_t w[4];
                blk(J->key, ion, st >> 2, 0, D_CAND, w);
                if (u01(w[st & 3]) < J->nsig_dt * bound) {
                    blk(J->key, ion, st, 0, D_COLL, w);
                    double r1 = sqrt(-2.0 * or_log(u01(w[0]))), r2 = sqrt(-2.0 * or_log(u01(w[2])));
                    double s1, c1, s2, c2;
                    or_sincos2pi(u01(w[1]), &s1, &c1);
                    or_sincos2pi(u01(w[3]), &s2, &c2);
                    double gv[3] = {v[0] - J->s_gas * (r1 * c1), v[1] - J->s_gas * (r1 * s1), v[2] - J->s_gas * (r2 * c2)};
                    double gl = norm3(gv);
                    blk(J->key, ion, st, 1, D_COLL, w);
                    if (u01(w[0]) * bound < gl) {
                        double c = 2.0 * u01(w[1]) - 1.0, sp, cp;
                        or_sincos2pi(u01(w[2]), &sp, &cp);
                        double sn = sqrt(fmax(0.0, 1.0 - c * c));
                        double nn[3] = {sn * cp, sn * sp, c};
                        double mg = J->mr * gl;
                        for (int d = 0; d < 3; ++d) v[d] = (v[d] - J->mr * gv[d]) + mg * nn[d];
                        ncoll += 1;
                    }
                }
            }
        } else {
            cd -= 1;
            if (cd == 0) {
                uint32_t w[4];
                blk(J->key, ion, st, 0, D_COLL, w);
                double vm = norm3(v);
                if (vm > J->floor_v) {
                    double c = 2.0 * u01(w[0]) - 1.0, sp, cp;
                    or_sincos2pi(u01(w[1]), &sp, &cp);
                    double sn = sqrt(fmax(0.0, 1.0 - c * c));
                    double nn[3] = {sn * cp, sn * sp, c};
                    double mv = J->mr * vm;
                    for (int d = 0; d < 3; ++d) v[d] = v[d] * J->one_m_mr + mv * nn[d];
                    ncoll += 1;
                }
                double mfp = -or_log(fmax(u01(w[2]), 1e-7)) * J->mfp_mean;
                cd = cd_from(floor(mfp / fmax(norm3(v), 1e-3) / dt));
            }
        }
        /* classification (channel.metal:114-120) */
        if (x[2] < C->z_ret) status = MQ_RETURNED;
        else if (x[2] > C->z_exit) status = MQ_TRANSMITTED;
        else if (x[0] < 0.0 || x[1] < 0.0 || x[2] < 0.0 || x[0] > C->lim[0] || x[1] > C->lim[1] || x[2] > C->lim[2])
            status = MQ_LEFT;
        else if (x[2] > C->cz0 && x[2] < C->cz1 && f->mask &&
                 f->mask[(size_t)rint(x[0] * C->inv_h) +
                         (size_t)f->nx * ((size_t)rint(x[1] * C->inv_h) + (size_t)f->ny * (size_t)rint(x[2] * C->inv_h))])
            status = MQ_CRASHED;
        else if ((uint64_t)n >= C->max_steps) status = MQ_TIMEOUT;
    }
    s->x[0] = x[0]; s->x[1] = x[1]; s->x[2] = x[2];
    s->v[0] = v[0]; s->v[1] = v[1]; s->v[2] = v[2];
    s->steps = n;
    s->ncoll = ncoll;
    s->cd = cd;
    s->status = (uint8_t)status;
}

/* ------------------------------------------------------------- batch run */

static double* combine_dc(const mq_fields* f, const double dc[4])
{
    size_t n3 = 3 * (size_t)f->nx * (size_t)f->ny * (size_t)f->nz;
    double* e = (double*)calloc(n3, sizeof(double));
    if (!e) return NULL;
    for (int b = 0; b < 4; ++b) {
        if (!f->dc[b] || dc[b] == 0.0) continue;
        for (size_t i = 0; i < n3; ++i) e[i] = e[i] + dc[b] * f->dc[b][i];
    }
    return e;
}

typedef struct {
    const or_ctx* C;
    const mq_job* jobs;
    const or_jp* jp;
    const uint64_t* off; /* job offsets, njobs + 1 */
    size_t njobs, total;
    const mq_physics* ph;
    const mq_fields* f;
    or_ion* out;
    uint32_t* out_job;
    uint32_t i0; /* first ion index inside the job (range runs) */
    atomic_size_t next;
} or_work;

static void* worker(void* arg)
{
    or_work* W = (or_work*)arg;
    for (;;) {
        size_t a = atomic_fetch_add(&W->next, 16);
        if (a >= W->total) break;
        size_t b = a + 16 < W->total ? a + 16 : W->total;
        size_t j = 0;
        while (j + 1 < W->njobs && W->off[j + 1] <= a) ++j;
        for (size_t i = a; i < b; ++i) {
            while (j + 1 < W->njobs && W->off[j + 1] <= i) ++j;
            uint32_t ion = (uint32_t)(i - W->off[j]) + W->i0;
            if (!W->C->compat) or_inject(W->f, &W->jobs[j], W->ph, ion, &W->out[i]);
            advance(W->C, &W->jp[j], &W->out[i], ion);
            if (W->out_job) W->out_job[i] = (uint32_t)j;
        }
    }
    return NULL;
}

static int run_common(const mq_fields* f, const mq_job* jobs, size_t njobs, const double dc[4], const mq_physics* ph,
                      int nthreads, or_ion* out, uint32_t* out_job, uint32_t i0, size_t total_override)
{
    double* dcc = combine_dc(f, dc);
    if (!dcc) return -1;
    or_ctx C;
    ctx_init(&C, f, dcc, ph);
    or_jp* jp = (or_jp*)malloc(njobs * sizeof(or_jp));
    uint64_t* off = (uint64_t*)malloc((njobs + 1) * sizeof(uint64_t));
    off[0] = 0;
    for (size_t j = 0; j < njobs; ++j) {
        job_params(&jobs[j], ph, &jp[j]);
        off[j + 1] = off[j] + jobs[j].n;
    }
    if (C.compat) {
        /* the reference's injection stream is sequential per job: inject on one thread first */
        if (i0 != 0 || total_override) {
            free(jp); free(off); free(dcc);
            return -2; /* a sub-range of a compat stream needs the whole job */
        }
        for (size_t j = 0; j < njobs; ++j) or_inject_reference(f, &jobs[j], ph, out + off[j]);
    }
    or_work W;
    W.C = &C; W.jobs = jobs; W.jp = jp; W.off = off; W.njobs = njobs;
    W.total = total_override ? total_override : (size_t)off[njobs];
    W.ph = ph; W.f = f; W.out = out; W.out_job = out_job; W.i0 = i0;
    atomic_init(&W.next, 0);
    if (nthreads < 1) nthreads = 1;
    pthread_t* th = (pthread_t*)malloc((size_t)nthreads * sizeof(pthread_t));
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, worker, &W);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
    free(jp);
    free(off);
    free(dcc);
    return 0;
}

int or_run_batch(const mq_fields* f, const mq_job* jobs, size_t njobs, const double dc[4], const mq_physics* ph,
                 int nthreads, or_ion* out, uint32_t* out_job)
{
    return run_common(f, jobs, njobs, dc, ph, nthreads, out, out_job, 0, 0);
}

int or_run_range(const mq_fields* f, const mq_job* job, const double dc[4], const mq_physics* ph, uint32_t i0,
                 uint32_t i1, int nthreads, or_ion* out)
{
    if (i1 <= i0) return 0;
    mq_job j = *job;
    j.n = i1 - i0;
    return run_common(f, &j, 1, dc, ph, nthreads, out, NULL, i0, 0);
}

void or_summarize(const or_ion* ions, const uint32_t* job, size_t n, size_t njobs, double dt, mq_outcome* out)
{
    memset(out, 0, njobs * sizeof(mq_outcome));
    for (size_t i = 0; i < n; ++i) {
        mq_outcome* o = &out[job[i]];
        o->n += 1;
        switch (ions[i].status) {
        case MQ_CRASHED: o->crashed += 1; break;
        case MQ_RETURNED: o->returned += 1; break;
        case MQ_TRANSMITTED: o->transmitted += 1; break;
        case MQ_LEFT: o->left += 1; break;
        default: o->timeout += 1; break;
        }
        o->mean_coll += (double)ions[i].ncoll;
        o->mean_time += (double)ions[i].steps * dt;
    }
    for (size_t j = 0; j < njobs; ++j)
        if (out[j].n) {
            out[j].mean_coll /= (double)out[j].n;
            out[j].mean_time /= (double)out[j].n;
        }
}

/* ------------------------------------------------ differentiable form (dual numbers) */

/* Value and Jacobian d out / d x of the trilinear sample at position x (clamped
 * coordinates have zero derivative). */
static void sample_jac_x(const double* F, const or_ctx* C, const double x[3], double out[3], double jac[3][3])
{
    const mq_fields* f = C->f;
    const int n[3] = {f->nx, f->ny, f->nz};
    int ijk[3];
    double t[3], dt[3];
    for (int a = 0; a < 3; ++a) {
        double g = x[a] * C->inv_h, gc = g;
        int cl = 0;
        if (!(gc >= 0.0)) { gc = 0.0; cl = 1; }
        if (gc > (double)(n[a] - 1)) { gc = (double)(n[a] - 1); cl = 1; }
        int ii = (int)gc;
        if (ii > n[a] - 2) ii = n[a] - 2;
        ijk[a] = ii;
        t[a] = C->nearest ? 0.0 : gc - (double)ii;
        dt[a] = (C->nearest || cl) ? 0.0 : C->inv_h;
    }
    size_t sy = (size_t)f->nx, sz = (size_t)f->nx * (size_t)f->ny;
    size_t b = (size_t)ijk[0] + sy * ((size_t)ijk[1] + (size_t)f->ny * (size_t)ijk[2]);
    for (int d = 0; d < 3; ++d) {
        double c000 = F[3 * b + d], c100 = F[3 * (b + 1) + d], c010 = F[3 * (b + sy) + d], c110 = F[3 * (b + 1 + sy) + d];
        double c001 = F[3 * (b + sz) + d], c101 = F[3 * (b + 1 + sz) + d];
        double c011 = F[3 * (b + sy + sz) + d], c111 = F[3 * (b + 1 + sy + sz) + d];
        double c00 = lerp(c000, c100, t[0]), c10 = lerp(c010, c110, t[0]);
        double c01 = lerp(c001, c101, t[0]), c11 = lerp(c011, c111, t[0]);
        double lo = lerp(c00, c10, t[1]), hi = lerp(c01, c11, t[1]);
        out[d] = lerp(lo, hi, t[2]);
        /* partial derivatives of the trilinear form in the cell coordinates */
        double w00 = (1.0 - t[1]) * (1.0 - t[2]), w10 = t[1] * (1.0 - t[2]), w01 = (1.0 - t[1]) * t[2], w11 = t[1] * t[2];
        double dtx = w00 * (c100 - c000) + w10 * (c110 - c010) + w01 * (c101 - c001) + w11 * (c111 - c011);
        double dty = (1.0 - t[2]) * (c10 - c00) + t[2] * (c11 - c01);
        double dtz = hi - lo;
        jac[d][0] = dtx * dt[0];
        jac[d][1] = dty * dt[1];
        jac[d][2] = dtz * dt[2];
    }
}

typedef struct {
    const or_ctx* C;
    const or_jp* J;
    const mq_job* job;
    const mq_physics* ph;
    const mq_objective* obj;
    uint32_t n;      /* number of output slots */
    const uint32_t* ions;  /* slot -> ion index, or NULL for slot == ion */
    double* Ji;      /* n */
    double* gi;      /* 5 * n */
    double* gsi;     /* 6 * n */
    atomic_size_t next;
} or_gwork;

static double sigm(double u) { return 1.0 / (1.0 + exp(-u)); }

static void grad_one(const or_gwork* W, uint32_t ion, uint32_t slot)
{
    const or_ctx* C = W->C;
    const or_jp* J = W->J;
    const mq_fields* f = C->f;
    enum { NT = 11 };
    or_ion s0;
    or_inject(f, W->job, W->ph, ion, &s0);
    double x[3] = {s0.x[0], s0.x[1], s0.x[2]}, v[3] = {s0.v[0], s0.v[1], s0.v[2]};
    double dx[NT][3], dv[NT][3];
    memset(dx, 0, sizeof dx);
    memset(dv, 0, sizeof dv);
    for (int a = 0; a < 3; ++a) { dx[5 + a][a] = 1.0; dv[8 + a][a] = 1.0; }
    const double dt = C->dt, two_pi = 6.283185307179586;
    for (uint32_t n = 0; n < W->obj->nsteps; ++n) {
        double er[3], ed[3], ea[3], jr[3][3], jd[3][3], ja[3][3];
        sample_jac_x(f->rf, C, x, er, jr);
        sample_jac_x(C->dcc, C, x, ed, jd);
        sample_jac_x(f->ac, C, x, ea, ja);
        double srf, crf, sac, cac;
        or_sincos2pi(frac(J->w_rf * (double)n), &srf, &crf);
        or_sincos2pi(frac(fma(J->w_ac, (double)n, s0.phi_ac)), &sac, &cac);
        double a_rf = J->v_rf * srf, a_ac = J->v_ac * sac;
        /* d a_rf / d f_rf and d a_ac / d f_ac */
        double da_rf = J->v_rf * crf * two_pi * dt * (double)n, da_ac = J->v_ac * cac * two_pi * dt * (double)n;
        double E[3], JE[3][3];
        for (int d = 0; d < 3; ++d) {
            E[d] = (a_rf * er[d] + J->s_dc * ed[d]) + a_ac * ea[d];
            for (int e = 0; e < 3; ++e) JE[d][e] = (a_rf * jr[d][e] + J->s_dc * jd[d][e]) + a_ac * ja[d][e];
        }
        for (int k = 0; k < NT; ++k) {
            double dE[3];
            for (int d = 0; d < 3; ++d) {   /* all components from the tangent before this step */
                dE[d] = JE[d][0] * dx[k][0] + JE[d][1] * dx[k][1] + JE[d][2] * dx[k][2];
                if (k == 0) dE[d] += srf * er[d];
                else if (k == 1) dE[d] += da_rf * er[d];
                else if (k == 2) dE[d] += sac * ea[d];
                else if (k == 3) dE[d] += da_ac * ea[d];
                else if (k == 4) dE[d] += ed[d];
            }
            for (int d = 0; d < 3; ++d) {
                dv[k][d] += J->kq * dE[d];
                dx[k][d] += dt * dv[k][d];
            }
        }
        for (int d = 0; d < 3; ++d) {
            v[d] = fma(J->kq, E[d], v[d]);
            x[d] = fma(v[d], dt, x[d]);
        }
        /* thermal collision of step n with frozen draws (the Philox branch of
         * advance above): the candidate and acceptance decisions are taken on the
         * primal velocity and held fixed, so the tangent goes through the map
         * v' = v - mr g + mr |g| d, g = v - gas velocity, whose Jacobian is
         * (1 - mr) I + mr d (g / |g|)^T (pathwise derivative). */
        if (!C->compat && !C->legacy_coll && J->nsig_dt > 0.0) {
            double bound = norm3(v) + J->v_bound;
            uint32_t w[4];
            blk(J->key, ion, n >> 2, 0, D_CAND, w);
            if (u01(w[n & 3]) < J->nsig_dt * bound) {
                blk(J->key, ion, n, 0, D_COLL, w);
                double r1 = sqrt(-2.0 * or_log(u01(w[0]))), r2 = sqrt(-2.0 * or_log(u01(w[2])));
                double s1, c1, s2, c2;
                or_sincos2pi(u01(w[1]), &s1, &c1);
                or_sincos2pi(u01(w[3]), &s2, &c2);
                double gv[3] = {v[0] - J->s_gas * (r1 * c1), v[1] - J->s_gas * (r1 * s1), v[2] - J->s_gas * (r2 * c2)};
                double gl = norm3(gv);
                blk(J->key, ion, n, 1, D_COLL, w);
                if (u01(w[0]) * bound < gl) {
                    double c = 2.0 * u01(w[1]) - 1.0, sp, cp;
                    or_sincos2pi(u01(w[2]), &sp, &cp);
                    double sn = sqrt(fmax(0.0, 1.0 - c * c));
                    double nn[3] = {sn * cp, sn * sp, c};
                    double mg = J->mr * gl;
                    for (int k = 0; k < NT; ++k) {
                        double gd = (gv[0] * dv[k][0] + gv[1] * dv[k][1] + gv[2] * dv[k][2]) / gl;
                        for (int d = 0; d < 3; ++d) dv[k][d] = (1.0 - J->mr) * dv[k][d] + J->mr * nn[d] * gd;
                    }
                    for (int d = 0; d < 3; ++d) v[d] = (v[d] - J->mr * gv[d]) + mg * nn[d];
                }
            }
        }
    }
    /* objective and its gradient at the final position */
    const mq_objective* O = W->obj;
    double u = (x[2] - O->z_exit) / O->scale, st = sigm(u);
    double Ji = st, gJ[3] = {0.0, 0.0, st * (1.0 - st) / O->scale};
    if (O->w_leak != 0.0) {
        double rx = x[0] - f->axis_x, ry = x[1] - f->axis_y, r = sqrt(rx * rx + ry * ry);
        double sl = sigm((r - O->r_leak) / O->scale);
        Ji -= O->w_leak * sl;
        if (r > 0.0) {
            double k = -O->w_leak * sl * (1.0 - sl) / O->scale / r;
            gJ[0] = k * rx;
            gJ[1] = k * ry;
        }
    }
    W->Ji[slot] = Ji;
    for (int k = 0; k < NT; ++k) {
        double g = gJ[0] * dx[k][0] + gJ[1] * dx[k][1] + gJ[2] * dx[k][2];
        if (k < 5) W->gi[(size_t)k * W->n + slot] = g;
        else if (W->gsi) W->gsi[(size_t)(k - 5) * W->n + slot] = g;
    }
}

static void* grad_worker(void* arg)
{
    or_gwork* W = (or_gwork*)arg;
    for (;;) {
        size_t a = atomic_fetch_add(&W->next, 1);
        if (a >= W->n) break;
        grad_one(W, W->ions ? W->ions[a] : (uint32_t)a, (uint32_t)a);
    }
    return NULL;
}

/* Collisions in the differentiable form: the thermal model on the Philox
 * streams only (frozen draws). */
static int grad_physics_ok(const mq_job* job, const mq_physics* ph)
{
    const int gas = ph->pressure_pa > 0.0 || job->pressure_pa > 0.0;
    return !gas || (!ph->legacy_collisions && ph->rng == MQ_RNG_PHILOX);
}

int or_grad_reference(const mq_fields* f, const mq_job* job, const double dc[4], const mq_physics* ph,
                      const mq_objective* obj, int nthreads, double* Jout, double grad[5], double* grad_state)
{
    if (!grad_physics_ok(job, ph)) return -3;
    double* dcc = combine_dc(f, dc);
    if (!dcc) return -1;
    or_ctx C;
    ctx_init(&C, f, dcc, ph);
    or_jp J;
    job_params(job, ph, &J);
    uint32_t n = (uint32_t)job->n;
    or_gwork W;
    W.C = &C; W.J = &J; W.job = job; W.ph = ph; W.obj = obj; W.n = n; W.ions = NULL;
    W.Ji = (double*)calloc(n, sizeof(double));
    W.gi = (double*)calloc((size_t)5 * n, sizeof(double));
    W.gsi = (double*)calloc((size_t)6 * n, sizeof(double));
    atomic_init(&W.next, 0);
    if (nthreads < 1) nthreads = 1;
    pthread_t* th = (pthread_t*)malloc((size_t)nthreads * sizeof(pthread_t));
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, grad_worker, &W);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
    double sJ = 0.0;
    for (uint32_t i = 0; i < n; ++i) sJ += W.Ji[i];
    *Jout = sJ / (double)n;
    for (int k = 0; k < 5; ++k) {
        double s = 0.0;
        for (uint32_t i = 0; i < n; ++i) s += W.gi[(size_t)k * n + i];
        grad[k] = s / (double)n;
    }
    if (grad_state)
        for (size_t i = 0; i < (size_t)6 * n; ++i) grad_state[i] = W.gsi[i] / (double)n;
    free(W.Ji);
    free(W.gi);
    free(W.gsi);
    free(dcc);
    return 0;
}

/* Per-ion form of or_grad_reference for a list of ion indices of the job:
 * Ji[s] = objective of ion ions[s], gi[k * nions + s] = its d/dtheta_k (not
 * divided by job->n). Used to examine single trajectories of a large batch. */
int or_grad_ions(const mq_fields* f, const mq_job* job, const double dc[4], const mq_physics* ph,
                 const mq_objective* obj, int nthreads, const uint32_t* ions, uint32_t nions, double* Ji, double* gi)
{
    if (!grad_physics_ok(job, ph)) return -3;
    for (uint32_t s = 0; s < nions; ++s)
        if (ions[s] >= (uint32_t)job->n) return -2;
    double* dcc = combine_dc(f, dc);
    if (!dcc) return -1;
    or_ctx C;
    ctx_init(&C, f, dcc, ph);
    or_jp J;
    job_params(job, ph, &J);
    or_gwork W;
    W.C = &C; W.J = &J; W.job = job; W.ph = ph; W.obj = obj; W.n = nions; W.ions = ions;
    W.Ji = Ji;
    W.gi = gi;
    W.gsi = NULL;
    atomic_init(&W.next, 0);
    if (nthreads < 1) nthreads = 1;
    pthread_t* th = (pthread_t*)malloc((size_t)nthreads * sizeof(pthread_t));
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, grad_worker, &W);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
    free(dcc);
    return 0;
}

/* ================================================================ tandem batch
 * Restates trapsim's Model::advance / enter / leave
 * (simulation/trapsim/src/lib.rs:176-385) with chsim's step and bath, in the
 * operation order of the GPU's FP64 tandem kernel: the cube's conductor
 * nodes, the mouths above the seam, the top, the entrance aperture and tube on
 * the y face, then the other faces (lib.rs:206-243). Counters per port:
 * entered, returned, forwarded, revisits, from_other, reverse entries, lost,
 * retry_from[9], by_job[8] (TC_* layout of the product). */

enum { TC_PORT = 24, TC_ENTERED = 0, TC_RETURNED = 1, TC_FORWARDED = 2, TC_REVISITS = 3, TC_FROM_OTHER = 4,
       TC_REVERSE = 5, TC_LOST = 6, TC_RETRY = 7, TC_BYJOB = 16 };
#define OR_NO_PORT 255u

typedef struct {
    const double *rf, *ac;
    const uint8_t* mask;
    double inv_h, lim[3];
    int nx, ny, nz;
} or_tgrid;

typedef struct {
    or_tgrid cube, port;
    const double* dc[4]; /* splitter, combiner, driven port, idle port */
    double dt, seam, ap_r2, ent_x, ent_z, l_ext, z_ret, z_exit, ax_p, ay_p, mx[9], my[9];
    double v_rf, w_rf, v_ac, w_ac[9];
    uint64_t max_steps;
    int nearest;
} or_tctx;

static void tgrid_init(or_tgrid* G, const mq_fields* f)
{
    G->rf = f->rf; G->ac = f->ac; G->mask = f->mask;
    G->inv_h = 1.0 / f->h;
    G->lim[0] = (double)(f->nx - 1) * f->h;
    G->lim[1] = (double)(f->ny - 1) * f->h;
    G->lim[2] = (double)(f->nz - 1) * f->h;
    G->nx = f->nx; G->ny = f->ny; G->nz = f->nz;
}

/* conductor bit of the nearest node, the index clamped to the grid (a cube ion
 * may be past its y face, inside the entrance tube) */
static size_t clamp_node(double v, double inv_h, int n)
{
    return (size_t)fmin(fmax(rint(v * inv_h), 0.0), (double)(n - 1));
}

static int tgrid_crash(const or_tgrid* G, const double x[3])
{
    if (!G->mask) return 0;
    return G->mask[clamp_node(x[0], G->inv_h, G->nx) +
                   (size_t)G->nx * (clamp_node(x[1], G->inv_h, G->ny) +
                                    (size_t)G->ny * clamp_node(x[2], G->inv_h, G->nz))] != 0;
}

static int outside(const or_tgrid* G, const double x[3])
{
    return x[0] < 0.0 || x[1] < 0.0 || x[2] < 0.0 || x[0] > G->lim[0] || x[1] > G->lim[1] || x[2] > G->lim[2];
}

/* classification and region crossing after a step (lib.rs:206-265, 326-377) */
static int tandem_cross(const or_tctx* T, uint32_t* reg, uint32_t* last, double x[3], double v[3], uint32_t job,
                        uint64_t* ctr)
{
    if (*reg == 0 || *reg == 10) {
        if (tgrid_crash(&T->cube, x)) return 1; /* a conductor node of the cube (lib.rs:208-212) */
        if (x[2] > T->seam) {
            uint32_t p = 0;
            double best = 0.0;
            for (uint32_t k = 0; k < 9; ++k) {
                double dx = x[0] - T->mx[k], dy = x[1] - T->my[k];
                double d2 = dx * dx + dy * dy;
                if (k == 0 || d2 < best) { best = d2; p = k; }
            }
            if (best < T->ap_r2) {
                uint64_t* c = ctr + (size_t)p * TC_PORT;
                x[0] = x[0] + (T->ax_p - T->mx[p]);
                x[1] = x[1] + (T->ay_p - T->my[p]);
                if (*reg == 0) {
                    x[2] = T->z_ret + (x[2] - T->seam);
                    c[TC_ENTERED] += 1;
                    if (*last != OR_NO_PORT) {
                        c[TC_REVISITS] += 1;
                        c[TC_RETRY + *last] += 1;
                        if (*last != p) c[TC_FROM_OTHER] += 1;
                    }
                } else {
                    x[2] = T->z_exit - (x[2] - T->seam);
                    v[2] = -v[2];
                    c[TC_REVERSE] += 1;
                }
                *reg = 1 + p;
                return 0;
            }
            /* outside every mouth: the ion keeps moving above the board, up to the top (lib.rs:226-229) */
            if (x[2] > T->cube.lim[2]) return 1;
        }
        if (x[1] > T->cube.lim[1]) {
            /* the entrance face: a wall outside the aperture (lib.rs:231-236); through it, the tube, and
               l_ext beyond the face the ion is back or detected (lib.rs:237-243) */
            const double ex = x[0] - T->ent_x, ez = x[2] - T->ent_z;
            if (!(ex * ex + ez * ez < T->ap_r2)) return 1;
            if (x[1] > T->cube.lim[1] + T->l_ext) return *reg == 0 ? 2 : 3;
        }
        if (x[0] < 0.0 || x[1] < 0.0 || x[2] < 0.0 || x[0] > T->cube.lim[0]) return 1;
        return 0;
    }
    const uint32_t p = *reg - 1;
    uint64_t* c = ctr + (size_t)p * TC_PORT;
    if (outside(&T->port, x) || tgrid_crash(&T->port, x)) {
        c[TC_LOST] += 1;
        return 1;
    }
    if (x[2] < T->z_ret) {
        x[0] = x[0] + (T->mx[p] - T->ax_p);
        x[1] = x[1] + (T->my[p] - T->ay_p);
        x[2] = T->seam + (x[2] - T->z_ret);
        *last = p;
        *reg = 0;
        c[TC_RETURNED] += 1;
    } else if (x[2] > T->z_exit) {
        x[0] = x[0] + (T->mx[p] - T->ax_p);
        x[1] = x[1] + (T->my[p] - T->ay_p);
        x[2] = T->seam - (x[2] - T->z_exit);
        v[2] = -v[2];
        *last = OR_NO_PORT;
        *reg = 10;
        c[TC_FORWARDED] += 1;
        if (job < 8) c[TC_BYJOB + job] += 1;
    }
    return 0;
}

static void tandem_advance(const or_tctx* T, const or_jp* J, or_ion* s, uint32_t ion, uint32_t job, uint8_t* reg_out,
                           uint8_t* last_out, uint64_t* ctr)
{
    const double dt = T->dt;
    double x[3] = {s->x[0], s->x[1], s->x[2]}, v[3] = {s->v[0], s->v[1], s->v[2]};
    uint32_t n = s->steps, ncoll = s->ncoll, reg = 0, last = OR_NO_PORT;
    int status = 0;
    while (status == 0) {
        const int cube = reg == 0 || reg == 10;
        const or_tgrid* G = cube ? &T->cube : &T->port;
        const double wa = cube ? 0.0 : T->w_ac[reg - 1];
        const double* dcp = reg == 0 ? T->dc[0] : reg == 10 ? T->dc[1] : (wa > 0.0 ? T->dc[2] : T->dc[3]);
        double g[3] = {x[0] * G->inv_h, x[1] * G->inv_h, x[2] * G->inv_h};
        double er[3], ed[3], ea[3] = {0.0, 0.0, 0.0};
        or_sample(G->rf, G->nx, G->ny, G->nz, g, T->nearest, er);
        or_sample(dcp, G->nx, G->ny, G->nz, g, T->nearest, ed);
        double a_rf = T->v_rf * sin2pi(frac(T->w_rf * (double)n));
        double a_ac = 0.0;
        if (!cube) {
            or_sample(G->ac, G->nx, G->ny, G->nz, g, T->nearest, ea);
            if (wa > 0.0) a_ac = T->v_ac * sin2pi(frac(wa * (double)n));
        }
        for (int d = 0; d < 3; ++d) {
            double E = (a_rf * er[d] + ed[d]) + a_ac * ea[d];
            v[d] = fma(J->kq, E, v[d]);
            x[d] = fma(v[d], dt, x[d]);
        }
        const uint32_t st = n;
        n += 1;
        if (J->nsig_dt > 0.0) {
            double bound = norm3(v) + J->v_bound;
            uint32_t w[4];
            blk(J->key, ion, st >> 2, 0, D_CAND, w);
            if (u01(w[st & 3]) < J->nsig_dt * bound) {
                blk(J->key, ion, st, 0, D_COLL, w);
                double r1 = sqrt(-2.0 * or_log(u01(w[0]))), r2 = sqrt(-2.0 * or_log(u01(w[2])));
                double s1, c1, s2, c2;
                or_sincos2pi(u01(w[1]), &s1, &c1);
                or_sincos2pi(u01(w[3]), &s2, &c2);
                double gv[3] = {v[0] - J->s_gas * (r1 * c1), v[1] - J->s_gas * (r1 * s1), v[2] - J->s_gas * (r2 * c2)};
                double gl = norm3(gv);
                blk(J->key, ion, st, 1, D_COLL, w);
                if (u01(w[0]) * bound < gl) {
                    double c = 2.0 * u01(w[1]) - 1.0, sp, cp;
                    or_sincos2pi(u01(w[2]), &sp, &cp);
                    double sn = sqrt(fmax(0.0, 1.0 - c * c));
                    double nn[3] = {sn * cp, sn * sp, c};
                    double mg = J->mr * gl;
                    for (int d = 0; d < 3; ++d) v[d] = (v[d] - J->mr * gv[d]) + mg * nn[d];
                    ncoll += 1;
                }
            }
        }
        status = tandem_cross(T, &reg, &last, x, v, job, ctr);
        if (status == 0 && (uint64_t)n >= T->max_steps) status = 5;
    }
    s->x[0] = x[0]; s->x[1] = x[1]; s->x[2] = x[2];
    s->v[0] = v[0]; s->v[1] = v[1]; s->v[2] = v[2];
    s->steps = n;
    s->ncoll = ncoll;
    s->status = (uint8_t)status;
    *reg_out = (uint8_t)reg;
    *last_out = (uint8_t)last;
}

typedef struct {
    const or_tctx* T;
    const mq_fields* cube;
    const mq_job* jobs;
    const or_jp* jp;
    const uint64_t* off;
    size_t njobs, total;
    const mq_physics* ph;
    or_ion* out;
    uint32_t* out_job;
    uint8_t *region, *last;
    uint64_t* ctr; /* per thread: 9 * TC_PORT */
    atomic_size_t next;
} or_twork;

typedef struct {
    or_twork* W;
    uint64_t* ctr;
} or_targ;

static void* tandem_worker(void* arg)
{
    or_targ* A = (or_targ*)arg;
    or_twork* W = A->W;
    for (;;) {
        size_t a = atomic_fetch_add(&W->next, 16);
        if (a >= W->total) break;
        size_t b = a + 16 < W->total ? a + 16 : W->total;
        for (size_t i = a; i < b; ++i) {
            size_t j = 0;
            while (j + 1 < W->njobs && W->off[j + 1] <= i) ++j;
            uint32_t ion = (uint32_t)(i - W->off[j]);
            or_inject(W->cube, &W->jobs[j], W->ph, ion, &W->out[i]);
            tandem_advance(W->T, &W->jp[j], &W->out[i], ion, (uint32_t)j, &W->region[i], &W->last[i], A->ctr);
            if (W->out_job) W->out_job[i] = (uint32_t)j;
        }
    }
    return NULL;
}

int or_tandem_run(const mq_fields* cube, const mq_fields* port, const mq_tandem_geom* g, const mq_job* jobs,
                  size_t njobs, const mq_tandem_drive* d, const mq_physics* ph, int nthreads, or_ion* out,
                  uint32_t* out_job, uint8_t* region, uint8_t* last, uint64_t* counts, uint64_t* totals)
{
    or_tctx T;
    memset(&T, 0, sizeof T);
    tgrid_init(&T.cube, cube);
    tgrid_init(&T.port, port);
    double idle[4];
    for (int b = 0; b < 4; ++b) idle[b] = d->dc_port[b] + d->bias_idle[b];
    double* dcs[4] = {combine_dc(cube, d->dc_split), combine_dc(cube, d->dc_combine), combine_dc(port, d->dc_port),
                      combine_dc(port, idle)};
    for (int k = 0; k < 4; ++k)
        if (!dcs[k]) return -1;
    for (int k = 0; k < 4; ++k) T.dc[k] = dcs[k];
    T.dt = ph->dt;
    T.seam = g->seam_z;
    T.ap_r2 = g->ap_r * g->ap_r;
    T.ent_x = g->entrance[0];
    T.ent_z = g->entrance[1];
    T.l_ext = g->l_ext;
    T.z_ret = port->z_return;
    T.z_exit = port->z_exit;
    T.ax_p = port->axis_x;
    T.ay_p = port->axis_y;
    for (int k = 0; k < 9; ++k) {
        T.mx[k] = g->mouth[k][0];
        T.my[k] = g->mouth[k][1];
        T.w_ac[k] = d->tones[k] * ph->dt;
    }
    T.v_rf = d->v_rf;
    T.w_rf = d->f_rf * ph->dt;
    T.v_ac = d->v_ac;
    T.max_steps = max_steps_of(ph);
    T.nearest = ph->legacy_fields != 0;
    or_jp* jp = (or_jp*)malloc(njobs * sizeof(or_jp));
    uint64_t* off = (uint64_t*)malloc((njobs + 1) * sizeof(uint64_t));
    off[0] = 0;
    for (size_t j = 0; j < njobs; ++j) {
        job_params(&jobs[j], ph, &jp[j]);
        off[j + 1] = off[j] + jobs[j].n;
    }
    if (nthreads < 1) nthreads = 1;
    or_twork W;
    W.T = &T; W.cube = cube; W.jobs = jobs; W.jp = jp; W.off = off; W.njobs = njobs; W.total = (size_t)off[njobs];
    W.ph = ph; W.out = out; W.out_job = out_job; W.region = region; W.last = last;
    atomic_init(&W.next, 0);
    uint64_t* ctrs = (uint64_t*)calloc((size_t)nthreads * 9 * TC_PORT, sizeof(uint64_t));
    or_targ* args = (or_targ*)malloc((size_t)nthreads * sizeof(or_targ));
    pthread_t* th = (pthread_t*)malloc((size_t)nthreads * sizeof(pthread_t));
    for (int t = 0; t < nthreads; ++t) {
        args[t].W = &W;
        args[t].ctr = ctrs + (size_t)t * 9 * TC_PORT;
        pthread_create(&th[t], NULL, tandem_worker, &args[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    for (int k = 0; k < 9 * TC_PORT; ++k) {
        uint64_t s = 0;
        for (int t = 0; t < nthreads; ++t) s += ctrs[(size_t)t * 9 * TC_PORT + k];
        counts[k] = s;
    }
    for (int k = 0; k < 8; ++k) totals[k] = 0;
    for (size_t i = 0; i < W.total; ++i) {
        totals[0] += 1;
        int st = out[i].status, r = region[i];
        int k = st == 1 ? 4 : st == 2 ? 5 : st == 3 ? 6 : (r == 0 ? 1 : r == 10 ? 3 : 2);
        totals[k] += 1;
        totals[7] += out[i].steps;
    }
    free(th);
    free(args);
    free(ctrs);
    free(jp);
    free(off);
    for (int k = 0; k < 4; ++k) free(dcs[k]);
    return 0;
}
