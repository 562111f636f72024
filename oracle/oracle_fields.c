/*
 * oracle_fields.c -- CPU restatement of the basis-field pipeline.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker of the GPU field
 * build, never part of the product.
 *
 * What it restates (reference file:line):
 *   Geom::default, layers, nz, center, axis  simulation/chsim/src/channel.rs:38-57
 *   Geom::classify (electrode rasterization) channel.rs:59-103
 *   basis_value, rf_sign, ac_sign            channel.rs:101-119
 *   the discrete problem of fieldsolve::solve (7-point Laplacian, masked
 *   Dirichlet nodes, box faces at outer_bc, true max-norm residual scaled
 *   by max(|fixed|, |outer_bc|) or 1)        fieldsolve/src/lib.rs:150-166,
 *                                            mg.rs:962-965
 *   efield (central differences, one-sided at the faces) channel.rs:168-184
 *
 * The solver here is red-black successive over-relaxation, a different
 * method from the reference's multigrid: both stop on the same residual of
 * the same linear system, whose solution is unique, so the potentials agree
 * to the tolerance. It is meant for small grids only.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void or_geom_default(mq_geom* g)
{
    memset(g, 0, sizeof *g);
    g->de = 6.35; g->coef = 0.323; g->q1_start = 3.0; g->q1_len = 12.0; g->alpha = 0.02; g->gap = 1.0;
    g->plate = 1.6; g->hole_d = 7.0; g->q2_len = 3.0; g->stem_d = 4.6; g->q2_flare_bug = 1; g->grow = 0.0;
    g->tilt_mode = 0; g->ac_mode = 0; g->n_xy = 160; g->h_mm = 0.1;
}

void or_geom_layers(const mq_geom* g, double l[8])
{
    l[0] = g->q1_start;
    l[1] = l[0] + g->q1_len;
    l[2] = l[1] + g->gap;
    l[3] = l[2] + g->plate;
    l[4] = l[3] + g->gap;
    l[5] = l[4] + g->q2_len;
    l[6] = l[5] + g->gap;
    l[7] = l[6] + g->plate;
}

int or_geom_nz(const mq_geom* g)
{
    double l[8];
    or_geom_layers(g, l);
    return (int)(l[7] / g->h_mm + 1e-6);
}

/* Quadrant of (x, y) about the axis a: 0 LL, 1 RL, 2 LU, 3 RU, -1 on an axis plane. */
static int quadrant(double x, double y, double a)
{
    if (x < a && y < a) return 0;
    if (x > a && y < a) return 1;
    if (x < a && y > a) return 2;
    if (x > a && y > a) return 3;
    return -1;
}

/* Is (x, y) inside rod q (diameter d, squeeze sq)? Q1 rods add the
 * misalignment of mis[q] about the Q1 midpoint. */
static int in_rod(const mq_geom* g, int q, double sq, double d, int q1, double x, double y, double z, double a,
                  double c)
{
    double cx, cy;
    const double s2 = sq / sqrt(2.0);
    if (g->tilt_mode == 0) {
        switch (q) {
            case 0: cx = a - c; cy = a - c + sq; break;
            case 1: cx = a + c; cy = a - c + sq; break;
            case 2: cx = a - c; cy = a + c - sq; break;
            default: cx = a + c; cy = a + c - sq; break;
        }
    } else if (g->tilt_mode == 2) {
        switch (q) {
            case 0: cx = a - c + s2; cy = a - c + s2; break;
            case 1: cx = a + c - s2; cy = a - c + s2; break;
            case 2: cx = a - c + s2; cy = a + c - s2; break;
            default: cx = a + c - s2; cy = a + c - s2; break;
        }
    } else {
        switch (q) {
            case 0: cx = a - c; cy = a - c; break;
            case 1: cx = a + c; cy = a - c; break;
            case 2: cx = a - c + sq; cy = a + c; break;
            default: cx = a + c - sq; cy = a + c; break;
        }
    }
    if (q1) {
        const double zmid = g->q1_start + g->q1_len / 2.0;
        const double* m = g->mis[q];
        cx += m[0] + m[2] * (z - zmid);
        cy += m[1] + m[3] * (z - zmid);
    }
    const double dx = x - cx, dy = y - cy;
    return sqrt(dx * dx + dy * dy) < d / 2.0 + g->grow * g->h_mm;
}

int or_geom_classify_node(const mq_geom* g, int i, int j, int k)
{
    const double x = (double)i * g->h_mm, y = (double)j * g->h_mm, z = (double)k * g->h_mm;
    const double a = (double)(g->n_xy / 2) * g->h_mm;
    const double c = g->coef * g->de / 2.0 + g->de / 2.0;
    double l[8];
    or_geom_layers(g, l);
    const double eps = 1e-9;
    const int q = quadrant(x, y, a);
    if (z <= l[0] + eps) return MQ_ELEC_NONE;
    if (z <= l[1] + eps) {
        const double sq = g->alpha * (l[1] - z);
        return (q >= 0 && in_rod(g, q, sq, g->de, 1, x, y, z, a, c)) ? MQ_ELEC_Q1 + q : MQ_ELEC_NONE;
    }
    if (z <= l[2] + eps) return (q >= 0 && in_rod(g, q, 0.0, g->stem_d, 0, x, y, z, a, c)) ? MQ_ELEC_STEM3 + q : 0;
    if (z <= l[3] + eps) {
        const double r2 = (x - a) * (x - a) + (y - a) * (y - a);
        const double hr = g->hole_d / 2.0 - g->grow * g->h_mm;
        return r2 < hr * hr ? MQ_ELEC_NONE : MQ_ELEC_PLATE;
    }
    if (z <= l[4] + eps) return (q >= 0 && in_rod(g, q, 0.0, g->stem_d, 0, x, y, z, a, c)) ? MQ_ELEC_STEM5 + q : 0;
    if (z <= l[5] + eps) {
        const double sq = g->q2_flare_bug ? g->alpha * (l[1] - z) : 0.0;
        return (q >= 0 && in_rod(g, q, sq, g->de, 0, x, y, z, a, c)) ? MQ_ELEC_Q2 + q : MQ_ELEC_NONE;
    }
    if (z <= l[6] + eps) return (q >= 0 && in_rod(g, q, 0.0, g->stem_d, 0, x, y, z, a, c)) ? MQ_ELEC_STEM7 + q : 0;
    if (z <= l[7] + eps) return MQ_ELEC_SLAB;
    return MQ_ELEC_NONE;
}

void or_geom_classify(const mq_geom* g, uint8_t* codes)
{
    const int nx = g->n_xy, ny = g->n_xy, nz = or_geom_nz(g);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                codes[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k)] =
                    (uint8_t)or_geom_classify_node(g, i, j, k);
}

double or_basis_value(int basis, int code, int ac_mode)
{
    const int rod = code != MQ_ELEC_NONE && code != MQ_ELEC_PLATE && code != MQ_ELEC_SLAB;
    const int q = code >= MQ_ELEC_STEM7 ? code - MQ_ELEC_STEM7
                : code >= MQ_ELEC_Q2 ? code - MQ_ELEC_Q2
                : code >= MQ_ELEC_STEM5 ? code - MQ_ELEC_STEM5
                : code >= MQ_ELEC_STEM3 ? code - MQ_ELEC_STEM3 : code - MQ_ELEC_Q1;
    const int upstream = code >= MQ_ELEC_Q1 && code < MQ_ELEC_PLATE;  /* Q1 or Stem3 */
    switch (basis) {
        case MQ_BASIS_RF: return rod ? ((q == 0 || q == 3) ? -1.0 : 1.0) : 0.0;
        case MQ_BASIS_AC:
            if (!upstream) return 0.0;
            if (ac_mode == 1 && q >= 2) return 0.0;
            return (q == 0 || q == 2) ? -1.0 : 1.0;
        case MQ_BASIS_DC_Q1: return upstream ? 1.0 : 0.0;
        case MQ_BASIS_DC_PLATE: return code == MQ_ELEC_PLATE ? 1.0 : 0.0;
        case MQ_BASIS_DC_DOWN: return (code >= MQ_ELEC_STEM5) ? 1.0 : 0.0;
        default: return 0.0;
    }
}

static double resid_at(const double* u, size_t id, size_t sy, size_t sz)
{
    return ((((u[id - 1] + u[id + 1]) + (u[id - sy] + u[id + sy])) + (u[id - sz] + u[id + sz]))) - 6.0 * u[id];
}

int or_laplace_sor(int nx, int ny, int nz, const uint8_t* mask, const double* fixed, double outer_bc, double tol,
                   uint32_t max_iter, double* phi, uint32_t* iters, double* residual)
{
    const size_t n = (size_t)nx * ny * nz, sy = (size_t)nx, sz = (size_t)nx * ny;
    uint8_t* fr = malloc(n);
    if (!fr) return -1;
    double maxfix = 0.0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const size_t id = (size_t)i + sy * j + sz * k;
                const int face = i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1;
                if (mask[id]) {
                    phi[id] = fixed[id];
                    fr[id] = 0;
                    if (fabs(fixed[id]) > maxfix) maxfix = fabs(fixed[id]);
                } else if (face) {
                    phi[id] = outer_bc;
                    fr[id] = 0;
                } else {
                    phi[id] = 0.0;
                    fr[id] = 1;
                }
            }
    double scale = fmax(maxfix, fabs(outer_bc));
    if (!(scale > 0.0)) scale = 1.0;
    /* optimal SOR factor of the model problem on the largest dimension */
    const int m = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
    const double w = 2.0 / (1.0 + sin(3.14159265358979323846 / (double)m));
    double res = INFINITY;
    uint32_t it = 0;
    for (; it < max_iter; ++it) {
        for (int color = 0; color < 2; ++color)
            for (int k = 1; k < nz - 1; ++k)
                for (int j = 1; j < ny - 1; ++j)
                    for (int i = 1 + ((j + k + color) & 1); i < nx - 1; i += 2) {
                        const size_t id = (size_t)i + sy * j + sz * k;
                        if (fr[id]) phi[id] += w * resid_at(phi, id, sy, sz) / 6.0;
                    }
        if ((it & 15) == 15 || it + 1 == max_iter) {
            double mx = 0.0;
            for (size_t id = 0; id < n; ++id)
                if (fr[id]) {
                    const double r = fabs(resid_at(phi, id, sy, sz));
                    if (r > mx) mx = r;
                }
            res = mx / scale;
            if (res <= tol) { ++it; break; }
        }
    }
    free(fr);
    if (iters) *iters = it;
    if (residual) *residual = res;
    return res <= tol ? 0 : 1;
}

void or_efield(const double* phi, int nx, int ny, int nz, double h, double* e)
{
#define IDX(i, j, k) ((size_t)(i) + (size_t)nx * ((size_t)(j) + (size_t)ny * (size_t)(k)))
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int im = i > 0 ? i - 1 : 0, ip = i + 1 < nx ? i + 1 : nx - 1;
                const int jm = j > 0 ? j - 1 : 0, jp = j + 1 < ny ? j + 1 : ny - 1;
                const int km = k > 0 ? k - 1 : 0, kp = k + 1 < nz ? k + 1 : nz - 1;
                double* o = e + 3 * IDX(i, j, k);
                o[0] = -(phi[IDX(ip, j, k)] - phi[IDX(im, j, k)]) / ((double)(ip - im) * h);
                o[1] = -(phi[IDX(i, jp, k)] - phi[IDX(i, jm, k)]) / ((double)(jp - jm) * h);
                o[2] = -(phi[IDX(i, j, kp)] - phi[IDX(i, j, km)]) / ((double)(kp - km) * h);
            }
#undef IDX
}
