/*
 * oracle_legacy.c -- CPU restatement of the Java sketch's potential export.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker of the GPU
 * regeneration of the legacy potentials, never part of the product.
 *
 * What it restates (reference/MQ_MZ_ejection_singleQ/, file:line):
 *   dimensions, electrode constants   MQ_MZ_ejection_singleQ.pde:27-73, 118, 152-162
 *   calculateElectrodeGeometry        GEOMETRY.pde:11-47
 *   define_RF, define_DC, define_AC   GEOMETRY.pde:54-127, 129-204, 206-288
 *   mapLayer                          GEOMETRY.pde:291-318
 *   withinQuadElectrode               GEOMETRY.pde:362-393
 *   SIGN                              GEOMETRY.pde:407-414
 *   withinHole                        GEOMETRY.pde:463-467
 *   withinSqueezedQuadElectrode       GEOMETRY.pde:512-550
 *   calculatePotential (sweeps, stop) POTENTIAL.pde:31-84
 *   convertMillimetersToArray         CONVERSIONS.pde:12-15
 *
 * Arithmetic: IEEE single precision in the source's operation order, no
 * contraction (-ffp-contract=off), as Java evaluates float expressions. The
 * Processing preprocessor turns every decimal literal into a float literal,
 * so 0.323, 0.02 and 0.0000000001 are float constants here too. Processing's
 * sqrt(float) is (float)Math.sqrt, the correctly rounded float square root
 * (sqrtf). The sweep is the sequential loop, one node after the other.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    float scl, step;
    float xDevice, yDevice, zDevice;
    float Qstart, Q1, GAP, PLATE, Q2, DH, De, spacing, U, U0;
    int nx, ny, nz;
} jgeo;

/* The sketch's globals (MQ_MZ_ejection_singleQ.pde:27-73, 152-162). */
static void jgeo_init(jgeo* g, float scl)
{
    g->scl = scl;
    g->xDevice = 16.0f;
    g->yDevice = 16.0f;
    g->Qstart = 3.0f;
    g->Q1 = 12.0f;
    g->GAP = 1.0f;
    g->PLATE = 1.6f;
    g->Q2 = 3.0f;
    g->zDevice = g->Qstart + g->Q1 + g->GAP + g->PLATE + g->GAP + g->Q2 + g->GAP + g->PLATE;
    g->DH = 7.0f;
    g->De = 6.35f;
    g->spacing = 0.323f * g->De;
    g->U = 1.0f;
    g->U0 = 0.0f;
    g->nx = (int)(g->xDevice * scl);
    g->ny = (int)(g->yDevice * scl);
    g->nz = (int)(g->zDevice * scl);
    g->step = 1.0f / scl;
}

static int map_layer(const jgeo* g, float z)
{
    const float L1 = 0.0f + g->Qstart;
    const float L2 = L1 + g->Q1;
    const float L3 = L2 + g->GAP;
    const float L4 = L3 + g->PLATE;
    const float L5 = L4 + g->GAP;
    const float L6 = L5 + g->Q2;
    const float L7 = L6 + g->GAP;
    const float L8 = L7 + g->PLATE;
    if (0.0f < z && z <= L1) return 1;
    if (L1 < z && z <= L2) return 2;
    if (L2 < z && z <= L3) return 3;
    if (L3 < z && z <= L4) return 4;
    if (L4 < z && z <= L5) return 5;
    if (L5 < z && z <= L6) return 6;
    if (L6 < z && z <= L7) return 7;
    if (L7 < z && z <= L8) return 8;
    return 0;
}

static int in_disc(float x, float y, float cx, float cy, float d)
{
    return sqrtf((x - cx) * (x - cx) + (y - cy) * (y - cy)) < d / 2.0f;
}

static int within_quad(const jgeo* g, float x, float y, float d)
{
    const float center = g->spacing / 2.0f + g->De / 2.0f;
    const float hx = g->xDevice / 2.0f, hy = g->yDevice / 2.0f;
    if (x < hx && y < hy) return in_disc(x, y, hx - center, hy - center, d);
    if (x > hx && y < hy) return in_disc(x, y, hx + center, hy - center, d);
    if (x < hx && y > hy) return in_disc(x, y, hx - center, hy + center, d);
    if (x > hx && y > hy) return in_disc(x, y, hx + center, hy + center, d);
    return 0;
}

static int within_squeezed(const jgeo* g, float x, float y, float z)
{
    const float alfa = 0.02f;
    const float d = g->De;
    const float center = g->spacing / 2.0f + g->De / 2.0f;
    const float sq = alfa * (g->Qstart + g->Q1 - z);
    const float hx = g->xDevice / 2.0f, hy = g->yDevice / 2.0f;
    if (x < hx && y < hy) return in_disc(x, y, hx - center, hy - center + sq, d);
    if (x > hx && y < hy) return in_disc(x, y, hx + center, hy - center + sq, d);
    if (x < hx && y > hy) return in_disc(x, y, hx - center, hy + center - sq, d);
    if (x > hx && y > hy) return in_disc(x, y, hx + center, hy + center - sq, d);
    return 0;
}

static int within_hole(const jgeo* g, float x, float y, float r)
{
    const float X = g->xDevice / 2.0f - x;
    const float Y = g->yDevice / 2.0f - y;
    return (X * X + Y * Y) < r * r;
}

static float sign_of(const jgeo* g, float x, float y)
{
    const float hx = g->xDevice / 2.0f, hy = g->yDevice / 2.0f;
    return ((x <= hx && y <= hy) || (x >= hx && y >= hy)) ? -1.0f : 1.0f;
}

/* define_RF / define_DC / define_AC of one node: potential and electrode flag. */
static void define_node(const jgeo* g, int basis, float x, float y, float z, float* pot, uint8_t* el)
{
    const float U = g->U, U0 = g->U0;
    const int slice = map_layer(g, z);
    const float hx = g->xDevice / 2.0f, hy = g->yDevice / 2.0f;
    int e = 0;
    float v = 0.0f;
    switch (slice) {
        case 2:
        case 3:
        case 5:
        case 6:
        case 7: {
            e = (slice == 2 || slice == 6) ? within_squeezed(g, x, y, z) : within_quad(g, x, y, 4.6f);
            if (!e) break;
            if (basis == MQ_LEGACY_RF) {
                v = sign_of(g, x, y) * U;
            } else if (basis == MQ_LEGACY_DC) {
                v = slice <= 3 ? U0 : -U;
            } else if (slice <= 3) {
                /* define_AC cases 2, 3: the first two branches are x >= xDevice/2;
                 * the third, ((x <= xD/2 && y <= yD/2) || y >= yD/2), holds for
                 * every remaining electrode node */
                if (x >= hx && y >= hy) v = U;
                else if (x >= hx && y <= hy) v = U;
                else if ((x <= hx && y <= hy) || y >= hy) v = -U;
                else e = 0;
            } else {
                v = U0;
            }
            break;
        }
        case 4:
            if (!within_hole(g, x, y, g->DH / 2.0f)) {
                e = 1;
                v = basis == MQ_LEGACY_DC ? U : U0;
            }
            break;
        case 8:
            e = 1;
            v = basis == MQ_LEGACY_DC ? -U : U0;
            break;
        default:
            break;
    }
    *pot = e ? v : 0.0f;
    *el = (uint8_t)e;
}

int or_legacy_dims(float scl, int* nx, int* ny, int* nz, int check[3])
{
    jgeo g;
    jgeo_init(&g, scl);
    *nx = g.nx;
    *ny = g.ny;
    *nz = g.nz;
    if (check) {
        const float x_check = g.xDevice / 2.0f + 1.0f;
        const float y_check = g.yDevice / 2.0f + 0.5f;
        const float z_check = g.Qstart + g.Q1 - 1.0f;
        check[0] = (int)(x_check * scl);
        check[1] = (int)(y_check * scl);
        check[2] = (int)(z_check * scl);
    }
    return 0;
}

void or_legacy_rasterize(float scl, int basis, float* phi, uint8_t* el)
{
    jgeo g;
    jgeo_init(&g, scl);
    for (int k = 0; k < g.nz; ++k) {
        const float z = (float)k * g.step;
        for (int i = 0; i < g.nx; ++i) {
            const float x = (float)i * g.step;
            for (int j = 0; j < g.ny; ++j) {
                const float y = (float)j * g.step;
                const size_t n = (size_t)i + (size_t)g.nx * ((size_t)j + (size_t)g.ny * k);
                define_node(&g, basis, x, y, z, &phi[n], &el[n]);
            }
        }
    }
}

/* One sweep of calculatePotential: z outer, x, y inner, in place. */
void or_legacy_sweep(int nx, int ny, int nz, float* p, const uint8_t* el)
{
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    for (int k = 1; k < nz - 1; ++k)
        for (int i = 1; i < nx - 1; ++i)
            for (int j = 1; j < ny - 1; ++j) {
                const size_t n = (size_t)i + sy * j + sz * k;
                if (el[n]) continue;
                p[n] = (p[n - 1] + p[n + 1] + p[n - sy] + p[n + sy] + p[n - sz] + p[n + sz]) / 6.0f;
            }
}

/* The stop rule's change at the check point, in percent. */
float or_legacy_dif(float prev, float now)
{
    return fabsf(100.0f * ((prev + 1e-10f) - now) / (prev + 1e-10f));
}

uint32_t or_legacy_sweeps(int nx, int ny, int nz, float* p, const uint8_t* el, const int check[3], float precision,
                          uint32_t max_iter, float* dif_out, float* trace)
{
    const size_t ch = (size_t)check[0] + (size_t)nx * ((size_t)check[1] + (size_t)ny * check[2]);
    float dif = 100.0f, prev = 1.0f;
    uint32_t nit = 0;
    while (nit < max_iter && dif > precision) {
        or_legacy_sweep(nx, ny, nz, p, el);
        dif = or_legacy_dif(prev, p[ch]);
        prev = p[ch];
        if (trace) trace[nit] = p[ch];
        nit++;
    }
    if (dif_out) *dif_out = dif;
    return nit;
}
