/*
 * lbm_oracle.c -- CPU oracle for the Inamuro two-phase D3Q15 time step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 *
 * A from-scratch C restatement of the reference CPU solver.  Internal arrays
 * use the reference's own traversal order, [x][y][z] with z fastest, so every
 * global sum runs x-outer / y / z-inner exactly as CPU/Inamuro.cpp does, and
 * every expression keeps the reference's left-to-right association.  Built
 * with gcc -O3 -std=c11 (ISO mode: no FMA contraction), it reproduces the
 * compiled reference bit for bit; oracle/ref_probe.cpp checks that.
 *
 * Extensions that the reference does not have (dual residual on the CPU,
 * tolerance / interval parameters, iteration watchdog, droplet velocity
 * vectors) are marked as such.
 *
 * How each part is pinned (none is left unpinned):
 *   the step, its stages and the initialisers -- bit for bit against the
 *     compiled reference (tests/test_oracle_golden.py, golden fixtures);
 *   the dual residual (a CUDA-only path of the reference) -- a closed form
 *     and a rebuild from the pinned stages (tests/test_oracle_dual_residual.py);
 *   the Poisson accelerators (CUDA-only) -- equivalences with the strict
 *     loop, closed forms and rebuilds (tests/test_oracle_accel.py);
 *   z walls and gravity (no reference counterpart) -- closed forms
 *     (tests/test_oracle_walls.py);
 *   the adjoint and the objectives -- central differences of the oracle's
 *     own forward map (tests/test_oracle_adjoint.py).
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define Q 15

/* D3Q15 tables, CPU/common.hpp:3-45. */
static const int EX[Q] = {0, 1, 0, 0, -1, 0, 0, 1, -1, 1, 1, -1, 1, -1, -1};
static const int EY[Q] = {0, 0, 1, 0, 0, -1, 0, 1, 1, -1, 1, -1, -1, 1, -1};
static const int EZ[Q] = {0, 0, 0, 1, 0, 0, -1, 1, 1, 1, -1, -1, -1, -1, 1};
static const double UC[Q] = {0.0, 1.0, 0.0, 0.0, -1.0, 0.0, 0.0,
                             1.0, -1.0, 1.0, 1.0, -1.0, 1.0, -1.0, -1.0};
static const double VC[Q] = {0.0, 0.0, 1.0, 0.0, 0.0, -1.0, 0.0,
                             1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, -1.0};
static const double WC[Q] = {0.0, 0.0, 0.0, 1.0, 0.0, 0.0, -1.0,
                             1.0, 1.0, 1.0, -1.0, -1.0, -1.0, -1.0, 1.0};
static const double EW[Q] = {2.0 / 9.0,
                             1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                             1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0,
                             1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0};
static const double HW[Q] = {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
                             0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
static const double FW[Q] = {-7.0 / 3.0,
                             1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0,
                             1.0 / 24.0, 1.0 / 24.0, 1.0 / 24.0, 1.0 / 24.0,
                             1.0 / 24.0, 1.0 / 24.0, 1.0 / 24.0, 1.0 / 24.0};

/* x-wall specular pairs (+x dir, -x dir), CPU/Inamuro.cpp:592-593. */
static const int SLIP_POS[5] = {1, 7, 9, 10, 12};
static const int SLIP_NEG[5] = {4, 8, 14, 13, 11};
/* Opposite direction c_opp = -c (bounce-back at the z walls, extension). */
static const int OPP[Q] = {0, 4, 5, 6, 1, 2, 3, 11, 12, 13, 14, 7, 8, 9, 10};

struct orc_sim {
    orc_params P;
    int lx, ly, lz;
    size_t n;
    double* f; double* g; double* h; double* tmp;
    double* arr[ORC_NUM_SCALAR_ARRAYS];
    int* xn; int* yn; int* zn;          /* stencil neighbour tables per axis, [coord*3 + (c+1)] */
    int first_step;                     /* function-static is_first_step, CPU/Inamuro.cpp:150 */
    int z_walls;                        /* extension: no-slip walls at z = 0 and lz-1 */
    double gz;                          /* extension: gravitational acceleration along z */
    int objective;                      /* extension: ORC_OBJ_* of orc_objective / orc_adjoint */
    int nthreads;
    uint64_t check_interval;
    double tol;
    uint64_t iter_cap;
    int dual;
    uint64_t last_iteration_count;
    int pressure_solve_call_count;
    uint64_t* schedule;                 /* frozen per-step iteration counts (extension) */
    int schedule_len, schedule_pos;
    orc_poisson_accel accel;            /* reference GPU Poisson accelerators */
    orc_poisson_diag diag;
    char err[256];
};

#define IX(s, x, y, z) ((((size_t)(x)) * (size_t)(s)->ly + (size_t)(y)) * (size_t)(s)->lz + (size_t)(z))

#define A(s, id) ((s)->arr[(id)])

static void set_err(orc_sim* s, const char* msg)
{
    snprintf(s->err, sizeof(s->err), "%s", msg);
}

const char* orc_last_error(const orc_sim* s) { return s->err; }

void orc_set_threads(orc_sim* s, int nthreads) { s->nthreads = nthreads > 0 ? nthreads : 1; }

/* z neighbours of the stencils: periodic (CPU/Inamuro.hpp:240-247), or, with
 * the z walls of the extension, the cell itself beyond a wall (zero normal
 * gradient about a wall half a cell outside the first and last layer). */
static void build_z_neighbours(orc_sim* s)
{
    for (int z = 0; z < s->lz; ++z)
        for (int c = -1; c <= 1; ++c) {
            int zp = z + c;
            if (s->z_walls) {
                if (zp >= s->lz) zp = s->lz - 1;
                if (zp < 0) zp = 0;
            } else {
                if (zp >= s->lz) zp = 0;
                if (zp < 0) zp = s->lz - 1;
            }
            s->zn[z * 3 + c + 1] = zp;
        }
}

orc_sim* orc_create(const orc_params* params)
{
    if (!params || params->lx < 3 || params->ly < 1 || params->lz < 1) return NULL;
    orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
    if (!s) return NULL;
    s->P = *params;
    s->lx = params->lx; s->ly = params->ly; s->lz = params->lz;
    s->n = (size_t)s->lx * (size_t)s->ly * (size_t)s->lz;
    s->f = (double*)calloc((size_t)Q * s->n, sizeof(double));
    s->g = (double*)calloc((size_t)Q * s->n, sizeof(double));
    s->h = (double*)calloc((size_t)Q * s->n, sizeof(double));
    s->tmp = (double*)calloc((size_t)Q * s->n, sizeof(double));
    int ok = s->f && s->g && s->h && s->tmp;
    for (int i = 0; i < ORC_NUM_SCALAR_ARRAYS; ++i) {
        A(s, i) = (double*)calloc(s->n, sizeof(double));
        ok = ok && A(s, i);
    }
    s->xn = (int*)malloc(sizeof(int) * 3 * (size_t)s->lx);
    s->yn = (int*)malloc(sizeof(int) * 3 * (size_t)s->ly);
    s->zn = (int*)malloc(sizeof(int) * 3 * (size_t)s->lz);
    ok = ok && s->xn && s->yn && s->zn;
    if (!ok) { orc_destroy(s); return NULL; }
    /* Stencil neighbours, CPU/Inamuro.hpp:230-248: x mirror (lx -> lx-2,
     * -1 -> 1), y and z periodic. */
    for (int x = 0; x < s->lx; ++x)
        for (int c = -1; c <= 1; ++c) {
            int xp = x + c;
            if (xp >= s->lx) xp = s->lx - 2;
            if (xp < 0) xp = 1;
            s->xn[x * 3 + c + 1] = xp;
        }
    for (int y = 0; y < s->ly; ++y)
        for (int c = -1; c <= 1; ++c) {
            int yp = y + c;
            if (yp >= s->ly) yp = 0;
            if (yp < 0) yp = s->ly - 1;
            s->yn[y * 3 + c + 1] = yp;
        }
    build_z_neighbours(s);
    /* initializeArrays, CPU/Inamuro.cpp:511-548. */
    for (size_t i = 0; i < s->n; ++i) {
        A(s, ORC_RHO)[i] = params->rho_G;
        A(s, ORC_FEI)[i] = params->fei_min;
    }
    s->first_step = 1;
    s->nthreads = 1;
    s->check_interval = 100;
    s->tol = 1.0e-3;
    s->iter_cap = 0;
    s->dual = 0;
    orc_default_poisson_accel(&s->accel);
    return s;
}

void orc_destroy(orc_sim* s)
{
    if (!s) return;
    free(s->f); free(s->g); free(s->h); free(s->tmp);
    for (int i = 0; i < ORC_NUM_SCALAR_ARRAYS; ++i) free(A(s, i));
    free(s->xn); free(s->yn); free(s->zn);
    free(s->schedule);
    free(s);
}

int orc_set_poisson(orc_sim* s, uint64_t check_interval, double tol, uint64_t iter_cap, int dual)
{
    if (check_interval == 0) { set_err(s, "poisson check interval must be positive"); return -1; }
    if (!(tol > 0.0)) { set_err(s, "poisson tolerance must be positive"); return -1; }
    if (iter_cap != 0 && iter_cap % check_interval != 0) {
        set_err(s, "Poisson iteration limit must be a multiple of check interval");
        return -1;
    }
    s->check_interval = check_interval;
    s->tol = tol;
    s->iter_cap = iter_cap;
    s->dual = dual ? 1 : 0;
    return 0;
}

int orc_is_first_step(const orc_sim* s) { return s->first_step; }
void orc_set_first_step(orc_sim* s, int first) { s->first_step = first ? 1 : 0; }

/* ------------------------------------------------------------------ */
/* Layout conversion at the API boundary                               */
/* ------------------------------------------------------------------ */

int orc_get(const orc_sim* s, int id, double* out)
{
    const int lx = s->lx, ly = s->ly, lz = s->lz;
    const size_t n = s->n;
    if (id >= 0 && id < ORC_NUM_SCALAR_ARRAYS) {
        const double* src = s->arr[id];
        for (int z = 0; z < lz; ++z)
            for (int y = 0; y < ly; ++y)
                for (int x = 0; x < lx; ++x)
                    out[((size_t)z * ly + y) * lx + x] = src[IX(s, x, y, z)];
        return 0;
    }
    const double* d = id == ORC_F ? s->f : id == ORC_G ? s->g : id == ORC_H ? s->h : NULL;
    if (!d) return -1;
    for (int q = 0; q < Q; ++q)
        for (int z = 0; z < lz; ++z)
            for (int y = 0; y < ly; ++y)
                for (int x = 0; x < lx; ++x)
                    out[(size_t)q * n + ((size_t)z * ly + y) * lx + x] = d[(size_t)q * n + IX(s, x, y, z)];
    return 0;
}

int orc_set(orc_sim* s, int id, const double* in)
{
    const int lx = s->lx, ly = s->ly, lz = s->lz;
    const size_t n = s->n;
    if (id >= 0 && id < ORC_NUM_SCALAR_ARRAYS) {
        double* dst = s->arr[id];
        for (int z = 0; z < lz; ++z)
            for (int y = 0; y < ly; ++y)
                for (int x = 0; x < lx; ++x)
                    dst[IX(s, x, y, z)] = in[((size_t)z * ly + y) * lx + x];
        return 0;
    }
    double* d = id == ORC_F ? s->f : id == ORC_G ? s->g : id == ORC_H ? s->h : NULL;
    if (!d) return -1;
    for (int q = 0; q < Q; ++q)
        for (int z = 0; z < lz; ++z)
            for (int y = 0; y < ly; ++y)
                for (int x = 0; x < lx; ++x)
                    d[(size_t)q * n + IX(s, x, y, z)] = in[(size_t)q * n + ((size_t)z * ly + y) * lx + x];
    return 0;
}

/* ------------------------------------------------------------------ */
/* Initial conditions                                                  */
/* ------------------------------------------------------------------ */

/* Inamuro::initializeDroplets of the CPU tree, CPU/Inamuro.cpp:550-587. */
void orc_init_two_droplets_sharp(orc_sim* s)
{
    const orc_params* P = &s->P;
    const double radius = P->DD / 2.0;
    const double cy = s->ly / 2.0;
    const double cz1 = s->lz / 2.0 + P->DD;
    const double cz2 = s->lz / 2.0 - P->DD;
    const double velocity = 0.035;
    for (int x = 0; x < s->lx; ++x)
        for (int y = 0; y < s->ly; ++y)
            for (int z = 0; z < s->lz; ++z) {
                const double dx = x, dy = y - cy;
                const double dz1 = z - cz1, dz2 = z - cz2;
                const double r1 = sqrt(dx * dx + dy * dy + dz1 * dz1);
                const double r2 = sqrt(dx * dx + dy * dy + dz2 * dz2);
                const size_t i = IX(s, x, y, z);
                if (r1 < radius) {
                    A(s, ORC_FEI)[i] = P->fei_max;
                    A(s, ORC_RHO)[i] = P->rho_L;
                    A(s, ORC_W)[i] = -velocity;
                }
                if (r2 < radius) {
                    A(s, ORC_FEI)[i] = P->fei_max;
                    A(s, ORC_RHO)[i] = P->rho_L;
                    A(s, ORC_W)[i] = velocity;
                }
            }
}

/* place_droplet of the GPU tree, GPU/Inamuro.cpp:499-533.  The reference sets
 * only w = vz * liquid_fraction; vx and vy (extension) follow the same rule. */
void orc_place_droplet(orc_sim* s, double cx, double cy, double cz,
                       double vx, double vy, double vz, int profile, double interface_width)
{
    const orc_params* P = &s->P;
    const double radius = P->DD / 2.0;
    const double width = interface_width > 0.0 ? interface_width : 2.0;
    for (int x = 0; x < s->lx; ++x)
        for (int y = 0; y < s->ly; ++y)
            for (int z = 0; z < s->lz; ++z) {
                const double dx = x - cx;
                const double dy = y - cy;
                const double dz = z - cz;
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                double lf = 0.0;
                if (profile == 1)
                    lf = 0.5 * (1.0 - tanh((r - radius) / width));
                else if (r < radius)
                    lf = 1.0;
                if (lf > 0.0) {
                    const double local_fei = P->fei_min + lf * (P->fei_max - P->fei_min);
                    const size_t i = IX(s, x, y, z);
                    if (local_fei > A(s, ORC_FEI)[i]) {
                        A(s, ORC_FEI)[i] = local_fei;
                        A(s, ORC_RHO)[i] = P->rho_G + lf * (P->rho_L - P->rho_G);
                        if (vx != 0.0) A(s, ORC_U)[i] = vx * lf;
                        if (vy != 0.0) A(s, ORC_V)[i] = vy * lf;
                        A(s, ORC_W)[i] = vz * lf;
                    }
                }
            }
}

/* ------------------------------------------------------------------ */
/* Scalar helpers                                                      */
/* ------------------------------------------------------------------ */

/* Inamuro::EOS, CPU/Inamuro.cpp:843-847: p0 = fei*T/(1 - b*fei) - a*fei^2. */
double orc_eos(const orc_params* P, double fei)
{
    return fei * P->T / (1.0 - P->b * fei) - P->a * fei * fei;
}

/* Density map of getMacro, CPU/Inamuro.cpp:296, 313-327. */
double orc_rho_of_phi(const orc_params* P, double fei_local)
{
    const double fei_avg = (P->fei_L + P->fei_G) / 2.0;
    if (fei_local <= P->fei_G) return P->rho_G;
    if (fei_local >= P->fei_L) return P->rho_L;
    const double arg = (fei_local - fei_avg) / (P->fei_L - P->fei_G) * M_PI;
    return (P->rho_L - P->rho_G) / 2.0 * (sin(arg) + 1.0) + P->rho_G;
}

/* ------------------------------------------------------------------ */
/* Stencils                                                            */
/* ------------------------------------------------------------------ */

/* Inamuro::firstord, CPU/Inamuro.hpp:213-266: the 14 moving directions in
 * k order, accumulated from 0.0, then divided by 10.0. */
static void firstord(orc_sim* s, const double* var, double* vx_, double* vy_, double* vz_)
{
    const int lx = s->lx, ly = s->ly, lz = s->lz;
#pragma omp parallel for num_threads(s->nthreads) schedule(static) if (s->nthreads > 1)
    for (int x = 0; x < lx; ++x)
        for (int y = 0; y < ly; ++y)
            for (int z = 0; z < lz; ++z) {
                double ax = 0.0, ay = 0.0, az = 0.0;
                for (int k = 1; k < Q; ++k) {
                    const int xp = s->xn[x * 3 + EX[k] + 1];
                    const int yp = s->yn[y * 3 + EY[k] + 1];
                    const int zp = s->zn[z * 3 + EZ[k] + 1];
                    const double val = var[IX(s, xp, yp, zp)];
                    ax += val * UC[k];
                    ay += val * VC[k];
                    az += val * WC[k];
                }
                const size_t i = IX(s, x, y, z);
                vx_[i] = ax / 10.0;
                vy_[i] = ay / 10.0;
                vz_[i] = az / 10.0;
            }
}

/* Inamuro::secondord, CPU/Inamuro.hpp:268-314: (sum - 14*centre) / 5. */
static void secondord(orc_sim* s, const double* var, double* lap)
{
    const int lx = s->lx, ly = s->ly, lz = s->lz;
#pragma omp parallel for num_threads(s->nthreads) schedule(static) if (s->nthreads > 1)
    for (int x = 0; x < lx; ++x)
        for (int y = 0; y < ly; ++y)
            for (int z = 0; z < lz; ++z) {
                double acc = 0.0;
                for (int k = 1; k < Q; ++k) {
                    const int xp = s->xn[x * 3 + EX[k] + 1];
                    const int yp = s->yn[y * 3 + EY[k] + 1];
                    const int zp = s->zn[z * 3 + EZ[k] + 1];
                    acc += var[IX(s, xp, yp, zp)];
                }
                const size_t i = IX(s, x, y, z);
                lap[i] = (acc - 14.0 * var[i]) / 5.0;
            }
}

static void gradients(orc_sim* s)
{
    /* CPU/Inamuro.cpp:140-148, same call order. */
    firstord(s, A(s, ORC_U), A(s, ORC_U_X), A(s, ORC_U_Y), A(s, ORC_U_Z));
    firstord(s, A(s, ORC_V), A(s, ORC_V_X), A(s, ORC_V_Y), A(s, ORC_V_Z));
    firstord(s, A(s, ORC_W), A(s, ORC_W_X), A(s, ORC_W_Y), A(s, ORC_W_Z));
    firstord(s, A(s, ORC_RHO), A(s, ORC_RHO_X), A(s, ORC_RHO_Y), A(s, ORC_RHO_Z));
    firstord(s, A(s, ORC_FEI), A(s, ORC_FEI_X), A(s, ORC_FEI_Y), A(s, ORC_FEI_Z));
    secondord(s, A(s, ORC_U), A(s, ORC_U_LAP));
    secondord(s, A(s, ORC_V), A(s, ORC_V_LAP));
    secondord(s, A(s, ORC_W), A(s, ORC_W_LAP));
    secondord(s, A(s, ORC_FEI), A(s, ORC_FEI_LAP));
}

/* ------------------------------------------------------------------ */
/* Collision of f and g                                                */
/* ------------------------------------------------------------------ */

/* Inamuro::collision, CPU/Inamuro.cpp:134-250. */
static void collision(orc_sim* s)
{
    const orc_params* P = &s->P;
    gradients(s);
    const double tmp1 = (P->mu_L - P->mu_G) / (P->rho_L - P->rho_G);
    const size_t n = s->n;
    const int first = s->first_step;
#pragma omp parallel for num_threads(s->nthreads) schedule(static) if (s->nthreads > 1)
    for (size_t i = 0; i < n; ++i) {
        const double rho = A(s, ORC_RHO)[i];
        const double fei = A(s, ORC_FEI)[i];
        const double u = A(s, ORC_U)[i], v = A(s, ORC_V)[i], w = A(s, ORC_W)[i];
        const double fx = A(s, ORC_FEI_X)[i], fy = A(s, ORC_FEI_Y)[i], fz = A(s, ORC_FEI_Z)[i];
        const double rx = A(s, ORC_RHO_X)[i], ry = A(s, ORC_RHO_Y)[i], rz = A(s, ORC_RHO_Z)[i];
        const double ux = A(s, ORC_U_X)[i], uy = A(s, ORC_U_Y)[i], uz = A(s, ORC_U_Z)[i];
        const double vx = A(s, ORC_V_X)[i], vy = A(s, ORC_V_Y)[i], vz = A(s, ORC_V_Z)[i];
        const double wx = A(s, ORC_W_X)[i], wy = A(s, ORC_W_Y)[i], wz = A(s, ORC_W_Z)[i];
        const double flap = A(s, ORC_FEI_LAP)[i];
        const double ulap = A(s, ORC_U_LAP)[i], vlap = A(s, ORC_V_LAP)[i], wlap = A(s, ORC_W_LAP)[i];

        const double mu = (rho - P->rho_G) * tmp1 + P->mu_G;
        const double sum_fei = fx * fx + fy * fy + fz * fz;
        const double sum_rho = rx * rx + ry * ry + rz * rz;
        const double fei2sum = flap;
        const double usq = u * u + v * v + w * w;
        const double p0 = fei * P->T / (1.0 - P->b * fei) - P->a * fei * fei;

        double fequ[Q], gequ[Q];
        for (int k = 0; k < Q; ++k) {
            const double un = UC[k] * u + VC[k] * v + WC[k] * w;
            const double fei_ei = UC[k] * fx + VC[k] * fy + WC[k] * fz;
            const double rho_ei = UC[k] * rx + VC[k] * ry + WC[k] * rz;
            const double ci_sq = UC[k] * UC[k] + VC[k] * VC[k] + WC[k] * WC[k];
            const double Gfei = 4.5 * (fei_ei * fei_ei) - 1.5 * sum_fei * ci_sq;
            const double Grho = 4.5 * (rho_ei * rho_ei) - 1.5 * sum_rho * ci_sq;
            fequ[k] = HW[k] * fei +
                      FW[k] * (p0 - P->k_f * fei * fei2sum - P->k_f * sum_fei / 6.0) +
                      3.0 * EW[k] * fei * un +
                      EW[k] * P->k_f * Gfei;
            const double velPart = 1.0 + 3.0 * un - 1.5 * usq + 4.5 * un * un +
                1.5 * (P->taug - 0.5) * 2.0 *
                    (UC[k] * UC[k] * ux + UC[k] * VC[k] * uy + UC[k] * WC[k] * uz +
                     VC[k] * UC[k] * vx + VC[k] * VC[k] * vy + VC[k] * WC[k] * vz +
                     WC[k] * UC[k] * wx + WC[k] * VC[k] * wy + WC[k] * WC[k] * wz);
            gequ[k] = EW[k] * velPart +
                      EW[k] * P->k_g / rho * Grho -
                      2.0 / 3.0 * FW[k] * P->k_g / rho * sum_rho;
        }
        if (first) {
            for (int k = 0; k < Q; ++k) {
                s->f[(size_t)k * n + i] = fequ[k];
                s->g[(size_t)k * n + i] = gequ[k];
            }
        } else {
            for (int k = 0; k < Q; ++k) {
                double* fk = &s->f[(size_t)k * n + i];
                double* gk = &s->g[(size_t)k * n + i];
                *fk = *fk - (*fk - fequ[k]) / P->tauf;
                *gk = *gk - (*gk - gequ[k]) / P->taug +
                      3.0 * EW[k] / rho * mu * (UC[k] * ulap + VC[k] * vlap + WC[k] * wlap);
            }
        }
        if (s->gz != 0.0) {
            /* extension: buoyant acceleration a = g (1 - rho_G/rho) along z on
             * the predicted velocity, sum_k c_k 3 E_k c_kz a = a e_z */
            const double az = s->gz * (1.0 - P->rho_G / rho);
            for (int k = 0; k < Q; ++k)
                if (EZ[k] != 0) s->g[(size_t)k * n + i] += 3.0 * EW[k] * WC[k] * az;
        }
    }
    s->first_step = 0;
}

/* ------------------------------------------------------------------ */
/* Streaming and boundary                                              */
/* ------------------------------------------------------------------ */

/* Inamuro::stream, CPU/Inamuro.cpp:252-287: periodic on all three axes.  The
 * reference pushes into a temporary; the pull below is the same permutation. */
static void stream(orc_sim* s, double* dist)
{
    const int lx = s->lx, ly = s->ly, lz = s->lz;
    const size_t n = s->n;
#pragma omp parallel for num_threads(s->nthreads) schedule(static) if (s->nthreads > 1)
    for (int q = 0; q < Q; ++q) {
        const double* src = dist + (size_t)q * n;
        double* dst = s->tmp + (size_t)q * n;
        for (int x = 0; x < lx; ++x) {
            const int xs = (x - EX[q] + lx) % lx;
            for (int y = 0; y < ly; ++y) {
                const int ys = (y - EY[q] + ly) % ly;
                for (int z = 0; z < lz; ++z) {
                    const int zr = z - EZ[q];
                    if (s->z_walls && (zr < 0 || zr >= lz)) {
                        /* extension: halfway bounce-back, the population that
                         * left this cell towards the wall returns reversed */
                        dst[IX(s, x, y, z)] = dist[(size_t)OPP[q] * n + IX(s, x, y, z)];
                        continue;
                    }
                    const int zs = (zr + lz) % lz;
                    dst[IX(s, x, y, z)] = src[IX(s, xs, ys, zs)];
                }
            }
        }
    }
    memcpy(dist, s->tmp, sizeof(double) * (size_t)Q * n);
}

/* Inamuro::slipBounceBack, CPU/Inamuro.cpp:589-623. */
static void slip_bounce_back(orc_sim* s, double* dist)
{
    const size_t n = s->n;
    for (int y = 0; y < s->ly; ++y)
        for (int z = 0; z < s->lz; ++z)
            for (int r = 0; r < 5; ++r)
                dist[(size_t)SLIP_POS[r] * n + IX(s, 0, y, z)] = dist[(size_t)SLIP_NEG[r] * n + IX(s, 0, y, z)];
    for (int y = 0; y < s->ly; ++y)
        for (int z = 0; z < s->lz; ++z)
            for (int r = 0; r < 5; ++r)
                dist[(size_t)SLIP_NEG[r] * n + IX(s, s->lx - 1, y, z)] =
                    dist[(size_t)SLIP_POS[r] * n + IX(s, s->lx - 1, y, z)];
}

/* ------------------------------------------------------------------ */
/* Macro fields                                                        */
/* ------------------------------------------------------------------ */

/* Inamuro::getMacro, CPU/Inamuro.cpp:294-346. */
static void get_macro(orc_sim* s)
{
    const size_t n = s->n;
    const orc_params* P = &s->P;
#pragma omp parallel for num_threads(s->nthreads) schedule(static) if (s->nthreads > 1)
    for (size_t i = 0; i < n; ++i) {
        double fei_local = 0.0;
        for (int k = 0; k < Q; ++k) fei_local += s->f[(size_t)k * n + i];
        A(s, ORC_FEI)[i] = fei_local;
        A(s, ORC_RHO)[i] = orc_rho_of_phi(P, fei_local);
        double u = 0.0, v = 0.0, w = 0.0;
        for (int k = 1; k < Q; ++k) {
            const double g_val = s->g[(size_t)k * n + i];
            u += UC[k] * g_val;
            v += VC[k] * g_val;
            w += WC[k] * g_val;
        }
        A(s, ORC_U)[i] = u;
        A(s, ORC_V)[i] = v;
        A(s, ORC_W)[i] = w;
    }
}

/* ------------------------------------------------------------------ */
/* Pressure Poisson                                                    */
/* ------------------------------------------------------------------ */

/* Inamuro::collision_p, CPU/Inamuro.cpp:755-778. */
static void collision_p_on(orc_sim* s, double* hdist)
{
    const size_t n = s->n;
    const double* rho = A(s, ORC_RHO);
    const double* p = A(s, ORC_P);
    const double* ux = A(s, ORC_U_X);
    const double* vy = A(s, ORC_V_Y);
    const double* wz = A(s, ORC_W_Z);
#pragma omp parallel for num_threads(s->nthreads) schedule(static) if (s->nthreads > 1)
    for (size_t i = 0; i < n; ++i) {
        const double tauh = 1.0 / rho[i] + 0.5;
        for (int k = 0; k < Q; ++k) {
            const double hequ = EW[k] * p[i];
            double* hk = &hdist[(size_t)k * n + i];
            *hk = *hk - (*hk - hequ) / tauh - (EW[k] / 3.0) * (ux[i] + vy[i] + wz[i]);
        }
    }
}

/* Inamuro::getp, CPU/Inamuro.cpp:779-793. */
static void sum_h_into(orc_sim* s, const double* hdist, double* out)
{
    const size_t n = s->n;
#pragma omp parallel for num_threads(s->nthreads) schedule(static) if (s->nthreads > 1)
    for (size_t i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int k = 0; k < Q; ++k) acc += hdist[(size_t)k * n + i];
        out[i] = acc;
    }
}

/* Inamuro::getError, CPU/Inamuro.cpp:726-753.  Returns 0 on success. */
static int get_error(orc_sim* s, double* error)
{
    const size_t n = s->n;
    double err1 = 0.0, err2 = 0.0;
    double* p = A(s, ORC_P);
    double* pp = A(s, ORC_P_PREV);
    for (size_t i = 0; i < n; ++i) {
        const double pc = p[i];
        err1 += fabs(pc - pp[i]);
        err2 += fabs(pc);
        pp[i] = pc;
    }
    s->diag.pressure_l1_delta = err1;
    s->diag.pressure_l1_norm = err2;
    if (!isfinite(err1) || !isfinite(err2)) {
        set_err(s, "Poisson pressure/residual contains NaN/Inf");
        return -1;
    }
    *error = (err2 > 0.0) ? (err1 / err2) : ((err1 == 0.0) ? 0.0 : INFINITY);
    return 0;
}

static double rel_or_inf(double num, double den)
{
    if (den > 0.0) return num / den;
    return num == 0.0 ? 0.0 : INFINITY;
}

/* Kinetic fixed-point residual of the frozen split map.  The reference has it
 * only on the GPU (GPU/InamuroCUDA.cu:1987-2057, 1114-1171,
 * DUAL_RESIDUAL_CONTRACT.md:24-48); this is its CPU restatement, evaluated
 * after the gauge, with sums in x-outer order.  Returns 0 on success. */
static int fixed_point_residual(orc_sim* s, double gauge_target)
{
    const size_t n = s->n;
    /* F(h, p) = BS[C(h, p)] in the scratch buffer s->tmp via a copy. */
    double* image = (double*)malloc(sizeof(double) * (size_t)Q * n);
    double* p_from_h = (double*)malloc(sizeof(double) * n);
    if (!image || !p_from_h) { free(image); free(p_from_h); set_err(s, "out of memory"); return -1; }
    memcpy(image, s->h, sizeof(double) * (size_t)Q * n);
    collision_p_on(s, image);
    stream(s, image);
    slip_bounce_back(s, image);
    sum_h_into(s, s->h, p_from_h);
    double S[6] = {0, 0, 0, 0, 0, 0};
    const double* p = A(s, ORC_P);
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < Q; ++k) {
            const double hv = s->h[(size_t)k * n + i];
            const double iv = image[(size_t)k * n + i];
            const double ge = EW[k] * gauge_target;
            S[0] += fabs(hv - iv);
            S[1] += fabs(hv - ge);
            S[2] += fabs(iv - ge);
        }
        S[3] += fabs(p[i] - p_from_h[i]);
        S[4] += fabs(p[i] - gauge_target);
        S[5] += fabs(p_from_h[i] - gauge_target);
    }
    free(image);
    free(p_from_h);
    for (int t = 0; t < 6; ++t)
        if (!isfinite(S[t])) { set_err(s, "kinetic fixed-point residual contains NaN/Inf"); return -1; }
    orc_poisson_diag* d = &s->diag;
    d->fixed_point_h_l1 = S[0];
    d->fixed_point_h_scale = S[1] > S[2] ? S[1] : S[2];
    d->fixed_point_h_relative = rel_or_inf(S[0], d->fixed_point_h_scale);
    d->fixed_point_p_l1 = S[3];
    d->fixed_point_p_scale = S[4] > S[5] ? S[4] : S[5];
    d->fixed_point_p_relative = rel_or_inf(S[3], d->fixed_point_p_scale);
    d->fixed_point_relative = d->fixed_point_h_relative > d->fixed_point_p_relative
                                  ? d->fixed_point_h_relative : d->fixed_point_p_relative;
    if (!isfinite(d->fixed_point_relative)) {
        set_err(s, "normalized kinetic fixed-point residual contains NaN/Inf");
        return -1;
    }
    d->fixed_point_evaluated = 1;
    return 0;
}

int orc_fixed_point_residual(orc_sim* s, double gauge_target)
{
    return fixed_point_residual(s, gauge_target);
}

void orc_poisson_iterations(orc_sim* s, int k)
{
    for (int it = 0; it < k; ++it) {
        collision_p_on(s, s->h);
        stream(s, s->h);
        slip_bounce_back(s, s->h);
        sum_h_into(s, s->h, A(s, ORC_P));
    }
}

/* Inamuro::solvePressurePoisson, CPU/Inamuro.cpp:624-724.  p_prev is the
 * reference's function-static pressure_prev, zero-initialised once per
 * simulation and kept across steps. */
static int solve_pressure_poisson_rec(orc_sim* s, double* rec_h, double* rec_p);
static int solve_pressure_poisson_accel(orc_sim* s);
static int accel_active(const orc_sim* s);

static int solve_pressure_poisson(orc_sim* s)
{
    if (accel_active(s)) return solve_pressure_poisson_accel(s);
    return solve_pressure_poisson_rec(s, NULL, NULL);
}

/* Start of a solve (CPU/Inamuro.cpp:624-668, GPU/InamuroCUDA.cu:1829-1925):
 * diagnostics reset, source statistics, projection of the source mean out of
 * u_x, and the gauge target mean(p).  0 = ok. */
static int poisson_prologue(orc_sim* s, double* gauge_target_out)
{
    const size_t n = s->n;
    ++s->pressure_solve_call_count;
    const uint64_t prev = s->last_iteration_count;
    orc_poisson_diag* d = &s->diag;
    memset(d, 0, sizeof(*d));
    d->pressure_solve_call_count = s->pressure_solve_call_count;
    d->prev_iteration_count = prev;
    d->fixed_point_h_l1 = d->fixed_point_h_scale = d->fixed_point_h_relative = NAN;
    d->fixed_point_p_l1 = d->fixed_point_p_scale = d->fixed_point_p_relative = NAN;
    d->fixed_point_relative = NAN;
    d->block_low_frequency_fraction = NAN;
    d->anderson_alpha = NAN;

    const double cell_count = (double)s->lx * s->ly * s->lz;
    double divergence_sum = 0.0, pressure_sum = 0.0, divergence_abs_sum = 0.0;
    double* ux = A(s, ORC_U_X);
    const double* vy = A(s, ORC_V_Y);
    const double* wz = A(s, ORC_W_Z);
    double* p = A(s, ORC_P);
    for (size_t i = 0; i < n; ++i) {
        const double div = ux[i] + vy[i] + wz[i];
        divergence_sum += div;
        divergence_abs_sum += fabs(div);
        pressure_sum += p[i];
    }
    const double divergence_mean = divergence_sum / cell_count;
    const double gauge_target = pressure_sum / cell_count;
    d->divergence_mean = divergence_mean;
    d->pressure_gauge_target = gauge_target;
    /* GPU-only source statistics, GPU/InamuroCUDA.cu:1891-1893. */
    d->source_mean = -divergence_mean / 3.0;
    d->source_abs_mean = divergence_abs_sum / (3.0 * cell_count);
    if (!isfinite(divergence_mean) || !isfinite(gauge_target)) {
        set_err(s, "Poisson source/gauge contains NaN/Inf");
        return -1;
    }
    for (size_t i = 0; i < n; ++i) ux[i] -= divergence_mean;
    /* GPU/InamuroCUDA.cu:1897-1920: the source's sum after the projection. */
    double projected_sum = 0.0;
    for (size_t i = 0; i < n; ++i) projected_sum += ux[i] + vy[i] + wz[i];
    d->projected_source_mean = -projected_sum / (3.0 * cell_count);
    *gauge_target_out = gauge_target;
    return 0;
}

/* rec_h / rec_p (extension, for the adjoint): when non-NULL, receive the h and
 * p each iteration starts from (after any gauge), iteration-major. */
static int solve_pressure_poisson_rec(orc_sim* s, double* rec_h, double* rec_p)
{
    const size_t n = s->n;
    orc_poisson_diag* d = &s->diag;
    const double cell_count = (double)s->lx * s->ly * s->lz;
    double gauge_target = 0.0;
    if (poisson_prologue(s, &gauge_target) != 0) return -1;
    double* p = A(s, ORC_P);

    const uint64_t forced = s->schedule_pos < s->schedule_len ? s->schedule[s->schedule_pos] : 0;
    uint64_t iter = 0;
    for (;;) {
        if (iter == UINT64_MAX) { set_err(s, "Poisson iteration counter overflow"); return -2; }
        if (rec_h) {
            memcpy(rec_h + iter * (size_t)Q * n, s->h, sizeof(double) * (size_t)Q * n);
            memcpy(rec_p + iter * n, p, sizeof(double) * n);
        }
        ++iter;
        collision_p_on(s, s->h);
        stream(s, s->h);
        slip_bounce_back(s, s->h);
        sum_h_into(s, s->h, p);
        if (iter % s->check_interval == 0) {
            double cur = 0.0;
            for (size_t i = 0; i < n; ++i) cur += p[i];
            const double shift = cur / cell_count - gauge_target;
            if (!isfinite(shift)) { set_err(s, "pressure gauge is NaN/Inf"); return -1; }
            if (fabs(shift) > d->pressure_gauge_max_shift) d->pressure_gauge_max_shift = fabs(shift);
            for (size_t i = 0; i < n; ++i) {
                p[i] -= shift;
                for (int k = 0; k < Q; ++k) s->h[(size_t)k * n + i] -= EW[k] * shift;
            }
            int fp_ok = 1;
            if (s->dual) {
                if (fixed_point_residual(s, gauge_target) != 0) return -1;
                fp_ok = d->fixed_point_relative < s->tol;
            }
            double error = 0.0;
            if (get_error(s, &error) != 0) return -1;
            if (!isfinite(error)) { set_err(s, "Poisson residual is NaN/Inf"); return -1; }
            d->relative_error = error;
            d->iteration_count = iter;
            if (forced) {
                if (iter >= forced) break;
                continue;
            }
            if (error < s->tol && fp_ok) { d->converged = 1; break; }
            if (s->iter_cap != 0 && iter >= s->iter_cap) {
                d->hit_max_iterations = 1;
                d->fallback_count = 1;  /* GPU/InamuroCUDA.cu:2335-2338 */
                break;
            }
        }
    }
    s->last_iteration_count = iter;
    d->iteration_count = iter;
    if (forced) ++s->schedule_pos;
    return 0;
}

/* ------------------------------------------------------------------ */
/* The reference GPU solver's Poisson accelerators                     */
/* ------------------------------------------------------------------ */
/* Restated from the CUDA kernels (GPU/InamuroCUDA.cu:503-583, 732-1013,     */
/* 1198-1323) and their driver loop doPressurePoisson (:1816-2349).  The CPU */
/* reference has none of them.  Global sums run in this oracle's x-outer     */
/* order (the reference uses atomics, i.e. no fixed order).                  */

void orc_default_poisson_accel(orc_poisson_accel* a)
{
    memset(a, 0, sizeof(*a));
    a->variant = ORC_POISSON_SPLIT;
    a->pressure_relax_scale = 1.0;
    a->fixed_point_relax = 1.0;
    a->scalar_source_scale = 2.0;  /* GPU/InamuroCUDA.hpp:152-157 */
    a->source_aware_hh_scale = 1.0;
    a->anderson_beta_max = 1.0;
    a->two_grid_strength = 0.5;
}

/* The setters' validation, GPU/InamuroCUDA.cu:2488-2540. */
int orc_set_poisson_accel(orc_sim* s, const orc_poisson_accel* a)
{
    if (a->variant < ORC_POISSON_SPLIT || a->variant > ORC_POISSON_SCALAR) {
        set_err(s, "--poisson must be split, fused, onepass, or scalar");
        return -1;
    }
    if (!(a->pressure_relax_scale > 0.0)) { set_err(s, "pressure relaxation scale must be positive"); return -1; }
    if (!(a->fixed_point_relax > 0.0)) { set_err(s, "Poisson fixed-point relaxation must be positive"); return -1; }
    if (!(a->anderson_beta_max >= 0.0)) { set_err(s, "Poisson Anderson beta max must be non-negative"); return -1; }
    if (!(a->two_grid_strength >= 0.0)) { set_err(s, "Poisson two-grid strength must be non-negative"); return -1; }
    s->accel = *a;
    return 0;
}

static int accel_active(const orc_sim* s)
{
    const orc_poisson_accel* a = &s->accel;
    return a->variant != ORC_POISSON_SPLIT || a->fused_boundary || a->pressure_relax_scale != 1.0 ||
           a->fixed_point_relax != 1.0 || a->source_aware_hh_init || a->anderson_m1 || a->two_grid ||
           a->spatial_diagnostics;
}

/* collisionPressureKernel with the relaxation scale (GPU/InamuroCUDA.cu:
 * 503-537): h <- h - s (h - E p) / tau_h - (E / 3) div, in place. */
static void collision_p_scaled(orc_sim* s, double* hdist, double scale)
{
    const size_t n = s->n;
    const double* rho = A(s, ORC_RHO);
    const double* p = A(s, ORC_P);
    const double* ux = A(s, ORC_U_X);
    const double* vy = A(s, ORC_V_Y);
    const double* wz = A(s, ORC_W_Z);
#pragma omp parallel for num_threads(s->nthreads) schedule(static) if (s->nthreads > 1)
    for (size_t i = 0; i < n; ++i) {
        const double div_u = ux[i] + vy[i] + wz[i];
        const double tauh = 1.0 / rho[i] + 0.5;
        for (int k = 0; k < Q; ++k) {
            const double hequ = EW[k] * p[i];
            double* hk = &hdist[(size_t)k * n + i];
            *hk = *hk - scale * (*hk - hequ) / tauh - (EW[k] / 3.0) * div_u;
        }
    }
}

/* The image of cell (x, y, z) under the pull form of one iteration: the
 * collision of collisionPressureStreamKernel (GPU/InamuroCUDA.cu:539-583)
 * evaluated at the source cell x - c_q with periodic wrap, then the x-slip
 * overwrite of boundaryAndComputePressureKernel (:732-776), which
 * collisionStreamBoundaryPressureKernel (:778-846) applies the same way. */
static void accel_image_cell(const orc_sim* s, const double* h, const double* p, int x, int y, int z,
                             double* img)
{
    const size_t n = s->n;
    const double* rho = A(s, ORC_RHO);
    const double* ux = A(s, ORC_U_X);
    const double* vy = A(s, ORC_V_Y);
    const double* wz = A(s, ORC_W_Z);
    const double scale = s->accel.pressure_relax_scale;
    for (int q = 0; q < Q; ++q) {
        const int xs = (x - EX[q] + s->lx) % s->lx;
        const int ys = (y - EY[q] + s->ly) % s->ly;
        const int zs = (z - EZ[q] + s->lz) % s->lz;
        const size_t src = IX(s, xs, ys, zs);
        const double tauh = 1.0 / rho[src] + 0.5;
        const double hequ = EW[q] * p[src];
        const double div_u = ux[src] + vy[src] + wz[src];
        const double h_old = h[(size_t)q * n + src];
        img[q] = h_old - scale * (h_old - hequ) / tauh - (EW[q] / 3.0) * div_u;
    }
    if (x == 0) {
        for (int r = 0; r < 5; ++r) img[SLIP_POS[r]] = img[SLIP_NEG[r]];
    } else if (x == s->lx - 1) {
        for (int r = 0; r < 5; ++r) img[SLIP_NEG[r]] = img[SLIP_POS[r]];
    }
}

/* pressure_bounce_direction, GPU/InamuroCUDA.cu:848-866: at the x walls a
 * neighbour beyond the wall is replaced by its mirror image. */
static int pressure_bounce_direction(int q, int x, int lx)
{
    if (x == 0) {
        for (int r = 0; r < 5; ++r)
            if (q == SLIP_POS[r]) return SLIP_NEG[r];
    } else if (x == lx - 1) {
        for (int r = 0; r < 5; ++r)
            if (q == SLIP_NEG[r]) return SLIP_POS[r];
    }
    return q;
}

/* Work arrays of one accelerated solve. */
typedef struct accel_work {
    double* hn;        /* next h (15 n)                              */
    double* pn;        /* next p (n)                                 */
    double* prev_res;  /* Anderson: previous residual (15 n)         */
    double* prev_img;  /* Anderson: previous image (15 n)            */
    double* blocks;    /* 4^3 block sums of p - p_prev               */
    int has_hist;
} accel_work;

/* One iteration of the selected variant: h and p of the simulation advance
 * in place (GPU/InamuroCUDA.cu:2216-2323). */
static void accel_iteration(orc_sim* s, accel_work* w)
{
    const int lx = s->lx, ly = s->ly, lz = s->lz;
    const size_t n = s->n;
    const orc_poisson_accel* a = &s->accel;
    double* h = s->h;
    double* p = A(s, ORC_P);
    const double* rho = A(s, ORC_RHO);
    const double* ux = A(s, ORC_U_X);
    const double* vy = A(s, ORC_V_Y);
    const double* wz = A(s, ORC_W_Z);

    if (a->variant == ORC_POISSON_SCALAR) {
        /* scalarPressureJacobiKernel (:868-908). */
        const double scale = a->scalar_source_scale;
#pragma omp parallel for num_threads(s->nthreads) schedule(static) if (s->nthreads > 1)
        for (int x = 0; x < lx; ++x)
            for (int y = 0; y < ly; ++y)
                for (int z = 0; z < lz; ++z) {
                    double neighbor_sum = 0.0;
                    for (int q = 1; q < Q; ++q) {
                        const int qb = pressure_bounce_direction(q, x, lx);
                        const int xs = (x - EX[qb] + lx) % lx;
                        const int ys = (y - EY[qb] + ly) % ly;
                        const int zs = (z - EZ[qb] + lz) % lz;
                        neighbor_sum += p[IX(s, xs, ys, zs)];
                    }
                    const size_t i = IX(s, x, y, z);
                    const double div_u = ux[i] + vy[i] + wz[i];
                    const double rhs = scale * rho[i] * div_u;
                    w->pn[i] = (neighbor_sum - 5.0 * rhs) / 14.0;
                }
        memcpy(p, w->pn, sizeof(double) * n);
        return;
    }

    if (a->variant == ORC_POISSON_SPLIT) {
        /* collisionPressureKernel, streamKernel, then slipBounceBackKernel +
         * computePressureKernel or boundaryAndComputePressureKernel: the same
         * values either way (:2296-2322). */
        collision_p_scaled(s, h, a->pressure_relax_scale);
        stream(s, h);
        slip_bounce_back(s, h);
        sum_h_into(s, h, p);
        return;
    }

    const int onepass = a->variant == ORC_POISSON_ONEPASS;
    const double omega = a->fixed_point_relax;
#pragma omp parallel for num_threads(s->nthreads) schedule(static) if (s->nthreads > 1)
    for (int x = 0; x < lx; ++x)
        for (int y = 0; y < ly; ++y)
            for (int z = 0; z < lz; ++z) {
                double img[Q];
                accel_image_cell(s, h, p, x, y, z, img);
                const size_t i = IX(s, x, y, z);
                double p_sum = 0.0;
                for (int q = 0; q < Q; ++q) {
                    double v = img[q];
                    if (onepass) {
                        /* collisionStreamBoundaryPressureKernel (:836-842). */
                        const double h_old = h[(size_t)q * n + i];
                        v = h_old + omega * (v - h_old);
                    }
                    w->hn[(size_t)q * n + i] = v;
                    p_sum += v;
                }
                w->pn[i] = p_sum;
            }

    if (onepass && a->anderson_m1) {
        const size_t nv = (size_t)Q * n;
        if (!w->has_hist) {
            /* andersonStoreHistoryKernel (:930-942). */
            for (size_t k = 0; k < nv; ++k) {
                w->prev_res[k] = w->hn[k] - h[k];
                w->prev_img[k] = w->hn[k];
            }
            w->has_hist = 1;
        } else {
            /* andersonM1DotKernel (:944-981) and the host's coefficient
             * (:2254-2266). */
            double num = 0.0, den = 0.0;
            for (size_t k = 0; k < nv; ++k) {
                const double residual = w->hn[k] - h[k];
                const double delta = residual - w->prev_res[k];
                num += w->prev_res[k] * delta;
                den += delta * delta;
            }
            double alpha = 1.0;
            if (den > 1.0e-300) alpha = -num / den;
            if (!isfinite(alpha)) alpha = 1.0;
            const double lo = -a->anderson_beta_max;
            const double hi = 1.0 + a->anderson_beta_max;
            if (alpha > hi) alpha = hi;
            if (alpha < lo) alpha = lo;
            s->diag.anderson_alpha = alpha;
            /* andersonM1ApplyKernel (:983-1013). */
#pragma omp parallel for num_threads(s->nthreads) schedule(static) if (s->nthreads > 1)
            for (size_t i = 0; i < n; ++i) {
                double p_sum = 0.0;
                for (int q = 0; q < Q; ++q) {
                    const size_t k = (size_t)q * n + i;
                    const double current_image = w->hn[k];
                    const double accelerated = (1.0 - alpha) * w->prev_img[k] + alpha * current_image;
                    w->prev_res[k] = current_image - h[k];
                    w->prev_img[k] = current_image;
                    w->hn[k] = accelerated;
                    p_sum += accelerated;
                }
                w->pn[i] = p_sum;
            }
        }
    }
    memcpy(h, w->hn, sizeof(double) * (size_t)Q * n);
    memcpy(p, w->pn, sizeof(double) * n);
}

#define ACCEL_BLOCK 4 /* spatial_block_size, GPU/InamuroCUDA.cu:2061 */

/* check_residual of the accelerated loop (GPU/InamuroCUDA.cu:1960-2160)
 * without the dual residual, which the reference refuses with any
 * accelerator.  Returns 1 when converged, 0 when not, -1 on error. */
static int accel_check(orc_sim* s, accel_work* w, double gauge_target, uint64_t iter)
{
    const int lx = s->lx, ly = s->ly, lz = s->lz;
    const size_t n = s->n;
    const orc_poisson_accel* a = &s->accel;
    orc_poisson_diag* d = &s->diag;
    const double cell_count = (double)lx * ly * lz;
    double* p = A(s, ORC_P);
    double* pp = A(s, ORC_P_PREV);
    double* h = s->h;

    /* pressure gauge (:1963-1970) on p and h */
    double cur = 0.0;
    for (size_t i = 0; i < n; ++i) cur += p[i];
    const double shift = cur / cell_count - gauge_target;
    if (!isfinite(shift)) { set_err(s, "pressure gauge is NaN/Inf"); return -1; }
    if (fabs(shift) > d->pressure_gauge_max_shift) d->pressure_gauge_max_shift = fabs(shift);
    for (size_t i = 0; i < n; ++i) {
        p[i] -= shift;
        for (int k = 0; k < Q; ++k) h[(size_t)k * n + i] -= EW[k] * shift;
    }

    /* relative change, with 4^3 block sums of p - p_prev when the two-grid
     * correction or the spatial diagnostics need them
     * (pressureErrorSpatialKernel, :1198-1248) */
    const int need_blocks = a->two_grid || a->spatial_diagnostics;
    const int bx_count = (lx + ACCEL_BLOCK - 1) / ACCEL_BLOCK;
    const int by_count = (ly + ACCEL_BLOCK - 1) / ACCEL_BLOCK;
    const int bz_count = (lz + ACCEL_BLOCK - 1) / ACCEL_BLOCK;
    const int block_count = bx_count * by_count * bz_count;
    if (need_blocks) memset(w->blocks, 0, sizeof(double) * (size_t)block_count);
    double err1 = 0.0, err2 = 0.0;
    for (int x = 0; x < lx; ++x)
        for (int y = 0; y < ly; ++y)
            for (int z = 0; z < lz; ++z) {
                const size_t i = IX(s, x, y, z);
                const double current = p[i];
                const double delta = current - pp[i];
                err1 += fabs(delta);
                err2 += fabs(current);
                if (need_blocks) {
                    const int b = ((z / ACCEL_BLOCK) * by_count + y / ACCEL_BLOCK) * bx_count + x / ACCEL_BLOCK;
                    w->blocks[b] += delta;
                }
                pp[i] = current;
            }
    d->pressure_l1_delta = err1;
    d->pressure_l1_norm = err2;
    if (!isfinite(err1) || !isfinite(err2)) {
        set_err(s, "Poisson pressure/residual contains NaN/Inf");
        return -1;
    }
    const double rel = (err2 > 0.0) ? (err1 / err2) : ((err1 == 0.0) ? 0.0 : INFINITY);
    if (!isfinite(rel)) { set_err(s, "Poisson residual is NaN/Inf"); return -1; }
    d->relative_error = rel;
    d->iteration_count = iter;
    const int converged = rel < s->tol;
    if (need_blocks) {
        d->block_size = ACCEL_BLOCK;
        d->block_count = block_count;
    }
    if (a->spatial_diagnostics) {
        /* blockAbsSumKernel (:1250-1275) and the fraction (:2107-2110) */
        double block_error = 0.0;
        for (int b = 0; b < block_count; ++b) block_error += fabs(w->blocks[b]);
        d->block_low_frequency_fraction = (err1 > 0.0) ? (block_error / err1) : 0.0;
    }
    if (a->two_grid && !converged && err1 > 0.0) {
        /* coarsePressureCorrectionKernel (:1277-1323) */
        const double strength = a->two_grid_strength;
#pragma omp parallel for num_threads(s->nthreads) schedule(static) if (s->nthreads > 1)
        for (int x = 0; x < lx; ++x)
            for (int y = 0; y < ly; ++y)
                for (int z = 0; z < lz; ++z) {
                    const int bx = x / ACCEL_BLOCK, by = y / ACCEL_BLOCK, bz = z / ACCEL_BLOCK;
                    const int b = (bz * by_count + by) * bx_count + bx;
                    const int x0 = bx * ACCEL_BLOCK, y0 = by * ACCEL_BLOCK, z0 = bz * ACCEL_BLOCK;
                    const int x1 = x0 + ACCEL_BLOCK < lx ? x0 + ACCEL_BLOCK : lx;
                    const int y1 = y0 + ACCEL_BLOCK < ly ? y0 + ACCEL_BLOCK : ly;
                    const int z1 = z0 + ACCEL_BLOCK < lz ? z0 + ACCEL_BLOCK : lz;
                    int block_cells = (x1 - x0) * (y1 - y0) * (z1 - z0);
                    if (block_cells < 1) block_cells = 1;
                    const double correction = strength * w->blocks[b] / (double)block_cells;
                    const size_t i = IX(s, x, y, z);
                    const double corrected_p = p[i] + correction;
                    p[i] = corrected_p;
                    pp[i] = corrected_p;
                    for (int q = 0; q < Q; ++q) h[(size_t)q * n + i] += EW[q] * correction;
                }
    }
    return converged;
}

/* doPressurePoisson with accelerators (GPU/InamuroCUDA.cu:1816-2349). */
static int solve_pressure_poisson_accel(orc_sim* s)
{
    const orc_poisson_accel* a = &s->accel;
    if (s->dual) {
        /* GPU/InamuroCUDA.cu:1846-1865 */
        if (a->variant != ORC_POISSON_SPLIT || a->fused_boundary || a->source_aware_hh_init ||
            a->anderson_m1 || a->two_grid || a->pressure_relax_scale != 1.0 || a->fixed_point_relax != 1.0)
            set_err(s, "fixed-point residual requires frozen split+split Poisson with "
                       "unit relaxation and all accelerators disabled");
        else
            set_err(s, "dual-residual exit requires check_interval=100, tolerance=1e-3, "
                       "poisson_iteration_limit=0, deterministic reductions, no spatial diagnostics");
        return -1;
    }
    if (s->z_walls) { set_err(s, "Poisson accelerators support the reference's periodic z only"); return -1; }
    const size_t n = s->n;
    orc_poisson_diag* d = &s->diag;
    double gauge_target = 0.0;
    if (poisson_prologue(s, &gauge_target) != 0) return -1;

    accel_work w;
    memset(&w, 0, sizeof(w));
    const int bcount = ((s->lx + ACCEL_BLOCK - 1) / ACCEL_BLOCK) * ((s->ly + ACCEL_BLOCK - 1) / ACCEL_BLOCK) *
                       ((s->lz + ACCEL_BLOCK - 1) / ACCEL_BLOCK);
    w.hn = (double*)malloc(sizeof(double) * (size_t)Q * n);
    w.pn = (double*)malloc(sizeof(double) * n);
    w.blocks = (double*)malloc(sizeof(double) * (size_t)bcount);
    int ok = w.hn && w.pn && w.blocks;
    if (a->anderson_m1 && a->variant == ORC_POISSON_ONEPASS) {
        w.prev_res = (double*)malloc(sizeof(double) * (size_t)Q * n);
        w.prev_img = (double*)malloc(sizeof(double) * (size_t)Q * n);
        ok = ok && w.prev_res && w.prev_img;
    }
    int rc = 0;
    if (!ok) { set_err(s, "out of memory"); rc = -1; goto done; }

    if (a->source_aware_hh_init && a->variant != ORC_POISSON_SCALAR) {
        /* sourceAwareHHInitKernel (:905-928) */
        const double* rho = A(s, ORC_RHO);
        const double* ux = A(s, ORC_U_X);
        const double* vy = A(s, ORC_V_Y);
        const double* wz = A(s, ORC_W_Z);
        const double* p = A(s, ORC_P);
        const double source_scale = a->source_aware_hh_scale;
        for (size_t i = 0; i < n; ++i) {
            const double tauh = 1.0 / rho[i] + 0.5;
            const double div_u = ux[i] + vy[i] + wz[i];
            const double loc_p = p[i];
            for (int q = 0; q < Q; ++q) {
                const double hequ = EW[q] * loc_p;
                const double source = (EW[q] / 3.0) * div_u;
                s->h[(size_t)q * n + i] = hequ - source_scale * tauh * source;
            }
        }
    }

    const uint64_t forced = s->schedule_pos < s->schedule_len ? s->schedule[s->schedule_pos] : 0;
    uint64_t iter = 0;
    for (;;) {
        if (iter == UINT64_MAX) { set_err(s, "Poisson iteration counter overflow"); rc = -2; goto done; }
        ++iter;
        accel_iteration(s, &w);
        if (iter % s->check_interval == 0) {
            const int conv = accel_check(s, &w, gauge_target, iter);
            if (conv < 0) { rc = -1; goto done; }
            if (forced) {
                if (iter >= forced) break;
                continue;
            }
            if (conv) { d->converged = 1; break; }
            if (s->iter_cap != 0 && iter >= s->iter_cap) {
                d->hit_max_iterations = 1;
                d->fallback_count = 1;
                break;
            }
        }
    }
    s->last_iteration_count = iter;
    d->iteration_count = iter;
    if (forced) ++s->schedule_pos;
done:
    free(w.hn); free(w.pn); free(w.blocks); free(w.prev_res); free(w.prev_img);
    return rc;
}

int orc_poisson_accel_iterations(orc_sim* s, int k)
{
    const size_t n = s->n;
    accel_work w;
    memset(&w, 0, sizeof(w));
    w.hn = (double*)malloc(sizeof(double) * (size_t)Q * n);
    w.pn = (double*)malloc(sizeof(double) * n);
    w.prev_res = (double*)malloc(sizeof(double) * (size_t)Q * n);
    w.prev_img = (double*)malloc(sizeof(double) * (size_t)Q * n);
    int rc = 0;
    if (!w.hn || !w.pn || !w.prev_res || !w.prev_img) {
        set_err(s, "out of memory");
        rc = -1;
    } else {
        s->diag.anderson_alpha = NAN;
        for (int it = 0; it < k; ++it) accel_iteration(s, &w);
    }
    free(w.hn); free(w.pn); free(w.prev_res); free(w.prev_img);
    return rc;
}

/* Extension: no-slip z walls and gravity. */
int orc_set_z_walls(orc_sim* s, int enabled)
{
    if (enabled && s->lz < 2) { set_err(s, "z walls need lz >= 2"); return -1; }
    s->z_walls = enabled ? 1 : 0;
    build_z_neighbours(s);
    return 0;
}

void orc_set_gravity(orc_sim* s, double gz) { s->gz = gz; }

/* Extension: freeze the iteration count of the next `count` solves. */
int orc_set_schedule(orc_sim* s, const uint64_t* iters, int count)
{
    free(s->schedule);
    s->schedule = NULL;
    s->schedule_len = s->schedule_pos = 0;
    if (count <= 0) return 0;
    s->schedule = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)count);
    if (!s->schedule) { set_err(s, "out of memory"); return -1; }
    for (int i = 0; i < count; ++i) {
        if (iters[i] == 0 || iters[i] % s->check_interval != 0) {
            set_err(s, "scheduled iteration counts must be positive multiples of the check interval");
            return -1;
        }
        s->schedule[i] = iters[i];
    }
    s->schedule_len = count;
    return 0;
}

/* Inamuro::correct_uvw, CPU/Inamuro.cpp:794-823 (reciprocal-multiply form). */
static void correct_uvw(orc_sim* s)
{
    const int lx = s->lx, ly = s->ly, lz = s->lz;
    const double* p = A(s, ORC_P);
    const double* rho = A(s, ORC_RHO);
    double* u = A(s, ORC_U);
    double* v = A(s, ORC_V);
    double* w = A(s, ORC_W);
#pragma omp parallel for num_threads(s->nthreads) schedule(static) if (s->nthreads > 1)
    for (int x = 0; x < lx; ++x)
        for (int y = 0; y < ly; ++y)
            for (int z = 0; z < lz; ++z) {
                const int x_e = (x == lx - 1) ? lx - 2 : x + 1;
                const int x_w = (x == 0) ? 1 : x - 1;
                const int y_n = (y == ly - 1) ? 0 : y + 1;
                const int y_s = (y == 0) ? ly - 1 : y - 1;
                const int z_n = (z == lz - 1) ? (s->z_walls ? lz - 1 : 0) : z + 1;
                const int z_s = (z == 0) ? (s->z_walls ? 0 : lz - 1) : z - 1;
                const size_t i = IX(s, x, y, z);
                const double inv_2rho = 1.0 / (2.0 * rho[i]);
                u[i] -= (p[IX(s, x_e, y, z)] - p[IX(s, x_w, y, z)]) * inv_2rho;
                v[i] -= (p[IX(s, x, y_n, z)] - p[IX(s, x, y_s, z)]) * inv_2rho;
                w[i] -= (p[IX(s, x, y, z_n)] - p[IX(s, x, y, z_s)]) * inv_2rho;
            }
}

/* Inamuro::update_hh, CPU/Inamuro.cpp:825-840. */
static void update_hh(orc_sim* s)
{
    const size_t n = s->n;
    const double* p = A(s, ORC_P);
    for (int k = 0; k < Q; ++k)
        for (size_t i = 0; i < n; ++i) s->h[(size_t)k * n + i] = p[i] * EW[k];
}

int orc_stage(orc_sim* s, int stage)
{
    switch (stage) {
    case ORC_STAGE_GRADIENTS: gradients(s); return 0;
    case ORC_STAGE_COLLISION: collision(s); return 0;
    case ORC_STAGE_STREAM_FG: stream(s, s->f); stream(s, s->g); return 0;
    case ORC_STAGE_BC_FG: slip_bounce_back(s, s->f); slip_bounce_back(s, s->g); return 0;
    case ORC_STAGE_MACRO: get_macro(s); return 0;
    case ORC_STAGE_POISSON: return solve_pressure_poisson(s);
    case ORC_STAGE_CORRECT: correct_uvw(s); return 0;
    case ORC_STAGE_UPDATE_HH: update_hh(s); return 0;
    case ORC_STAGE_COLLISION_P: collision_p_on(s, s->h); return 0;
    case ORC_STAGE_STREAM_H: stream(s, s->h); return 0;
    case ORC_STAGE_BC_H: slip_bounce_back(s, s->h); return 0;
    case ORC_STAGE_GETP: sum_h_into(s, s->h, A(s, ORC_P)); return 0;
    default: set_err(s, "unknown stage"); return -1;
    }
}

/* Inamuro::performTimeStep, CPU/Inamuro.cpp:347-363. */
int orc_step(orc_sim* s)
{
    collision(s);
    stream(s, s->f);
    stream(s, s->g);
    slip_bounce_back(s, s->f);
    slip_bounce_back(s, s->g);
    get_macro(s);
    const int rc = solve_pressure_poisson(s);
    if (rc != 0) return rc;
    correct_uvw(s);
    update_hh(s);
    return 0;
}

void orc_poisson_diag_get(const orc_sim* s, orc_poisson_diag* out) { *out = s->diag; }


/* ================================================================== */
/* Discrete adjoint of the frozen-schedule step (extension)           */
/* ================================================================== */
/*
 * The reference has no adjoint.  Below is the reverse mode of the step above
 * with every Poisson solve's iteration count frozen (orc_set_schedule): each
 * primitive is transposed by hand -- correct_uvw, every executed Poisson
 * iteration with its gauge shifts and the source-mean projection, getMacro
 * with the clamped density map, slip and stream, the BGK collision with its
 * first-step equilibrium start, and the gradient / Laplacian stencils.  The
 * forward pass keeps each step's start state; the reverse pass recomputes one
 * step at a time from it, recording that step's Poisson iterates, so memory
 * holds one step's iterates (small grids: this is a checker).
 *
 * Objective J = sum (phi_T - target)^2, or the centroid height of
 * orc_set_objective, after `steps` steps from phi0, with
 * rho0 = rho_G + (phi0 - fei_min)(rho_L - rho_G)/(fei_max - fei_min) (the
 * density rule of place_droplet, GPU/Inamuro.cpp:521-527), zero u, v, w, p,
 * h = 0 and the first-step flag set.
 */

typedef struct {
    double* m[6];     /* step-start fei, rho, u, v, w, p */
    double* p_prev;   /* step-start pressure_prev (only feeds the residual) */
    double* f;        /* step-start streamed f, g (unused on the first step) */
    double* g;
    int first;
} adj_ckpt;

static void load_canonical(orc_sim* s, int id_scratch, const double* canon, double* internal)
{
    double* save = A(s, id_scratch);
    A(s, id_scratch) = internal;
    orc_set(s, id_scratch, canon);
    A(s, id_scratch) = save;
}

static void store_canonical(const orc_sim* s, int id_scratch, const double* internal, double* canon)
{
    orc_sim* w = (orc_sim*)s;
    double* save = A(w, id_scratch);
    A(w, id_scratch) = (double*)internal;
    orc_get(w, id_scratch, canon);
    A(w, id_scratch) = save;
}

static void adj_reset(orc_sim* s, const double* phi0_canon)
{
    const orc_params* P = &s->P;
    const size_t n = s->n;
    orc_set(s, ORC_FEI, phi0_canon);
    for (size_t i = 0; i < n; ++i) {
        A(s, ORC_RHO)[i] =
            P->rho_G + (A(s, ORC_FEI)[i] - P->fei_min) * (P->rho_L - P->rho_G) / (P->fei_max - P->fei_min);
        A(s, ORC_U)[i] = A(s, ORC_V)[i] = A(s, ORC_W)[i] = 0.0;
        A(s, ORC_P)[i] = A(s, ORC_P_PREV)[i] = 0.0;
    }
    memset(s->f, 0, sizeof(double) * (size_t)Q * n);
    memset(s->g, 0, sizeof(double) * (size_t)Q * n);
    memset(s->h, 0, sizeof(double) * (size_t)Q * n);
    s->first_step = 1;
    s->last_iteration_count = 0;
    s->pressure_solve_call_count = 0;
}

int orc_set_objective(orc_sim* s, int kind)
{
    if (kind != ORC_OBJ_MISFIT && kind != ORC_OBJ_CENTROID_Z) {
        set_err(s, "unknown objective");
        return -1;
    }
    s->objective = kind;
    return 0;
}

/* The objective at the current fields; with seed != NULL also dJ/dfei (internal
 * layout).  The centroid's z of internal index i is i mod lz. */
static double adj_value(const orc_sim* s, const double* target, double* seed)
{
    const double* fei = A(s, ORC_FEI);
    if (s->objective == ORC_OBJ_CENTROID_Z) {
        double W = 0.0, Z = 0.0;
        for (size_t i = 0; i < s->n; ++i) {
            const double w = fei[i] - s->P.fei_G;
            W += w;
            Z += (double)(i % (size_t)s->lz) * w;
        }
        const double zc = Z / W;
        if (seed)
            for (size_t i = 0; i < s->n; ++i) seed[i] = ((double)(i % (size_t)s->lz) - zc) / W;
        return zc;
    }
    double J = 0.0;
    for (size_t i = 0; i < s->n; ++i) {
        const double d = fei[i] - target[i];
        J += d * d;
        if (seed) seed[i] = 2.0 * d;
    }
    return J;
}

/* The adjoint transposes the strict split map only. */
#define ADJ_ACCEL_MSG "the adjoint differentiates the strict split Poisson map; turn the Poisson accelerators off"

int orc_objective(orc_sim* s, const double* phi0, const double* target, int steps,
                  const uint64_t* iters, double* J)
{
    if (accel_active(s)) { set_err(s, ADJ_ACCEL_MSG); return -1; }
    double* t = (double*)calloc(s->n, sizeof(double));
    if (!t) { set_err(s, "out of memory"); return -1; }
    if (s->objective == ORC_OBJ_MISFIT) load_canonical(s, ORC_P_PREV, target, t);
    adj_reset(s, phi0);
    int rc = orc_set_schedule(s, iters, steps);
    for (int k = 0; rc == 0 && k < steps; ++k) rc = orc_step(s);
    if (rc == 0) *J = adj_value(s, t, NULL);
    orc_set_schedule(s, NULL, 0);
    free(t);
    return rc;
}

/* Transpose of stream + slipBounceBack (CPU/Inamuro.cpp:252-292, 589-623),
 * with the z bounce-back of the extension, on a [q][n] adjoint array, in
 * place. */
static void adj_stream_slip(orc_sim* s, double* lam)
{
    const size_t n = s->n;
    const int lx = s->lx, ly = s->ly, lz = s->lz;
    /* slip wrote dst <- src: src += dst, dst = 0 (right wall last, so first) */
    for (int y = 0; y < ly; ++y)
        for (int z = 0; z < lz; ++z)
            for (int r = 0; r < 5; ++r) {
                const size_t ir = IX(s, lx - 1, y, z);
                lam[(size_t)SLIP_POS[r] * n + ir] += lam[(size_t)SLIP_NEG[r] * n + ir];
                lam[(size_t)SLIP_NEG[r] * n + ir] = 0.0;
            }
    for (int y = 0; y < ly; ++y)
        for (int z = 0; z < lz; ++z)
            for (int r = 0; r < 5; ++r) {
                const size_t il = IX(s, 0, y, z);
                lam[(size_t)SLIP_NEG[r] * n + il] += lam[(size_t)SLIP_POS[r] * n + il];
                lam[(size_t)SLIP_POS[r] * n + il] = 0.0;
            }
    /* stream wrote out_q(x) = in_q(x - c_q): in_q(y) = out_q(y + c_q); with
     * the z walls, in_q(y) whose y + c_q lies beyond a wall came back as
     * out_{opp q}(y) */
    for (int q = 0; q < Q; ++q) {
        const double* src = lam + (size_t)q * n;
        double* dst = s->tmp + (size_t)q * n;
        for (int x = 0; x < lx; ++x) {
            const int xs = (x + EX[q] + lx) % lx;
            for (int y = 0; y < ly; ++y) {
                const int ys = (y + EY[q] + ly) % ly;
                for (int z = 0; z < lz; ++z) {
                    const int zt = z + EZ[q];
                    if (s->z_walls && (zt < 0 || zt >= lz))
                        dst[IX(s, x, y, z)] = lam[(size_t)OPP[q] * n + IX(s, x, y, z)];
                    else
                        dst[IX(s, x, y, z)] = src[IX(s, xs, ys, (zt + lz) % lz)];
                }
            }
        }
    }
    memcpy(lam, s->tmp, sizeof(double) * (size_t)Q * n);
}

/* Transpose of firstord (gradient adjoints gx, gy, gz) and secondord
 * (Laplacian adjoint lp), CPU/Inamuro.hpp:213-314; any may be NULL.
 * Accumulates into out. */
static void adj_stencil(orc_sim* s, const double* gx, const double* gy, const double* gz,
                        const double* lp, double* out)
{
    for (int x = 0; x < s->lx; ++x)
        for (int y = 0; y < s->ly; ++y)
            for (int z = 0; z < s->lz; ++z) {
                const size_t i = IX(s, x, y, z);
                const double ax = gx ? gx[i] / 10.0 : 0.0;
                const double ay = gy ? gy[i] / 10.0 : 0.0;
                const double az = gz ? gz[i] / 10.0 : 0.0;
                const double al = lp ? lp[i] / 5.0 : 0.0;
                for (int k = 1; k < Q; ++k) {
                    const size_t j = IX(s, s->xn[x * 3 + EX[k] + 1], s->yn[y * 3 + EY[k] + 1],
                                        s->zn[z * 3 + EZ[k] + 1]);
                    out[j] += UC[k] * ax + VC[k] * ay + WC[k] * az + al;
                }
                out[i] -= 14.0 * al;
            }
}

/* d rho / d fei of the clamped sine map (CPU/Inamuro.cpp:313-327); zero on
 * the clamped branches, matching their one-sided derivative. */
static double drho_dfei(const orc_params* P, double fei)
{
    if (fei <= P->fei_G || fei >= P->fei_L) return 0.0;
    const double k = M_PI / (P->fei_L - P->fei_G);
    const double arg = (fei - (P->fei_L + P->fei_G) / 2.0) * k;
    return (P->rho_L - P->rho_G) / 2.0 * cos(arg) * k;
}

/* Gradient-array slots of the collision adjoint. */
enum { GA_UX, GA_UY, GA_UZ, GA_VX, GA_VY, GA_VZ, GA_WX, GA_WY, GA_WZ,
       GA_RX, GA_RY, GA_RZ, GA_FX, GA_FY, GA_FZ, GA_ULAP, GA_VLAP, GA_WLAP, GA_FLAP, GA_NUM };

int orc_adjoint(orc_sim* s, const double* phi0, const double* target, int steps,
                const uint64_t* iters, double* J, double* grad_phi0, double* grad_kg)
{
    if (accel_active(s)) { set_err(s, ADJ_ACCEL_MSG); return -1; }
    const orc_params* P = &s->P;
    const size_t n = s->n;
    const size_t qn = (size_t)Q * n;
    const int lx = s->lx, ly = s->ly, lz = s->lz;
    int rc = -1;
    uint64_t max_it = 0;
    for (int k = 0; k < steps; ++k) max_it = iters[k] > max_it ? iters[k] : max_it;

    adj_ckpt* ck = (adj_ckpt*)calloc((size_t)(steps > 0 ? steps : 1), sizeof(adj_ckpt));
    double* tgt = (double*)calloc(n, sizeof(double));
    double* rec_h = (double*)malloc(sizeof(double) * qn * (size_t)(max_it ? max_it : 1));
    double* rec_p = (double*)malloc(sizeof(double) * n * (size_t)(max_it ? max_it : 1));
    double* fei_new = (double*)malloc(sizeof(double) * n);
    double* rho_new = (double*)malloc(sizeof(double) * n);
    double* tau = (double*)malloc(sizeof(double) * n);
    double* pfin = (double*)malloc(sizeof(double) * n);
    double* lf = (double*)calloc(qn, sizeof(double));   /* adjoint of f (then of collision output) */
    double* lg = (double*)calloc(qn, sizeof(double));
    double* cf = (double*)calloc(qn, sizeof(double));   /* adjoint of the streamed f, g a step reads */
    double* cg = (double*)calloc(qn, sizeof(double));
    double* lh = (double*)calloc(qn, sizeof(double));
    double* lp = (double*)calloc(n, sizeof(double));
    double* ltau = (double*)calloc(n, sizeof(double));
    double* lS = (double*)calloc(n, sizeof(double));
    double* la[6] = {0, 0, 0, 0, 0, 0};               /* adjoints of fei, rho, u, v, w, p */
    double* ga[GA_NUM];
    double kg_bar = 0.0;
    int ok = ck && tgt && rec_h && rec_p && fei_new && rho_new && tau && pfin && lf && lg && cf && cg && lh &&
             lp && ltau && lS;
    for (int k = 0; k < 6; ++k) { la[k] = (double*)calloc(n, sizeof(double)); ok = ok && la[k]; }
    for (int k = 0; k < GA_NUM; ++k) { ga[k] = (double*)calloc(n, sizeof(double)); ok = ok && ga[k]; }
    for (int k = 0; ok && k < steps; ++k) {
        for (int j = 0; j < 6; ++j) { ck[k].m[j] = (double*)malloc(sizeof(double) * n); ok = ok && ck[k].m[j]; }
        ck[k].f = (double*)malloc(sizeof(double) * qn);
        ck[k].g = (double*)malloc(sizeof(double) * qn);
        ck[k].p_prev = (double*)malloc(sizeof(double) * n);
        ok = ok && ck[k].f && ck[k].g && ck[k].p_prev;
    }
    if (!ok) { set_err(s, "out of memory"); goto done; }
    if (orc_set_schedule(s, iters, steps) != 0) goto done;

    /* ---- forward, keeping every step's start state ---- */
    if (s->objective == ORC_OBJ_MISFIT) load_canonical(s, ORC_P_PREV, target, tgt);
    adj_reset(s, phi0);
    for (int k = 0; k < steps; ++k) {
        for (int j = 0; j < 6; ++j) memcpy(ck[k].m[j], A(s, ORC_FEI + j), sizeof(double) * n);
        memcpy(ck[k].p_prev, A(s, ORC_P_PREV), sizeof(double) * n);
        memcpy(ck[k].f, s->f, sizeof(double) * qn);
        memcpy(ck[k].g, s->g, sizeof(double) * qn);
        ck[k].first = s->first_step;
        if (orc_step(s) != 0) goto done;
    }
    *J = adj_value(s, tgt, la[0]);

    /* ---- reverse ---- */
    for (int k = steps - 1; k >= 0; --k) {
        const adj_ckpt* c = &ck[k];
        const uint64_t nit = iters[k];
        /* recompute step k up to the end of its Poisson solve */
        for (int j = 0; j < 6; ++j) memcpy(A(s, ORC_FEI + j), c->m[j], sizeof(double) * n);
        memcpy(A(s, ORC_P_PREV), c->p_prev, sizeof(double) * n);
        memcpy(s->f, c->f, sizeof(double) * qn);
        memcpy(s->g, c->g, sizeof(double) * qn);
        s->first_step = c->first;
        if (c->first) memset(s->h, 0, sizeof(double) * qn);
        else update_hh(s);
        orc_set_schedule(s, &iters[k], 1);
        collision(s);
        stream(s, s->f); stream(s, s->g);
        slip_bounce_back(s, s->f); slip_bounce_back(s, s->g);
        get_macro(s);
        memcpy(fei_new, A(s, ORC_FEI), sizeof(double) * n);
        memcpy(rho_new, A(s, ORC_RHO), sizeof(double) * n);
        for (size_t i = 0; i < n; ++i) tau[i] = 1.0 / rho_new[i] + 0.5;
        if (solve_pressure_poisson_rec(s, rec_h, rec_p) != 0) goto done;
        memcpy(pfin, A(s, ORC_P), sizeof(double) * n);

        /* 1. correct_uvw (CPU/Inamuro.cpp:794-823): u -= (p_e - p_w)/(2 rho') */
        memcpy(lp, la[5], sizeof(double) * n);
        for (int x = 0; x < lx; ++x)
            for (int y = 0; y < ly; ++y)
                for (int z = 0; z < lz; ++z) {
                    const size_t i = IX(s, x, y, z);
                    const size_t ie = IX(s, (x == lx - 1) ? lx - 2 : x + 1, y, z);
                    const size_t iw = IX(s, (x == 0) ? 1 : x - 1, y, z);
                    const size_t in = IX(s, x, (y == ly - 1) ? 0 : y + 1, z);
                    const size_t is = IX(s, x, (y == 0) ? ly - 1 : y - 1, z);
                    const size_t it = IX(s, x, y, (z == lz - 1) ? (s->z_walls ? lz - 1 : 0) : z + 1);
                    const size_t ib = IX(s, x, y, (z == 0) ? (s->z_walls ? 0 : lz - 1) : z - 1);
                    const double inv = 1.0 / (2.0 * rho_new[i]);
                    lp[ie] -= la[2][i] * inv; lp[iw] += la[2][i] * inv;
                    lp[in] -= la[3][i] * inv; lp[is] += la[3][i] * inv;
                    lp[it] -= la[4][i] * inv; lp[ib] += la[4][i] * inv;
                    la[1][i] += (la[2][i] * (pfin[ie] - pfin[iw]) + la[3][i] * (pfin[in] - pfin[is]) +
                                 la[4][i] * (pfin[it] - pfin[ib])) /
                                (2.0 * rho_new[i] * rho_new[i]);
                }

        /* 2. Poisson iterations, last first (CPU/Inamuro.cpp:673-716) */
        memset(lh, 0, sizeof(double) * qn);
        memset(ltau, 0, sizeof(double) * n);
        memset(lS, 0, sizeof(double) * n);
        double lam_target = 0.0;
        for (uint64_t it = nit; it >= 1; --it) {
            if (it % s->check_interval == 0) {
                /* p -= shift, h -= E shift, shift = mean(p) - target */
                double L = 0.0;
                for (size_t i = 0; i < n; ++i) {
                    double t = lp[i];
                    for (int q = 0; q < Q; ++q) t += EW[q] * lh[(size_t)q * n + i];
                    L += t;
                }
                for (size_t i = 0; i < n; ++i) lp[i] -= L / (double)n;
                lam_target += L;
            }
            /* p = sum_q h_q */
            for (int q = 0; q < Q; ++q)
                for (size_t i = 0; i < n; ++i) lh[(size_t)q * n + i] += lp[i];
            adj_stream_slip(s, lh);
            /* h* = h (1 - 1/tau) + E p / tau - E S / 3 at the iterate it-1 */
            const double* hk = rec_h + (it - 1) * qn;
            const double* pk = rec_p + (it - 1) * n;
            for (size_t i = 0; i < n; ++i) {
                const double t = tau[i];
                double np = 0.0, ns = 0.0, nt = 0.0;
                for (int q = 0; q < Q; ++q) {
                    double* l = &lh[(size_t)q * n + i];
                    np += *l * EW[q] / t;
                    ns -= *l * EW[q] / 3.0;
                    nt += *l * (hk[(size_t)q * n + i] - EW[q] * pk[i]) / (t * t);
                    *l = *l * (1.0 - 1.0 / t);
                }
                lp[i] = np;
                lS[i] += ns;
                ltau[i] += nt;
            }
        }
        /* solve entry: h_0 = E p_start (0 on the first step), p_0 = p_start,
         * target = mean(p_start); source projection S = div - mean(div) */
        for (size_t i = 0; i < n; ++i) {
            double b = lp[i] + lam_target / (double)n;
            if (!c->first)
                for (int q = 0; q < Q; ++q) b += EW[q] * lh[(size_t)q * n + i];
            la[5][i] = b;
        }
        double mS = 0.0;
        for (size_t i = 0; i < n; ++i) mS += lS[i];
        mS /= (double)n;
        for (int j = 0; j < GA_NUM; ++j) memset(ga[j], 0, sizeof(double) * n);
        for (size_t i = 0; i < n; ++i) {
            la[1][i] -= ltau[i] / (rho_new[i] * rho_new[i]);
            ga[GA_UX][i] = ga[GA_VY][i] = ga[GA_WZ][i] = lS[i] - mS;
        }

        /* 3. getMacro (CPU/Inamuro.cpp:294-346) plus the streamed f, g the
         *    next step read */
        for (size_t i = 0; i < n; ++i) {
            const double af = la[0][i] + la[1][i] * drho_dfei(P, fei_new[i]);
            for (int q = 0; q < Q; ++q) {
                lf[(size_t)q * n + i] = cf[(size_t)q * n + i] + af;
                lg[(size_t)q * n + i] = cg[(size_t)q * n + i] + UC[q] * la[2][i] + VC[q] * la[3][i] + WC[q] * la[4][i];
            }
        }
        /* 4. slip + stream of f and g */
        adj_stream_slip(s, lf);
        adj_stream_slip(s, lg);

        /* 5. collision (CPU/Inamuro.cpp:134-250) at the step-start fields */
        for (int j = 0; j < 5; ++j) memcpy(A(s, ORC_FEI + j), c->m[j], sizeof(double) * n);
        gradients(s);
        for (int j = 0; j < 5; ++j) memset(la[j], 0, sizeof(double) * n);
        const double tmp1 = (P->mu_L - P->mu_G) / (P->rho_L - P->rho_G);
        const double V = 1.5 * (P->taug - 0.5) * 2.0;
        for (size_t i = 0; i < n; ++i) {
            const double phi = c->m[0][i], rho = c->m[1][i];
            const double u[3] = {c->m[2][i], c->m[3][i], c->m[4][i]};
            const double gp[3] = {A(s, ORC_FEI_X)[i], A(s, ORC_FEI_Y)[i], A(s, ORC_FEI_Z)[i]};
            const double gr[3] = {A(s, ORC_RHO_X)[i], A(s, ORC_RHO_Y)[i], A(s, ORC_RHO_Z)[i]};
            const double Lu[3] = {A(s, ORC_U_LAP)[i], A(s, ORC_V_LAP)[i], A(s, ORC_W_LAP)[i]};
            const double Lphi = A(s, ORC_FEI_LAP)[i];
            const double sum_rho = gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2];
            const double mu = (rho - P->rho_G) * tmp1 + P->mu_G;
            const double dp0 = P->T / ((1.0 - P->b * phi) * (1.0 - P->b * phi)) - 2.0 * P->a * phi;
            double a_phi = 0.0, a_rho = 0.0, a_u[3] = {0, 0, 0}, a_gp[3] = {0, 0, 0}, a_gr[3] = {0, 0, 0};
            double a_Lu[3] = {0, 0, 0}, a_D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            double a_p0 = 0.0, a_sf = 0.0, a_sr = 0.0, a_vel = 0.0;
            for (int q = 0; q < Q; ++q) {
                const double cq[3] = {UC[q], VC[q], WC[q]};
                const size_t iq = (size_t)q * n + i;
                double af = lf[iq], ag = lg[iq];
                /* gravity: g_q += 3 E_q c_qz gz (1 - rho_G/rho) */
                if (EZ[q] != 0) a_rho += ag * 3.0 * EW[q] * WC[q] * s->gz * P->rho_G / (rho * rho);
                if (!c->first) {
                    cf[iq] = af * (1.0 - 1.0 / P->tauf);
                    cg[iq] = ag * (1.0 - 1.0 / P->taug);
                    const double cl = cq[0] * Lu[0] + cq[1] * Lu[1] + cq[2] * Lu[2];
                    for (int a = 0; a < 3; ++a) a_Lu[a] += ag * 3.0 * EW[q] * mu / rho * cq[a];
                    a_rho += ag * 3.0 * EW[q] * cl * (tmp1 / rho - mu / (rho * rho));
                    af /= P->tauf;
                    ag /= P->taug;
                } else {
                    cf[iq] = cg[iq] = 0.0;
                }
                const double un = cq[0] * u[0] + cq[1] * u[1] + cq[2] * u[2];
                const double fe = cq[0] * gp[0] + cq[1] * gp[1] + cq[2] * gp[2];
                const double re = cq[0] * gr[0] + cq[1] * gr[1] + cq[2] * gr[2];
                const double cs = cq[0] * cq[0] + cq[1] * cq[1] + cq[2] * cq[2];
                const double Gr = 4.5 * re * re - 1.5 * sum_rho * cs;
                /* feq = HW fei + FW (p0 - k_f fei lap - k_f |grad fei|^2/6)
                 *       + 3 E fei (c.u) + E k_f (4.5 (c.grad fei)^2 - 1.5 |grad fei|^2 |c|^2) */
                a_phi += af * (HW[q] + 3.0 * EW[q] * un);
                a_p0 += af * FW[q];
                a_sf += af * EW[q] * P->k_f * (-1.5 * cs);
                for (int a = 0; a < 3; ++a) {
                    a_gp[a] += af * EW[q] * P->k_f * 9.0 * fe * cq[a];
                    a_u[a] += af * 3.0 * EW[q] * phi * cq[a];
                }
                /* geq = E velPart + E k_g/rho Grho - 2/3 FW k_g/rho |grad rho|^2 */
                const double vb = ag * EW[q];
                a_vel += vb;
                for (int a = 0; a < 3; ++a) {
                    a_u[a] += vb * (3.0 + 9.0 * un) * cq[a];
                    for (int b = 0; b < 3; ++b) a_D[a * 3 + b] += vb * V * cq[a] * cq[b];
                }
                const double gb = vb * P->k_g / rho;
                for (int a = 0; a < 3; ++a) a_gr[a] += gb * 9.0 * re * cq[a];
                a_sr += gb * (-1.5 * cs) - ag * 2.0 / 3.0 * FW[q] * P->k_g / rho;
                a_rho += -vb * P->k_g * Gr / (rho * rho) + ag * 2.0 / 3.0 * FW[q] * P->k_g * sum_rho / (rho * rho);
                kg_bar += vb * Gr / rho - ag * 2.0 / 3.0 * FW[q] * sum_rho / rho;
            }
            for (int a = 0; a < 3; ++a) a_u[a] -= 3.0 * u[a] * a_vel;
            a_phi += a_p0 * (dp0 - P->k_f * Lphi);
            a_sf += a_p0 * (-P->k_f / 6.0);
            for (int a = 0; a < 3; ++a) {
                a_gp[a] += a_sf * 2.0 * gp[a];
                a_gr[a] += a_sr * 2.0 * gr[a];
            }
            la[0][i] = a_phi;
            la[1][i] = a_rho;
            la[2][i] = a_u[0]; la[3][i] = a_u[1]; la[4][i] = a_u[2];
            ga[GA_UX][i] += a_D[0]; ga[GA_UY][i] += a_D[1]; ga[GA_UZ][i] += a_D[2];
            ga[GA_VX][i] += a_D[3]; ga[GA_VY][i] += a_D[4]; ga[GA_VZ][i] += a_D[5];
            ga[GA_WX][i] += a_D[6]; ga[GA_WY][i] += a_D[7]; ga[GA_WZ][i] += a_D[8];
            ga[GA_RX][i] = a_gr[0]; ga[GA_RY][i] = a_gr[1]; ga[GA_RZ][i] = a_gr[2];
            ga[GA_FX][i] = a_gp[0]; ga[GA_FY][i] = a_gp[1]; ga[GA_FZ][i] = a_gp[2];
            ga[GA_ULAP][i] = a_Lu[0]; ga[GA_VLAP][i] = a_Lu[1]; ga[GA_WLAP][i] = a_Lu[2];
            ga[GA_FLAP][i] = a_p0 * (-P->k_f * phi);
        }
        /* 6. stencils (CPU/Inamuro.cpp:140-148) */
        adj_stencil(s, ga[GA_UX], ga[GA_UY], ga[GA_UZ], ga[GA_ULAP], la[2]);
        adj_stencil(s, ga[GA_VX], ga[GA_VY], ga[GA_VZ], ga[GA_VLAP], la[3]);
        adj_stencil(s, ga[GA_WX], ga[GA_WY], ga[GA_WZ], ga[GA_WLAP], la[4]);
        adj_stencil(s, ga[GA_RX], ga[GA_RY], ga[GA_RZ], NULL, la[1]);
        adj_stencil(s, ga[GA_FX], ga[GA_FY], ga[GA_FZ], ga[GA_FLAP], la[0]);
    }
    /* initial state: fei = phi0, rho = rho_G + (phi0 - fei_min) * scale */
    {
        const double scale = (P->rho_L - P->rho_G) / (P->fei_max - P->fei_min);
        for (size_t i = 0; i < n; ++i) la[0][i] += la[1][i] * scale;
        store_canonical(s, ORC_P_PREV, la[0], grad_phi0);
        *grad_kg = kg_bar;
    }
    rc = 0;
done:
    orc_set_schedule(s, NULL, 0);
    for (int k = 0; ck && k < steps; ++k) {
        for (int j = 0; j < 6; ++j) free(ck[k].m[j]);
        free(ck[k].f); free(ck[k].g); free(ck[k].p_prev);
    }
    free(ck); free(tgt); free(rec_h); free(rec_p); free(fei_new); free(rho_new); free(tau); free(pfin);
    free(lf); free(lg); free(cf); free(cg); free(lh); free(lp); free(ltau); free(lS);
    for (int k = 0; k < 6; ++k) free(la[k]);
    for (int k = 0; k < GA_NUM; ++k) free(ga[k]);
    return rc;
}
