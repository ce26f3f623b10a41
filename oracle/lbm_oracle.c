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

struct orc_sim {
    orc_params P;
    int lx, ly, lz;
    size_t n;
    double* f; double* g; double* h; double* tmp;
    double* arr[ORC_NUM_SCALAR_ARRAYS];
    int* xn; int* yn; int* zn;          /* stencil neighbour tables per axis, [coord*3 + (c+1)] */
    int first_step;                     /* function-static is_first_step, CPU/Inamuro.cpp:150 */
    int nthreads;
    uint64_t check_interval;
    double tol;
    uint64_t iter_cap;
    int dual;
    uint64_t last_iteration_count;
    int pressure_solve_call_count;
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
    for (int z = 0; z < s->lz; ++z)
        for (int c = -1; c <= 1; ++c) {
            int zp = z + c;
            if (zp >= s->lz) zp = 0;
            if (zp < 0) zp = s->lz - 1;
            s->zn[z * 3 + c + 1] = zp;
        }
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
    return s;
}

void orc_destroy(orc_sim* s)
{
    if (!s) return;
    free(s->f); free(s->g); free(s->h); free(s->tmp);
    for (int i = 0; i < ORC_NUM_SCALAR_ARRAYS; ++i) free(A(s, i));
    free(s->xn); free(s->yn); free(s->zn);
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
                    const int zs = (z - EZ[q] + lz) % lz;
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
static int solve_pressure_poisson(orc_sim* s)
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

    const double cell_count = (double)s->lx * s->ly * s->lz;
    double divergence_sum = 0.0, pressure_sum = 0.0;
    double* ux = A(s, ORC_U_X);
    const double* vy = A(s, ORC_V_Y);
    const double* wz = A(s, ORC_W_Z);
    double* p = A(s, ORC_P);
    for (size_t i = 0; i < n; ++i) {
        divergence_sum += ux[i] + vy[i] + wz[i];
        pressure_sum += p[i];
    }
    const double divergence_mean = divergence_sum / cell_count;
    const double gauge_target = pressure_sum / cell_count;
    d->divergence_mean = divergence_mean;
    d->pressure_gauge_target = gauge_target;
    if (!isfinite(divergence_mean) || !isfinite(gauge_target)) {
        set_err(s, "Poisson source/gauge contains NaN/Inf");
        return -1;
    }
    for (size_t i = 0; i < n; ++i) ux[i] -= divergence_mean;

    uint64_t iter = 0;
    for (;;) {
        if (iter == UINT64_MAX) { set_err(s, "Poisson iteration counter overflow"); return -2; }
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
            if (error < s->tol && fp_ok) { d->converged = 1; break; }
            if (s->iter_cap != 0 && iter >= s->iter_cap) { d->hit_max_iterations = 1; break; }
        }
    }
    s->last_iteration_count = iter;
    d->iteration_count = iter;
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
                const int z_n = (z == lz - 1) ? 0 : z + 1;
                const int z_s = (z == 0) ? lz - 1 : z - 1;
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
