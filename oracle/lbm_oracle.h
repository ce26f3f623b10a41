/*
 * lbm_oracle.h -- CPU oracle for the Inamuro two-phase D3Q15 time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker against which the
 * HIP path is compared.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library
 * (lbm_amd/liblbm_hip.so) never links or calls it.
 *
 * It restates the reference CPU solver (CPU/Inamuro.cpp, CPU/Inamuro.hpp,
 * CPU/common.hpp) operation for operation, including the association order of
 * every floating-point expression and the x-outer / y / z-inner order of every
 * global sum, so that gcc -O3 (no FMA contraction) reproduces the reference's
 * results bit for bit.  Each function cites the reference lines it follows.
 *
 * Layouts at this API boundary are canonical: macro fields are [z][y][x]
 * (x fastest, no ghost layers), distributions are [q][z][y][x].
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int lx, ly, lz;
    double rho_L, rho_G;
    double mu_L, mu_G;
    double tauf, taug;
    double k_f, k_g;
    double T, a, b;
    double fei_max, fei_min;
    double fei_L, fei_G;
    double DD;
} orc_params;

/* Mirrors the CPU PressureSolveDiagnostics (CPU/Inamuro.hpp:76-83) plus the
 * GPU PoissonStepDiagnostics fields (GPU/InamuroCUDA.hpp:19-43) that the
 * restated dual residual fills. */
typedef struct orc_poisson_diag {
    int32_t pressure_solve_call_count;
    int32_t hit_max_iterations;
    uint64_t prev_iteration_count;
    uint64_t iteration_count;
    double relative_error;
    double pressure_l1_delta;
    double pressure_l1_norm;
    double divergence_mean;
    double pressure_gauge_target;
    double pressure_gauge_max_shift;
    double fixed_point_h_l1, fixed_point_h_scale, fixed_point_h_relative;
    double fixed_point_p_l1, fixed_point_p_scale, fixed_point_p_relative;
    double fixed_point_relative;
    int32_t fixed_point_evaluated;
    int32_t converged;
    /* GPU PoissonStepDiagnostics source statistics (GPU/InamuroCUDA.cu:
     * 1891-1893, 1920), sums in x-outer order: -mean(div)/3, mean|div|/3 of
     * the unprojected source and -mean(div)/3 after the projection. */
    double source_mean;
    double source_abs_mean;
    double projected_source_mean;
    /* 1 when the solve stopped at a nonzero iteration cap
     * (GPU/InamuroCUDA.cu:2189, 2337). */
    int32_t fallback_count;
    int32_t reserved;
    /* Poisson accelerators (orc_set_poisson_accel), last check of the solve:
     * sum_b |sum_{cells of b} (p - p_prev)| / sum |p - p_prev| over the 4^3
     * blocks b when spatial diagnostics are on, else NaN
     * (GPU/InamuroCUDA.cu:2078-2110); the block edge and count when block
     * sums were formed (spatial diagnostics or two-grid), else 0; and the
     * last Anderson(1) mixing coefficient (NaN when none was computed). */
    double block_low_frequency_fraction;
    int32_t block_size;
    int32_t block_count;
    double anderson_alpha;
} orc_poisson_diag;

/* The reference GPU solver's optional Poisson accelerators
 * (GPU/InamuroCUDA.hpp:60-76, GPU/InamuroCUDA.cu:539-1013, 1198-1323,
 * 1816-2349).  They leave the strict contract; the CPU reference has none
 * of them, so this is a restatement of the CUDA kernels' arithmetic. */
enum {
    ORC_POISSON_SPLIT = 0,    /* collision_p, stream, slip, getp (the reference CPU)  */
    ORC_POISSON_FUSED = 1,    /* collisionPressureStreamKernel: pull with collision at the source */
    ORC_POISSON_ONEPASS = 2,  /* collisionStreamBoundaryPressureKernel: + fixed-point relaxation */
    ORC_POISSON_SCALAR = 3    /* scalarPressureJacobiKernel: Jacobi sweep on p alone   */
};
typedef struct orc_poisson_accel {
    int32_t variant;                 /* ORC_POISSON_*                                  */
    int32_t fused_boundary;          /* boundaryAndComputePressureKernel (same values) */
    double pressure_relax_scale;     /* s in h - s (h - E p) / tau_h - E S / 3 (> 0)    */
    double fixed_point_relax;        /* onepass: h_old + w (image - h_old) (> 0)        */
    double scalar_source_scale;      /* scalar: rhs = scale * rho * div                 */
    int32_t source_aware_hh_init;    /* h = E p - scale tau_h E S / 3 at solve start    */
    int32_t anderson_m1;             /* Anderson(1) mixing of the onepass images        */
    double source_aware_hh_scale;
    double anderson_beta_max;        /* alpha clamped to [-beta, 1 + beta] (>= 0)       */
    double two_grid_strength;        /* >= 0                                            */
    int32_t two_grid;                /* coarse correction from 4^3 block sums at checks */
    int32_t spatial_diagnostics;     /* block_low_frequency_fraction at checks          */
} orc_poisson_accel;


/* Array identifiers for orc_get / orc_set. */
enum {
    ORC_FEI = 0, ORC_RHO, ORC_U, ORC_V, ORC_W, ORC_P,
    ORC_FEI_X, ORC_FEI_Y, ORC_FEI_Z,
    ORC_RHO_X, ORC_RHO_Y, ORC_RHO_Z,
    ORC_U_X, ORC_U_Y, ORC_U_Z,
    ORC_V_X, ORC_V_Y, ORC_V_Z,
    ORC_W_X, ORC_W_Y, ORC_W_Z,
    ORC_FEI_LAP, ORC_U_LAP, ORC_V_LAP, ORC_W_LAP,
    ORC_P_PREV,
    ORC_NUM_SCALAR_ARRAYS,
    ORC_F = 100, ORC_G, ORC_H
};

/* Stage identifiers for orc_stage (single reference functions). */
enum {
    ORC_STAGE_GRADIENTS = 0,  /* the 5 firstord + 4 secondord calls of collision() */
    ORC_STAGE_COLLISION,      /* full Inamuro::collision() (stencils + BGK)        */
    ORC_STAGE_STREAM_FG,      /* stream(ff), stream(gg)                             */
    ORC_STAGE_BC_FG,          /* applyBoundaryConditions(ff), (gg)                  */
    ORC_STAGE_MACRO,          /* getMacro()                                         */
    ORC_STAGE_POISSON,        /* solvePressurePoisson()                             */
    ORC_STAGE_CORRECT,        /* correct_uvw()                                      */
    ORC_STAGE_UPDATE_HH,      /* update_hh()                                        */
    ORC_STAGE_COLLISION_P,    /* one collision_p()                                  */
    ORC_STAGE_STREAM_H,       /* stream(hh)                                         */
    ORC_STAGE_BC_H,           /* applyBoundaryConditions(hh)                        */
    ORC_STAGE_GETP            /* getp()                                             */
};

typedef struct orc_sim orc_sim;

/* Installs the accelerator settings (defaults: split, split boundary, unit
 * relaxations, scalar scale 2, source-aware scale 1, beta_max 1, two-grid
 * strength 0.5, everything off).  Validation as the reference's setters
 * (GPU/InamuroCUDA.cu:2488-2540).  0 = ok. */
int orc_set_poisson_accel(orc_sim* s, const orc_poisson_accel* a);
void orc_default_poisson_accel(orc_poisson_accel* a);
/* Test helper: k iterations of the configured variant from the current h
 * and p with u_x + v_y + w_z as the source (no projection, gauge or check;
 * the Anderson history starts empty).  0 = ok. */
int orc_poisson_accel_iterations(orc_sim* s, int k);

/* Allocates arrays and applies Inamuro::initializeArrays (CPU/Inamuro.cpp:511-548):
 * f=g=h=0, rho=rho_G, fei=fei_min, u=v=w=p=0. Returns NULL on bad sizes. */
orc_sim* orc_create(const orc_params* params);
void orc_destroy(orc_sim* s);
const char* orc_last_error(const orc_sim* s);

/* Worker threads for per-cell loops (global sums always stay sequential, so
 * results are bit-identical for any thread count). */
void orc_set_threads(orc_sim* s, int nthreads);

/* Poisson loop settings.  The reference hard-codes check_interval=100 and
 * tol=1e-3 (CPU/Inamuro.cpp:670-671); iter_cap=0 is the unbounded contract,
 * a nonzero cap is a test watchdog only.  dual=1 adds the restated GPU-only
 * kinetic fixed-point residual to the exit test (DUAL_RESIDUAL_CONTRACT.md). */
int orc_set_poisson(orc_sim* s, uint64_t check_interval, double tol,
                    uint64_t iter_cap, int dual);

/* Initial conditions. */
void orc_init_two_droplets_sharp(orc_sim* s);          /* CPU/Inamuro.cpp:550-587 */
/* GPU/Inamuro.cpp:499-533 place_droplet, generalised to a velocity vector:
 * profile 0 = sharp, 1 = tanh. */
void orc_place_droplet(orc_sim* s, double cx, double cy, double cz,
                       double vx, double vy, double vz,
                       int profile, double interface_width);

int orc_get(const orc_sim* s, int array_id, double* out);
int orc_set(orc_sim* s, int array_id, const double* in);

/* One full Inamuro::performTimeStep (CPU/Inamuro.cpp:347-363). 0 = ok,
 * negative = error (message in orc_last_error). */
int orc_step(orc_sim* s);
int orc_stage(orc_sim* s, int stage);
int orc_is_first_step(const orc_sim* s);
void orc_set_first_step(orc_sim* s, int first);
void orc_poisson_diag_get(const orc_sim* s, orc_poisson_diag* out);

/* Extension: the next `count` Poisson solves run exactly iters[i] iterations
 * (gauging at every check) -- the frozen map that finite differences of the
 * adjoint are taken on. */
int orc_set_schedule(orc_sim* s, const uint64_t* iters, int count);

/* Extension (no counterpart in the reference): interface-shape objective
 * J = sum (phi_T - target)^2 after `steps` steps from phi0 (rho0 linear in
 * phi0 as in place_droplet, u = v = w = p = 0, h = 0, first step) with Poisson
 * iteration counts iters[0..steps-1], and its discrete adjoint: dJ/dphi0
 * (canonical layout) and dJ/dk_g.  Both reset the simulation state. */
int orc_objective(orc_sim* s, const double* phi0, const double* target, int steps,
                  const uint64_t* iters, double* J);
int orc_adjoint(orc_sim* s, const double* phi0, const double* target, int steps,
                const uint64_t* iters, double* J, double* grad_phi0, double* grad_kg);

/* Objective of orc_objective / orc_adjoint (extension):
 *   ORC_OBJ_MISFIT      J = sum (phi_T - target)^2 (the default)
 *   ORC_OBJ_CENTROID_Z  J = sum z w / sum w with w = phi_T - fei_G and z the
 *                       cell's layer index: the height of the droplet's
 *                       centroid; target is not read (may be NULL). */
enum { ORC_OBJ_MISFIT = 0, ORC_OBJ_CENTROID_Z = 1 };
int orc_set_objective(orc_sim* s, int kind);

/* Extension (no counterpart in the reference): no-slip walls at z = 0 and
 * z = lz-1 -- halfway bounce-back of f, g and h in the streaming step, zero
 * normal gradient in the stencils and in correct_uvw -- and a gravitational
 * acceleration gz along z, applied to the predicted velocity as the buoyant
 * acceleration gz (1 - rho_G/rho).  Both off by default; orc_adjoint covers
 * both. */
int orc_set_z_walls(orc_sim* s, int enabled);
void orc_set_gravity(orc_sim* s, double gz);

/* Closed-form helpers exposed for unit tests. */
double orc_eos(const orc_params* p, double fei);                 /* CPU/Inamuro.cpp:843-847 */
double orc_rho_of_phi(const orc_params* p, double fei);           /* CPU/Inamuro.cpp:313-327 */

/* The kinetic fixed-point residual (DUAL_RESIDUAL_CONTRACT.md:24-41,
 * GPU/InamuroCUDA.cu:1987-2057, 1114-1171) of the current h, p, rho and
 * source (u_x + v_y + w_z) against the gauge target g, exactly as a dual-mode
 * check evaluates it after its gauge: the six sums and three ratios go to the
 * fixed_point_* fields of the diagnostics.  Exposed so tests can evaluate it
 * on states of their own.  0 = ok. */
int orc_fixed_point_residual(orc_sim* s, double gauge_target);

/* Timing helper for the CPU baseline: runs k bare Poisson iterations
 * (collision_p, stream, BC, getp) with no convergence test. */
void orc_poisson_iterations(orc_sim* s, int k);

#ifdef __cplusplus
}
#endif

#endif
