/*
 * lbm_hip.h -- C ABI of the MI355X (gfx950) Inamuro two-phase D3Q15 solver.
 *
 * This is the drop-in boundary for the reference's GPU backend class
 * InamuroCUDA (GPU/InamuroCUDA.hpp:16-248): every entry point below replaces
 * one member of that class, cited next to it.  Plain pointers and sizes only;
 * no exception crosses the ABI.  Every function returns LBM_OK (0) or a
 * negative status, and lbm_last_error() returns the message the reference
 * would have thrown (same text as CPU/Inamuro.cpp and GPU/InamuroCUDA.cu).
 *
 * Host layouts (caller-owned, copies are synchronous):
 *   macro field   double[lz][ly][lx]        (x fastest, no ghost layers)
 *   distribution  double[15][lz][ly][lx]    (q outer)
 * The reference keeps [x][y][z+1] nested vectors (CPU/Inamuro.hpp:21-28) and
 * a cell*15+q device layout (GPU/InamuroCUDA.hpp:236-247); the ABI uses the
 * canonical layout above and the device uses SoA [q][z][y][x].
 *
 * Threading: one context per host thread; calls on one context are
 * serialised by the caller.  A context owns one HIP stream on the device that
 * is current when lbm_create is called.  After a numeric error the context's
 * fields are undefined until the next lbm_upload.
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_ABI_VERSION 3

/* Status codes. */
#define LBM_OK            0
#define LBM_ERR_ARG      -1   /* invalid argument (reference: std::invalid_argument) */
#define LBM_ERR_HIP      -2   /* HIP runtime failure (reference: "CUDA failure")     */
#define LBM_ERR_NUMERIC  -3   /* non-finite source / gauge / residual (runtime_error) */
#define LBM_ERR_OVERFLOW -4   /* counter overflow (reference: std::overflow_error)   */
#define LBM_ERR_STATE    -5   /* call out of order, e.g. step before upload           */
#define LBM_ERR_NOMEM    -6   /* device allocation failed                             */

/* Arithmetic of the time step. */
#define LBM_PRECISION_FP64  0  /* plain FP64, reference op order, no FMA contraction */
#define LBM_PRECISION_MIXED 1  /* FP32 f/g/h deviations + FP32 collision; FP64 p,
                                   Poisson reductions, gauge and mass sums          */

/* The 14 positional parameters of params.in (CPU/Inamuro.cpp:37-53,
 * CPU/Inamuro.hpp:39-68).  `period`, max_time_steps and output_frequency
 * belong to the driver, not to the solver. */
typedef struct lbm_params_s {
    int32_t lx, ly, lz;
    double rho_L, rho_G;
    double mu_L, mu_G;
    double tauf, taug;
    double k_f, k_g;
    double T, a, b;
    double fei_max, fei_min;
    double fei_L, fei_G;
    double DD;
} lbm_params_t;

/* Mirrors InamuroCUDA::PoissonStepDiagnostics (GPU/InamuroCUDA.hpp:19-43),
 * every field with the reference's meaning, plus the CPU solver's
 * PressureSolveDiagnostics counters (CPU/Inamuro.hpp:76-83).  As in the
 * reference (GPU/InamuroCUDA.cu:1829-1844) the fixed_point_* values are NaN
 * when the kinetic residual was not evaluated; the reference's bools are
 * int32 here. */
typedef struct lbm_poisson_diag_s {
    uint64_t iterations;
    uint64_t prev_iterations;
    int32_t pressure_solve_call_count;
    int32_t hit_max_iterations;
    double relative_error;
    double pressure_l1_delta;
    double pressure_l1_norm;
    double fixed_point_h_l1;
    double fixed_point_h_scale;
    double fixed_point_h_relative;
    double fixed_point_p_l1;
    double fixed_point_p_scale;
    double fixed_point_p_relative;
    double fixed_point_relative;
    double source_mean;            /* -mean(div)/3, GPU/InamuroCUDA.cu:1891          */
    double source_abs_mean;        /* mean|div|/3 before the projection (:1893)        */
    double projected_source_mean;  /* -mean(div)/3 after the projection (:1920)        */
    double divergence_mean;        /* mean(u_x+v_y+w_z) removed from u_x              */
    double pressure_gauge_target;
    double pressure_gauge_max_shift;
    int32_t fixed_point_evaluated;
    int32_t pressure_converged;
    int32_t fixed_point_converged;
    int32_t dual_residual_enabled;
    int32_t converged;
    int32_t fallback_count;        /* 1: stopped at the iteration limit (:2189, :2337) */
} lbm_poisson_diag_t;

/* One convergence check of the Poisson loop: the values of one row of the
 * reference's Poisson diagnostics CSV (writePoissonDiagnostic,
 * GPU/InamuroCUDA.cu:2634-2682, called from check_residual at :2136-2147).
 * fixed_point_* are NaN where the kinetic residual is not evaluated. */
typedef struct lbm_poisson_check_s {
    int64_t step;                  /* 1-based step (perf.time_step_count + 1)          */
    uint64_t iteration;            /* iterations of this solve so far                  */
    double pressure_l1_delta;
    double pressure_l1_norm;
    double relative_error;
    double fixed_point_h_l1;
    double fixed_point_h_scale;
    double fixed_point_h_relative;
    double fixed_point_p_l1;
    double fixed_point_p_scale;
    double fixed_point_p_relative;
    double fixed_point_relative;
    int32_t fixed_point_evaluated;
    int32_t pressure_converged;
    int32_t fixed_point_converged;
    int32_t dual_residual_enabled;
    int32_t converged;
    /* The CSV's last three columns (GPU/InamuroCUDA.cu:2676-2681): the edge
     * and number of the 4^3 blocks when the check formed block sums (two-grid
     * correction or spatial diagnostics, lbm_set_poisson_accel), else 0, and
     * sum_b |block sum of p - p_prev| / sum |p - p_prev| with spatial
     * diagnostics, else NaN. */
    int32_t block_size;
    int32_t block_count;
    int32_t pad;
    double block_low_frequency_fraction;
    /* No CSV column: the last Anderson(1) mixing coefficient of the solve so
     * far, NaN when none was formed. */
    double anderson_alpha;
} lbm_poisson_check_t;

typedef void (*lbm_check_fn)(void* user, const lbm_poisson_check_t* check);

/* Event-timed totals (reference: PerformanceMetrics, GPU/InamuroCUDA.hpp:115-132). */
typedef struct lbm_perf_s {
    int64_t steps;
    uint64_t poisson_iterations;
    uint64_t poisson_checks;
    double collide_ms;          /* stencils + f/g collide-and-pull              */
    double macro_ms;            /* source reduction + macro pull                */
    double poisson_ms;          /* Poisson iteration kernels only               */
    double check_ms;            /* gauge / residual / dual kernels + read-back  */
    double correct_ms;          /* velocity correction                          */
    uint64_t poisson_iter_launches;
} lbm_perf_t;

/* Cost record of the last lbm_adjoint call (no counterpart in the reference). */
typedef struct lbm_adjoint_stats_s {
    int32_t steps;                   /* T */
    int32_t checkpoint_interval;     /* K */
    int64_t store_slots;             /* Poisson iterates recorded per block */
    uint64_t forward_iterations;     /* Poisson iterations of the forward pass */
    uint64_t recomputed_iterations;  /* Poisson iterations re-run in the reverse sweep */
    uint64_t adjoint_iterations;     /* adjoint Poisson iterations (the last step's
                                        solve carries no adjoint and is skipped) */
    double forward_ms;               /* forward pass with checkpoint saves */
    double reverse_ms;               /* reverse sweep, replays and recomputation included */
    /* Convergence of the reverse Poisson solves, as the forward check measures
     * its own (sum|p - p_prev| / sum|p| per check interval), on the
     * accumulated source adjoint lambda_S: the largest over the steps whose
     * solve spans more than one check interval of the relative change over
     * the last reversed interval, and of the reverse iterations until it first
     * fell below the tolerance (-1: some step's never did). */
    double final_relative_change;
    int64_t iterations_to_tolerance;
} lbm_adjoint_stats_t;

/* InamuroCUDA(const Inamuro&) (GPU/InamuroCUDA.hpp:46): allocates device state
 * for the grid in *params.  Fields start as Inamuro::initializeArrays leaves
 * them (CPU/Inamuro.cpp:511-548) until lbm_upload is called. */
int lbm_create(const lbm_params_t* params, int precision, void** out_ctx);

/* ~InamuroCUDA (GPU/InamuroCUDA.cu:1463-1467). */
int lbm_destroy(void* ctx);

/* Message of the last failure on ctx (or, when ctx is NULL, of the last
 * failed call made without a context: lbm_create, lbm_mass_noalloc). */
const char* lbm_last_error(const void* ctx);

/* initFromCPU (GPU/InamuroCUDA.cu:1628-1673): uploads the macro fields
 * fei, rho, u, v, w, p (any may be NULL to keep the current value); resets
 * h to 0, the first-step flag (f=feq, g=geq on the next step) and the
 * persistent p_prev to 0, as a freshly constructed reference solver has. */
int lbm_upload(void* ctx, const double* fei, const double* rho, const double* u,
               const double* v, const double* w, const double* p);

/* Inamuro::initializeArrays + initializeDroplets of the CPU tree
 * (CPU/Inamuro.cpp:511-587): background fei_min / rho_G / zero velocity and
 * pressure, then two sharp droplets of diameter DD at x=0, z=lz/2 +- DD with
 * w = -+0.035.  Computed on the host with libm, then uploaded as lbm_upload. */
int lbm_init_reference(void* ctx);

/* place_droplet of the GPU-tree constructor (GPU/Inamuro.cpp:499-533):
 * radius DD/2 centred at (cx,cy,cz); profile 0 = sharp, 1 = tanh with
 * interface_width (<= 0 selects 2.0); a cell is overwritten only where the
 * droplet's phi exceeds the current phi, with rho blended linearly and the
 * velocity scaled by the liquid fraction (the reference sets w only; vx, vy
 * extend it the same way).  Applies to the current fields and then behaves
 * as lbm_upload of the result.  Call after lbm_reset_fields to start from
 * the background state. */
int lbm_place_droplet(void* ctx, double cx, double cy, double cz, double vx, double vy,
                      double vz, int profile, double interface_width);

/* Inamuro::initializeArrays (CPU/Inamuro.cpp:511-548) followed by the
 * lbm_upload reset. */
int lbm_reset_fields(void* ctx);

/* initFromCPU for the distributions (GPU/InamuroCUDA.cu:1631-1648): f, g, h
 * as the reference stores them after a step (post stream + boundary).
 * first_step=1 keeps the f=feq initialisation of the next collision. */
int lbm_upload_dists(void* ctx, const double* f, const double* g, const double* h,
                     int first_step);

/* setPoissonConvergence / setPoissonIterationLimit / setUsePoissonDualResidual
 * / setEnablePoissonFixedPointShadow (GPU/InamuroCUDA.cu:2565-2608).
 * iteration_limit = 0 is the unbounded contract; a nonzero limit must be a
 * multiple of check_interval.  dual_residual: 0 off, 1 exit requires both the
 * relative pressure change and the kinetic fixed-point residual below the
 * tolerance, 2 evaluate the fixed-point residual without changing the exit.
 * All arguments are checked before anything changes: a rejected call leaves
 * the settings as they were.  It is rejected too when a stored iteration
 * schedule holds counts that are not multiples of the new check_interval. */
int lbm_set_poisson(void* ctx, int32_t check_interval, double tolerance,
                    uint64_t iteration_limit, int dual_residual);

/* The reference GPU solver's optional Poisson accelerators
 * (GPU/InamuroCUDA.hpp:60-76, 88; kernels GPU/InamuroCUDA.cu:503-1013,
 * 1198-1323, loop :1816-2349).  They leave the strict contract
 * (WHILE_BASELINE.md) and exist for the FP64 path only, the reference's
 * precision.  One struct replaces the setters:
 *   variant                setUseFusedPoisson / setUseOnePassPoisson /
 *                          setUseScalarPoisson (LBM_POISSON_*; the reference
 *                          gives scalar, then onepass, then fused precedence)
 *   fused_boundary         setUseFusedBoundaryPressure
 *   pressure_relax_scale   setPressureRelaxScale (> 0)
 *   fixed_point_relax      setPoissonFixedPointRelax (> 0; onepass)
 *   scalar_source_scale    setScalarPoissonSourceScale
 *   source_aware_hh_init,  setUseSourceAwareHHInit, setSourceAwareHHScale
 *   source_aware_hh_scale  (not with scalar)
 *   anderson_m1,           setUsePoissonAndersonM1, setPoissonAndersonBetaMax
 *   anderson_beta_max      (>= 0; acts with onepass only)
 *   two_grid,              setUsePoissonTwoGridCorrection,
 *   two_grid_strength      setPoissonTwoGridStrength (>= 0)
 *   spatial_diagnostics    setUsePoissonSpatialDiagnostics
 * Defaults (lbm_poisson_accel_defaults) are the reference's: split, split
 * boundary, unit relaxations, scalar scale 2, source-aware scale 1, beta 1,
 * two-grid strength 0.5, all switches off.  The fused variant and the fused
 * boundary compute the split values bit for bit, so on their own they run
 * the strict kernel.  As in the reference a step fails with
 * "fixed-point residual requires frozen split+split Poisson with unit
 * relaxation and all accelerators disabled" when the dual residual is on
 * together with an accelerator (spatial diagnostics give the dual-residual
 * message).  Rejected on mixed-precision contexts; a step with z walls
 * and an accelerator fails; lbm_adjoint refuses them.  Validation as the
 * reference's setters; a rejected call changes nothing. */
enum { LBM_POISSON_SPLIT = 0, LBM_POISSON_FUSED = 1, LBM_POISSON_ONEPASS = 2, LBM_POISSON_SCALAR = 3 };
typedef struct lbm_poisson_accel_s {
    int32_t variant;
    int32_t fused_boundary;
    double pressure_relax_scale;
    double fixed_point_relax;
    double scalar_source_scale;
    int32_t source_aware_hh_init;
    int32_t anderson_m1;
    double source_aware_hh_scale;
    double anderson_beta_max;
    double two_grid_strength;
    int32_t two_grid;
    int32_t spatial_diagnostics;
} lbm_poisson_accel_t;
void lbm_poisson_accel_defaults(lbm_poisson_accel_t* out);
int lbm_set_poisson_accel(void* ctx, const lbm_poisson_accel_t* settings);
int lbm_get_poisson_accel(const void* ctx, lbm_poisson_accel_t* out);

/* setUsePoissonGraph (GPU/InamuroCUDA.hpp:80, GPU/InamuroCUDA.cu:2550-2556):
 * replay each check_interval segment of Poisson iterations as a captured HIP
 * graph instead of launching its kernels one by one.  Results are identical
 * either way; only launch overhead changes.  On by default here. */
int lbm_set_poisson_graph(void* ctx, int enabled);

/* Launch shape of the Poisson iteration kernel (no counterpart in the
 * reference): -1, the default, tunes on the context's own solves -- the
 * check-interval segments that start from a check run the candidates in
 * turn, timed with HIP events, three segments each, and the fastest is kept
 * for this grid; a candidate's index fixes it (FP64, 0..2: 256 threads with
 * non-temporal loads of h, 256 plain, 768 plain; mixed, 0..3: 1024 threads
 * with non-temporal loads, 512 with, 512 plain, 1024 plain).
 * Every choice gives bit-identical results; only the speed differs with the
 * grid's shape.  lbm_get_poisson_launch reports the one in use (-1 while
 * still tuning). */
int lbm_set_poisson_launch(void* ctx, int config);
int lbm_get_poisson_launch(const void* ctx, int* config);

/* No counterpart in the reference (extensions for the falling-droplet
 * workload), both off by default:
 * lbm_set_z_walls: no-slip walls at z = 0 and z = lz-1 -- halfway bounce-back
 *   of f, g and h in streaming, zero normal gradient in the stencils and in
 *   correct_uvw.  Not combined with the dual residual.
 * lbm_set_gravity: gravitational acceleration gz along z (negative = down),
 *   applied to the predicted velocity as the buoyant acceleration
 *   gz (1 - rho_G/rho).
 * lbm_adjoint transposes both. */
int lbm_set_z_walls(void* ctx, int enabled);
int lbm_set_gravity(void* ctx, double gz);

/* Frozen Poisson iteration counts (no counterpart in the reference): step i
 * after this call runs exactly iterations[i] iterations, gauging at every
 * check, whatever the residual; later steps converge as usual.  This makes the
 * step a smooth map of its inputs (the stopping test is not), which is what
 * finite-difference checks of the adjoint need.  count = 0 clears it.  The
 * schedule survives lbm_upload (and the other calls that reset the fields),
 * which rewind it to its first entry; lbm_adjoint neither uses nor changes
 * it. */
int lbm_set_iteration_schedule(void* ctx, const uint64_t* iterations, int64_t count);

/* Poisson iterations of every step since the last upload (the CPU solver's
 * per-step pressure_iterations, CPU/InamuroSolver.cpp:318-324); *count gets
 * the number of steps, at most `capacity` entries are written. */
int lbm_iteration_history(const void* ctx, uint64_t* out, int64_t capacity, int64_t* count);

/* performTimeStepGPU (GPU/InamuroCUDA.cu:2372-2442), n times. */
int lbm_step(void* ctx, int32_t n);

/* downloadFieldsToCPU (GPU/InamuroCUDA.cu:1675-1701). NULL skips a field. */
int lbm_download(void* ctx, double* fei, double* rho, double* u, double* v,
                 double* w, double* p);

/* downloadHH (GPU/InamuroCUDA.cu:1703-1707), extended to f and g: the
 * distributions as the reference holds them after a step. NULL skips. */
int lbm_download_dists(void* ctx, double* f, double* g, double* h);

/* getLastPoissonDiagnostics (GPU/InamuroCUDA.cu:1714-1717). */
int lbm_poisson_diag(const void* ctx, lbm_poisson_diag_t* out);

/* Per-check records, the data behind setPoissonDiagnosticsPath
 * (GPU/InamuroCUDA.hpp:87, GPU/InamuroCUDA.cu:2610-2682): fn(user, check) is
 * called on the host after every completed convergence check whose values are
 * finite (the reference writes its row at the same point), in order, from
 * inside lbm_step.  The check's device work has finished when it is called.
 * The record is valid during the call only.  NULL switches it off. */
int lbm_set_check_observer(void* ctx, lbm_check_fn fn, void* user);

/* synchronize (GPU/InamuroCUDA.cu:1709-1712). */
int lbm_synchronize(void* ctx);

/* Event timing of each phase (GPU/InamuroCUDA.cu:2380-2437); off by default. */
int lbm_set_timing(void* ctx, int enabled);
int lbm_perf_get(const void* ctx, lbm_perf_t* out);
int lbm_perf_reset(void* ctx);

/* Device global sums in FP64: sum of phi (mass) and of |u| over the grid. */
int lbm_mass(void* ctx, double* sum_phi);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the FP64 sum of
 * lbm_mass over the caller's device array d_field[0..n), written to the
 * device double *d_sum, bit-identical to lbm_mass on the same values.  Two
 * kernel launches on `stream` (a hipStream_t; NULL = the null stream), no
 * allocation, memset, copy or synchronisation: *d_sum holds the sum once the
 * stream has run them.  Its block sums use a library-owned device array of
 * 1024 doubles (8 KiB per GPU, held until the process ends) shared by every
 * call on that GPU: calls that overlap on one GPU, on any streams, race.
 * 1 <= n <= 2^31 - 1; the message of a rejected call is lbm_last_error(NULL). */
int lbm_mass_noalloc(const double* d_field, int64_t n, double* d_sum, void* stream);

/* Discrete adjoint (no counterpart in the reference; FP64 contexts only).
 * On every return, successful or not, the iteration schedule of
 * lbm_set_iteration_schedule and its position, event timing and the check
 * observer are as they were before the call; no check records are reported
 * while it runs.
 * Objective J = sum (phi_T - target)^2 (or lbm_set_adjoint_objective's
 * centroid height) after `steps` steps from phi0 with
 * rho0 = rho_G + (phi0 - fei_min)(rho_L - rho_G)/(fei_max - fei_min) (the
 * density rule of place_droplet, GPU/Inamuro.cpp:521-527), u = v = w = p = 0
 * and the first-step flag set (CPU/Inamuro.cpp:150-151).  Returns J, dJ/dphi0
 * (canonical layout; NULL skips) and dJ/dk_g (NULL skips).  The Poisson
 * iteration counts are `iterations[0..steps-1]` when given (positive multiples
 * of the check interval), otherwise those the forward pass converges to; the
 * gradient is that of the step with these counts frozen, transposing every
 * executed iteration and gauge that J depends on (J reads phi_T, which is
 * formed before the last step's Poisson solve and velocity correction, so
 * the reverse sweep neither recomputes nor transposes that solve: its
 * adjoint is identically zero).  State is checkpointed every
 * `checkpoint_interval` steps and recomputed in the reverse sweep.  On return
 * the context holds the state at the end of the forward pass. */
int lbm_adjoint(void* ctx, const double* phi0, const double* target, int32_t steps,
                int32_t checkpoint_interval, const uint64_t* iterations, double* objective,
                double* grad_phi0, double* grad_kg);

/* Objective of lbm_adjoint (no counterpart in the reference):
 *   LBM_OBJ_MISFIT      J = sum (phi_T - target)^2 (the default)
 *   LBM_OBJ_CENTROID_Z  J = sum z w / sum w, w = phi_T - fei_G, z the layer
 *                       index: the droplet's centroid height (target unused,
 *                       may be NULL). */
enum { LBM_OBJ_MISFIT = 0, LBM_OBJ_CENTROID_Z = 1 };
int lbm_set_adjoint_objective(void* ctx, int32_t kind);

/* Progress of lbm_adjoint (no counterpart in the reference): fn(user, phase,
 * step, iterations) after each forward-pass step (its Poisson iterations),
 * before each recorded block of the reverse sweep (the block's first
 * iteration) and after each reversed step.  The stream is drained before each
 * call, so a report follows the work it names.  NULL switches reports off. */
enum { LBM_PROGRESS_FORWARD_STEP = 0, LBM_PROGRESS_REVERSE_BLOCK = 1, LBM_PROGRESS_REVERSE_STEP = 2 };
typedef void (*lbm_progress_fn)(void* user, int32_t phase, int32_t step, uint64_t iterations);
int lbm_set_progress(void* ctx, lbm_progress_fn fn, void* user);

/* Storage of the adjoint Poisson state (no counterpart in the reference):
 * LBM_ADJ_STATE_FP64 (the default) or LBM_ADJ_STATE_FP32, which keeps the
 * adjoint of h between iterations in FP32 while its arithmetic, the adjoint
 * gauge reductions, the source and tau adjoints and every other adjoint stay
 * FP64.  The forward recomputation is FP64 either way. */
enum { LBM_ADJ_STATE_FP64 = 0, LBM_ADJ_STATE_FP32 = 1 };
int lbm_set_adjoint_state_precision(void* ctx, int32_t precision);

/* Device memory for recorded Poisson iterates in the reverse sweep (0 =
 * automatic, 90% of what is free beyond 2 GiB).  A solve longer than the
 * recorded block is replayed in blocks from raw-h snapshots. */
int lbm_set_adjoint_memory(void* ctx, int64_t iterate_bytes);
int lbm_adjoint_stats(const void* ctx, lbm_adjoint_stats_t* out);

#ifdef __cplusplus
}
#endif

#endif /* LBM_HIP_H */
