// kernels.hpp -- launchers shared by the host orchestration (capi.cpp) and the
// kernel translation units (kernels_f64.hip, kernels_mixed.hip,
// kernels_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

#include "d3q15.hpp"

namespace lbm {

// Physical constants of one simulation, passed by value to every kernel.
struct Phys {
    double rho_L, rho_G, mu_L, mu_G, tauf, taug, k_f, k_g, T, a, b;
    double fei_L, fei_G;
    double tmp1;     // (mu_L - mu_G) / (rho_L - rho_G), CPU/Inamuro.cpp:137
    double fei_avg;  // (fei_L + fei_G) / 2.0, CPU/Inamuro.cpp:296
    double cell_count;
    double gz = 0.0;  // extension: gravitational acceleration along z
};

// Device-resident state of one Poisson solve; the host reads it back once per
// convergence check.  Status codes name the reference's error messages.
enum : int32_t {
    ST_OK = 0,
    ST_SOURCE_NONFINITE = 1,   // "Poisson source/gauge contains NaN/Inf"
    ST_GAUGE_NONFINITE = 2,    // "pressure gauge is NaN/Inf"
    ST_PRESSURE_NONFINITE = 3, // "Poisson pressure/residual contains NaN/Inf"
    ST_RESIDUAL_NONFINITE = 4, // "Poisson residual is NaN/Inf"
    ST_FP_NONFINITE = 5,       // "kinetic fixed-point residual contains NaN/Inf"
    ST_FP_REL_NONFINITE = 6,   // "normalized kinetic fixed-point residual contains NaN/Inf"
};

// Per-block terms of the collision kernels' source reduction: sum div,
// sum p, sum |div|.
constexpr int SOURCE_TERMS = 3;

struct DevState {
    double div_mean;
    double div_abs_sum;    // sum |u_x + v_y + w_z| before the projection
    double proj_div_sum;   // sum of the projected source
    double gauge_target;
    double shift;          // gauge shift of the latest check
    double max_shift;
    double e1, e2, rel;    // pressure relative change of the latest check
    double fp[6];          // dual-residual sums
    double fp_h_rel, fp_p_rel, fp_rel;
    double mass;           // scratch for lbm_mass
    double acc_alpha;      // Poisson accelerators: last Anderson(1) coefficient
    double acc_block_abs;  // Poisson accelerators: sum_b |4^3 block sum of p - p_prev|
    int32_t status;
    int32_t converged;
    int32_t pressure_converged;
    int32_t fp_converged;
    int32_t fp_evaluated;
    int32_t pad;
};

// How a Poisson iteration obtains h and p at its own cell.
//   H_STREAMED:    h from the streamed buffer with the x-slip rule applied on
//                  read; p = sum_q h_q (the previous iteration's getp)
//   H_EQUILIBRIUM: h = p * E_q (update_hh of the previous step); p from buffer
//   H_GAUGED:      streamed h with the slip rule, minus E_q * shift; gauged p
//                  from buffer (first iteration after a check)
//   H_EXPLICIT:    h from buffer as uploaded (already post-boundary); p from
//                  buffer (first iteration after lbm_upload)
enum : int { H_STREAMED = 0, H_EQUILIBRIUM = 1, H_GAUGED = 2, H_EXPLICIT = 3 };
// Collision modes for f and g.
enum : int { FG_FIRST = 0, FG_PULL = 1, FG_LOCAL = 2 };

inline int grid_blocks(int64_t n) { return static_cast<int>((n + BLOCK - 1) / BLOCK); }
// Blocks (tiles) of the collision kernels; each writes one pair of partials.
inline int stencil_blocks(const Grid& g)
{
    return ((g.lx + STENCIL_TX - 1) / STENCIL_TX) * ((g.ly + STENCIL_TY - 1) / STENCIL_TY) *
           ((g.lz + STENCIL_ZCHUNK - 1) / STENCIL_ZCHUNK);
}
inline int red_blocks(int64_t n)
{
    const int64_t b = (n + BLOCK - 1) / BLOCK;
    return static_cast<int>(b < RED_BLOCKS ? b : RED_BLOCKS);
}

namespace f64 {

void collide(hipStream_t s, int mode, const Grid& g, const Phys& ph,
             const double* fei, const double* rho, const double* u, const double* v,
             const double* w, const double* p,
             const double* fin, const double* gin, double* fout, double* gout,
             double* ux, double* vy, double* wz, double* partials);
void source_finalize(hipStream_t s, const Grid& g, const Phys& ph, const double* partials,
                     int nparts, DevState* st);
// Also reduces the projected source into st->proj_div_sum; partials holds
// grid_blocks(g.n) doubles.
void macro(hipStream_t s, int pull, const Grid& g, const Phys& ph, const double* fin,
           const double* gin, const double* ux, const double* vy, const double* wz,
           DevState* st, double* fei, double* rho, double* u, double* v, double* w,
           double* tau, double* src, double* partials);
void projected_finalize(hipStream_t s, const double* parts, int nparts, DevState* st);
// One Poisson iteration; config selects the launch shape (block size and
// load hint) among PUSH_CONFIGS, with results identical for all of them.
constexpr int PUSH_CONFIGS = 3;
// relax != 1 selects the reference's pressure relaxation scale (an
// accelerator; the strict map is relax = 1, a kernel of its own).
void poisson_iter(hipStream_t s, int mode, const Grid& g, const double* hin, double* hout,
                  const double* p, const double* tau, const double* src, const DevState* st,
                  int config, double relax = 1.0);
// getp of a check (p = sum_q of the streamed h under the slip view) with the
// block partial sums of p for the gauge, red_blocks(g.n) of them.
void getp_partials(hipStream_t s, const Grid& g, const double* hin, double* p, double* parts);
void check(hipStream_t s, const Grid& g, const Phys& ph, double* p, double* p_prev,
           const double* hin, const double* tau, const double* src, int dual, double tol,
           double* partials, DevState* st);
void streamed_to_post_boundary(hipStream_t s, const Grid& g, const double* hin, double* out);
// FP64 reductions shared with the mixed-precision path.
void gauge_finalize(hipStream_t s, const Phys& ph, const double* parts, int nparts, DevState* st);
void gauge_error(hipStream_t s, const Grid& g, double* p, double* p_prev, const DevState* st,
                 double* parts);
void check_finalize(hipStream_t s, const double* parts_err, const double* parts_fp, int nparts,
                    int dual, double tol, DevState* st);
void mass_finalize(hipStream_t s, const double* parts, int nparts, DevState* st);
void correct(hipStream_t s, const Grid& g, const double* p, const double* rho, double* u,
             double* v, double* w);
void materialize(hipStream_t s, const Grid& g, const double* in, double* out);
void equilibrium_h(hipStream_t s, const Grid& g, const double* p, double* h);
void sum_field(hipStream_t s, const Grid& g, const double* a, double* partials, DevState* st);
// The same deterministic sum of a[0..g.n) into the device double *out;
// partials holds red_blocks(g.n) <= RED_BLOCKS doubles.  Kernel launches only.
void sum_field_into(hipStream_t s, const Grid& g, const double* a, double* partials, double* out);
// getp + gauge of a check without the convergence test (recomputation of an
// already converged solve): p = sum_q view(h), shift, p -= shift.
void gauge(hipStream_t s, const Grid& g, const Phys& ph, const double* hin, double* p,
           double* p_prev, double* partials, DevState* st);

} // namespace f64

// Discrete adjoint of the FP64 step (no counterpart in the reference).
struct AdjState {
    double J;           // objective of the forward pass
    double kg;          // dJ/dk_g accumulated over the reverse sweep
    double lam_shift;   // pending gauge adjoint Lambda/N, applied on read
    double lam_target;  // adjoint of the gauge target mean(p_start), this step
    double lS_mean;     // mean of the source adjoint (projection S - mean)
    double weight;      // centroid objective: sum of the weights phi - fei_G
    double conv_change;  // relative change of lS over the last reversed check interval
    double conv_reached; // reverse iterations until it first fell below the tolerance (-1: not yet)
};
// Objective of the adjoint: sum (phi_T - target)^2, or the centroid height
// sum z w / sum w with w = phi_T - fei_G and z the layer index.
enum : int { ADJ_OBJ_MISFIT = 0, ADJ_OBJ_CENTROID_Z = 1 };
enum : int { ADJ_SRC_PLAIN = 0, ADJ_SRC_LAZY = 1, ADJ_SRC_INIT = 2 };

namespace adj {

// parts holds 2 * grid_blocks(g.n) doubles.
void seed(hipStream_t s, const Grid& g, const Phys& ph, int objective, const double* fei,
          const double* target, double* lfei, double* parts, AdjState* as);
void correct(hipStream_t s, const Grid& g, const Phys& ph, const double* p, const double* rho,
             const double* lu, const double* lv, const double* lw, const double* lp, double* lrho,
             double* lpf, double* parts, AdjState* as);
// Bin / Bout are [q][n] carriers of the adjoint, float when state_f32 is set
// and double otherwise; the first iteration (mode H_EQUILIBRIUM) writes
// lp_start instead of Bout.
// block: threads per block of the non-gauged iterations, BLOCK or 768 (the
// results do not depend on it).
void poisson(hipStream_t s, const Grid& g, const Phys& ph, int src, int mode, bool state_f32,
             const void* Bin, const double* lpf, AdjState* as, const double* raw, const double* pin,
             const double* shift, const double* tau, double* lS, double* ltau, void* Bout,
             double* lp_start, double* parts, int block);
void source(hipStream_t s, const Grid& g, const Phys& ph, const double* fei, const double* rho,
            const double* ltau, const double* lS, const double* lfei, AdjState* as, double* lrho,
            double* lps, double* lmac, double* parts);
void collide(hipStream_t s, int first, const Grid& g, const Phys& ph, const double* fei,
             const double* rho, const double* u, const double* v, const double* w, const double* cf,
             const double* cg, const double* lmac, const double* lu, const double* lv,
             const double* lw, const double* lS, AdjState* as, double* cf_o, double* cg_o,
             double* pc, double* parts);
// Convergence of the reverse solve: after the adjoints of a check interval
// (down to iteration k of nit), the relative change sum|lS - lS_prev| /
// sum|lS| of the accumulated source adjoint since the previous one; lS_prev
// takes lS.  The first call of a step passes first = 1 (resets the record).
void converge(hipStream_t s, const Grid& g, const double* lS, double* lS_prev, uint64_t done, double tol,
              int first, double* parts, AdjState* as);
// Appends (conv_change, conv_reached) of the step just reversed to out[2 * step].
void converge_record(hipStream_t s, const AdjState* as, double* out, int step);
void stencil(hipStream_t s, const Grid& g, const double* pc, double* lfei, double* lrho, double* lu,
             double* lv, double* lw);
constexpr int PC_SLOTS = 24;  // per-cell adjoint slots of collide -> stencil

} // namespace adj

// The reference GPU solver's optional Poisson accelerators, FP64
// (kernels_accel.hip).  h is explicit and post-boundary, p is written by
// every iteration.
namespace acc {

// One pull-form iteration h -> hn, p -> pn with relaxation scale `scale`;
// onepass adds the fixed-point relaxation omega.  anderson (onepass only):
// 0 none, 1 first iteration (also writes the history pr = hn - h, pi = hn),
// 2 later iterations (block partials of the Anderson dot products into
// parts, 2 * pull_blocks(g) doubles; pn is left to anderson_mix).
// Blocks of the staged pull and scalar kernels (tiles of the collision
// kernels, marching a grid-dependent number of layers).
int pull_blocks(const Grid& g);
void pull(hipStream_t s, const Grid& g, int onepass, int anderson, double scale, double omega,
          const double* h, const double* p, const double* tau, const double* src, double* hn,
          double* pn, double* pr, double* pi, double* parts);
// Anderson(1) coefficient from the partials (into st->acc_alpha), then the
// mixed state into pi_to_a (which becomes h), pr = img - h, pn = sum_q.
void anderson_mix(hipStream_t s, const Grid& g, double beta_max, const double* parts, DevState* st,
                  const double* h, const double* img, double* pi_to_a, double* pr, double* pn);
void scalar(hipStream_t s, const Grid& g, double scale, const double* p, const double* rho,
            const double* src, double* pn);
void source_init(hipStream_t s, const Grid& g, double scale, const double* p, const double* tau,
                 const double* src, double* h);
// Number of 4^3 blocks of the grid (and blocks along x and y).
int block_count(const Grid& g, int* nbx, int* nby);
// A convergence check on explicit p (and h unless nullptr): gauge, relative
// change (with 4^3 block sums into `blocks` when blocks_needed), convergence,
// then the spatial sum_b |B_b| and the two-grid correction when asked.
void check(hipStream_t s, const Grid& g, const Phys& ph, double tol, int blocks_needed,
           int spatial, int two_grid, double strength, double* p, double* p_prev, double* h,
           double* blocks, double* partials, DevState* st);
// The same check on the strict kernel's state (streamed h, gauge applied
// lazily by the next iteration): p = getp of the streamed h, the gauge and
// relative change with block sums, the spatial sum and the two-grid
// correction, which adds E correction to every stored population of the
// cell (slip partners share their weight, so the slip view moves by the
// same amount).
void check_streamed(hipStream_t s, const Grid& g, const Phys& ph, double tol, int spatial, int two_grid,
                    double strength, const double* hin_const, double* h_stored, double* p, double* p_prev,
                    double* blocks, double* partials, DevState* st);

} // namespace acc

// Message of the last failed call made without a context (lbm_create,
// lbm_mass_noalloc), returned by lbm_last_error(NULL).
void set_global_error(const std::string& msg);
// Device address of lbm_mass_noalloc's scratch array on device `dev` (which
// must be current), resolved with hipGetSymbolAddress once per device;
// nullptr with *err set on failure (noalloc.cpp).
double* noalloc_scratch(int dev, hipError_t* err);

// Mixed precision: FP32 storage and collision arithmetic for f, g, h (as
// deviations) and the macro fields; FP64 pressure, p_prev and every global sum.
namespace mx {

struct PhysF {
    float rho_L, rho_G, mu_G, tauf_inv, taug_inv, k_f, k_g, T, a, b, fei_L, fei_G;
    float tmp1, fei_avg, visc;  // visc = 1.5 * (taug - 0.5) * 2.0
    float drho = 0.f;            // rho_L - rho_G
    float rho_arg = 0.f;         // (pi/2) / (fei_L - fei_G)
    float gz = 0.f;             // extension: gravitational acceleration along z
};

void collide(hipStream_t s, int mode, const Grid& g, const PhysF& ph, const float* fei,
             const float* rho, const float* u, const float* v, const float* w, const double* p,
             const float* fin, const float* gin, float* fout, float* gout, float* ux, float* vy,
             float* wz, double* partials);
void macro(hipStream_t s, const Grid& g, const PhysF& ph, float* fpost, const float* gpost,
           const float* ux, const float* vy, const float* wz, DevState* st, float* fei,
           float* rho, float* u, float* v, float* w, float* omega, float* src3, double* partials);
constexpr int PUSH_CONFIGS = 4;
void poisson_iter(hipStream_t s, int mode, const Grid& g, const float* hin, float* hout,
                  const float* bbase, const double* pown, float* bout, const float* omega,
                  const float* src3, const DevState* st, int config);
void getp_sum(hipStream_t s, const Grid& g, const float* hin, const float* pbase, double* p,
              double* parts);
void fixed_point(hipStream_t s, const Grid& g, const float* hin, const float* pbase,
                 const double* p, const float* omega, const float* src3, const DevState* st,
                 double* parts);
void correct(hipStream_t s, const Grid& g, const double* p, const float* rho, float* u, float* v,
             float* w);
void sum_f32(hipStream_t s, const Grid& g, const float* a, double* parts);
void to_f64(hipStream_t s, int64_t n, const float* in, double* out);
void to_f32(hipStream_t s, int64_t n, const double* in, float* out);
// Distributions as the reference holds them after a step, in FP64.
void fg_full(hipStream_t s, const Grid& g, int pulled, int is_f, const float* in, const float* fei,
             double* out);
// Uploaded post-boundary f or g (FP64) to the local deviation form.
void fg_encode(hipStream_t s, const Grid& g, int is_f, const double* in, const float* fei,
               float* out);

} // namespace mx

} // namespace lbm
