// capi.cpp -- C ABI (include/lbm_hip.h) and host orchestration of the time
// step.  Mirrors InamuroCUDA::performTimeStepGPU (GPU/InamuroCUDA.cu:2372-2442)
// and doPressurePoisson (GPU/InamuroCUDA.cu:1816-2349) with the CPU solver's
// semantics (CPU/Inamuro.cpp:347-363, 624-724):
//
//   collide (stencils + BGK + pull of f,g)  -> source/gauge-target reduction
//   -> macro (pull of f,g, rho map, tau_h or 1-1/tau_h, projected source)
//   -> Poisson: segments of `check_interval` push-form iterations, then one
//      on-device check (getp, gauge, relative change, optional dual residual)
//      and a single device-to-host read of a small state struct
//   -> velocity correction.  h = E*p is folded into the next step's first
//      iteration instead of being written.
//
// No exception crosses the ABI; messages are the reference's.
#include "../../include/lbm_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace {

using lbm::DevState;
using lbm::Grid;
using lbm::Phys;

std::string g_create_error;

struct HipError {
    hipError_t code;
    const char* what;
};

#define HIPCHK(call)                                                \
    do {                                                            \
        hipError_t e__ = (call);                                    \
        if (e__ != hipSuccess) throw HipError{e__, #call};          \
    } while (0)

struct ApiError {
    int code;
    std::string msg;
};

enum class FG { First, Post, Local };
enum class HState { Explicit, Equilibrium };

struct Ctx {
    lbm_params_t P{};
    int precision = LBM_PRECISION_FP64;
    bool mixed = false;
    size_t esz = sizeof(double);  // bytes of one stored f/g/h/macro value
    Grid g{};
    Phys ph{};
    lbm::mx::PhysF phf{};
    int device = 0;
    hipStream_t stream = nullptr;

    // Precision-typed device buffers (double in FP64, float in mixed).
    void* f[2] = {nullptr, nullptr};
    void* gd[2] = {nullptr, nullptr};
    void* h[2] = {nullptr, nullptr};
    void *fei = nullptr, *rho = nullptr, *u = nullptr, *v = nullptr, *w = nullptr;
    void *ux = nullptr, *vy = nullptr, *wz = nullptr;
    void* tau = nullptr;  // FP64: tau_h;  mixed: 2 - 1/tau_h = 4/(rho + 2)
    void* src = nullptr;  // FP64: S;      mixed: S/3
    // FP64 in both precisions (p[1] unused).
    double* p[2] = {nullptr, nullptr};
    // Mixed: FP32 decode bases of the streamed h, ping-pong.
    float* base[2] = {nullptr, nullptr};
    double* p_prev = nullptr;
    double* partials = nullptr;
    double* scratch = nullptr;   // 15*n doubles, lazily, for distribution transfers
    double* scratch1 = nullptr;  // n doubles, mixed field transfers
    DevState* dst = nullptr;
    DevState* hst = nullptr;     // pinned mirror

    int fcur = 0, hcur = 0;
    int pc = 0;  // buffer holding the current pressure
    int pb = 0;  // mixed: base buffer that decodes the streamed h
    FG fg = FG::First;
    HState hs = HState::Explicit;

    int check_interval = 100;
    double tol = 1.0e-3;
    uint64_t iter_limit = 0;
    int dual = 0;

    lbm_poisson_diag_t diag{};
    uint64_t last_iterations = 0;
    int32_t solve_calls = 0;
    std::vector<uint64_t> history;   // iterations of every step since the last upload
    std::vector<uint64_t> schedule;  // frozen per-step iteration counts (empty = converge)
    size_t sched_pos = 0;
    int64_t adj_store_bytes = 0;     // device bytes for recorded Poisson iterates (0 = automatic)
    int adj_objective = lbm::ADJ_OBJ_MISFIT;
    // Threads per block of the adjoint Poisson kernel for FP64 / FP32 adjoint
    // state (-1: timed on the first reverse iterations of the next call).
    int adj_bs[2] = {-1, -1};
    // Launch shape of the reverse sweep's recomputation of a block, whose
    // iterates go to fresh slots instead of a cache-resident ping-pong pair
    // (-1: timed on its first iterations, like adj_bs).
    int rec_cfg = -1;
    bool adj_state_f32 = false;      // adjoint Poisson carriers stored in FP32
    lbm_check_fn check_observer = nullptr;  // per-check records (Poisson diagnostics CSV)
    void* check_user = nullptr;
    lbm_progress_fn progress = nullptr;  // lbm_adjoint progress reports
    void* progress_user = nullptr;
    lbm_adjoint_stats_t adj_stats{};

    // Captured Poisson segments (setUsePoissonGraph): one graph per entry
    // mode and buffer parity, each holding check_interval iterations.
    struct SegGraph {
        int mode, hcur, pb, ci, cfg;
        hipGraphExec_t exec;
        int hcur_after, pb_after;
    };
    bool use_graphs = true;
    // Launch shape of the Poisson push kernel (block size and load hint, one
    // of the precision's PUSH_CONFIGS; every one gives the same results).
    // push_request < 0: tune on the context's own segments and keep the
    // fastest (push_cfg < 0 until then, next_segment_config).
    int push_request = -1;
    int push_cfg = -1;
    int seg_cfg = 0;               // shape of the segment being launched
    std::vector<float> tune_best;  // tuning: each candidate's fastest segment (ms)
    int tune_done = 0;             // timed tuning segments so far
    std::vector<SegGraph> graphs;

    // The reference's optional Poisson accelerators (FP64; kernels_accel.hip).
    // Their loop holds h explicitly in h[0], h[1] (and acc_pi for Anderson)
    // and p in p[pc] / acc_p; AccRoles names which buffer plays which part.
    lbm_poisson_accel_t accel{};
    double* acc_p = nullptr;       // second pressure buffer
    double* acc_pi = nullptr;      // Anderson(1): previous image (15 n)
    double* acc_pr = nullptr;      // Anderson(1): previous residual (15 n)
    double* acc_blocks = nullptr;  // 4^3 block sums of p - p_prev
    struct AccRoles {
        int h, img, pi;  // 0, 1: h[0], h[1]; 2: acc_pi
        int pcur;        // 0: p[pc], 1: acc_p
        bool hist;       // Anderson history stored in this solve
        bool operator==(const AccRoles& o) const
        {
            return h == o.h && img == o.img && pi == o.pi && pcur == o.pcur && hist == o.hist;
        }
    };
    AccRoles acc{};
    uint64_t acc_iters = 0;  // iterations of the current accelerated solve
    struct AccGraph {
        AccRoles start, end;
        int ci;
        hipGraphExec_t exec;
    };
    std::vector<AccGraph> acc_graphs;

    bool timing = false;
    lbm_perf_t perf{};
    hipEvent_t ev[12] = {};

    std::string err;

    size_t n() const { return static_cast<size_t>(g.n); }
};

template <typename T>
T* D(void* p) { return static_cast<T*>(p); }

void free_all(Ctx& c)
{
    (void)hipSetDevice(c.device);
    auto F = [](auto*& ptr) {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
    };
    for (int i = 0; i < 2; ++i) {
        F(c.f[i]); F(c.gd[i]); F(c.h[i]); F(c.p[i]); F(c.base[i]);
    }
    F(c.fei); F(c.rho); F(c.u); F(c.v); F(c.w);
    F(c.tau); F(c.src); F(c.ux); F(c.vy); F(c.wz);
    F(c.p_prev); F(c.partials); F(c.scratch); F(c.scratch1);
    F(c.acc_p); F(c.acc_pi); F(c.acc_pr); F(c.acc_blocks);
    F(c.dst);
    if (c.hst) (void)hipHostFree(c.hst);
    c.hst = nullptr;
    for (auto& e : c.ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    for (auto& gph : c.graphs) (void)hipGraphExecDestroy(gph.exec);
    c.graphs.clear();
    for (auto& gph : c.acc_graphs) (void)hipGraphExecDestroy(gph.exec);
    c.acc_graphs.clear();
    if (c.stream) (void)hipStreamDestroy(c.stream);
    c.stream = nullptr;
}

void* dalloc_bytes(size_t bytes)
{
    void* ptr = nullptr;
    const hipError_t e = hipMalloc(&ptr, bytes);
    if (e != hipSuccess) throw ApiError{LBM_ERR_NOMEM, "device allocation failed"};
    return ptr;
}

double* dalloc(size_t count) { return static_cast<double*>(dalloc_bytes(count * sizeof(double))); }

void sync(Ctx& c) { HIPCHK(hipStreamSynchronize(c.stream)); }

// Host FP64 field -> device field of the context's precision.
void put_field(Ctx& c, void* dst, const double* src)
{
    if (!src) return;
    if (!c.mixed) {
        HIPCHK(hipMemcpyAsync(dst, src, c.n() * sizeof(double), hipMemcpyHostToDevice, c.stream));
        return;
    }
    HIPCHK(hipMemcpyAsync(c.scratch1, src, c.n() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    lbm::mx::to_f32(c.stream, c.g.n, c.scratch1, D<float>(dst));
    HIPCHK(hipGetLastError());
    sync(c);
}

// Device field -> host FP64.
void get_field(Ctx& c, double* dst, void* src)
{
    if (!dst) return;
    if (!c.mixed) {
        HIPCHK(hipMemcpyAsync(dst, src, c.n() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        return;
    }
    lbm::mx::to_f64(c.stream, c.g.n, D<float>(src), c.scratch1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(dst, c.scratch1, c.n() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    sync(c);
}

void put_p(Ctx& c, const double* src)
{
    if (src) HIPCHK(hipMemcpyAsync(c.p[c.pc], src, c.n() * sizeof(double), hipMemcpyHostToDevice, c.stream));
}

void reset_state(Ctx& c)
{
    // The state a freshly constructed reference solver has before its first
    // step (CPU/Inamuro.cpp:511-548): h = 0, f/g unset, p_prev = 0.
    const size_t dist_bytes = lbm::Q * c.n() * c.esz;
    for (int i = 0; i < 2; ++i) {
        HIPCHK(hipMemsetAsync(c.h[i], 0, dist_bytes, c.stream));
        HIPCHK(hipMemsetAsync(c.f[i], 0, dist_bytes, c.stream));
        HIPCHK(hipMemsetAsync(c.gd[i], 0, dist_bytes, c.stream));
    }
    HIPCHK(hipMemsetAsync(c.p_prev, 0, c.n() * sizeof(double), c.stream));
    HIPCHK(hipMemsetAsync(c.dst, 0, sizeof(DevState), c.stream));
    c.fcur = c.hcur = 0;
    c.pb = 0;
    c.fg = FG::First;
    c.hs = HState::Explicit;
    c.last_iterations = 0;
    c.solve_calls = 0;
    c.diag = lbm_poisson_diag_t{};
    c.history.clear();
    c.sched_pos = 0;
}

void check_params(const lbm_params_t& P)
{
    if (P.lx < 3 || P.ly < 1 || P.lz < 1)
        throw ApiError{LBM_ERR_ARG, "grid needs lx >= 3 (x-slip mirror) and ly, lz >= 1"};
    const int64_t n = static_cast<int64_t>(P.lx) * P.ly * P.lz;
    if (n > (int64_t(1) << 31) - 1)
        throw ApiError{LBM_ERR_ARG, "grid exceeds 2^31-1 cells"};
    if (!(P.rho_L != P.rho_G) || !(P.fei_L != P.fei_G))
        throw ApiError{LBM_ERR_ARG, "rho_L/rho_G and fei_L/fei_G must differ"};
}

Ctx* create(const lbm_params_t& P, int precision)
{
    check_params(P);
    if (precision != LBM_PRECISION_FP64 && precision != LBM_PRECISION_MIXED)
        throw ApiError{LBM_ERR_ARG, "unknown precision"};
    Ctx* c = new Ctx();
    try {
        c->P = P;
        c->precision = precision;
        c->mixed = precision == LBM_PRECISION_MIXED;
        c->esz = c->mixed ? sizeof(float) : sizeof(double);
        HIPCHK(hipGetDevice(&c->device));
        c->g = Grid{P.lx, P.ly, P.lz, static_cast<int64_t>(P.lx) * P.ly * P.lz};
        c->ph = Phys{P.rho_L, P.rho_G, P.mu_L, P.mu_G, P.tauf, P.taug, P.k_f, P.k_g, P.T, P.a, P.b,
                     P.fei_L, P.fei_G,
                     (P.mu_L - P.mu_G) / (P.rho_L - P.rho_G),
                     (P.fei_L + P.fei_G) / 2.0,
                     static_cast<double>(P.lx) * P.ly * P.lz};
        c->phf = lbm::mx::PhysF{
            static_cast<float>(P.rho_L), static_cast<float>(P.rho_G), static_cast<float>(P.mu_G),
            static_cast<float>(1.0 / P.tauf), static_cast<float>(1.0 / P.taug),
            static_cast<float>(P.k_f), static_cast<float>(P.k_g), static_cast<float>(P.T),
            static_cast<float>(P.a), static_cast<float>(P.b), static_cast<float>(P.fei_L),
            static_cast<float>(P.fei_G), static_cast<float>(c->ph.tmp1),
            static_cast<float>(c->ph.fei_avg), static_cast<float>(1.5 * (P.taug - 0.5) * 2.0)};
        c->phf.drho = static_cast<float>(P.rho_L - P.rho_G);
        c->phf.rho_arg = static_cast<float>(M_PI / 2.0 / (P.fei_L - P.fei_G));
        HIPCHK(hipStreamCreate(&c->stream));
        const size_t n = c->n();
        const size_t fb = n * c->esz;
        for (int i = 0; i < 2; ++i) {
            c->f[i] = dalloc_bytes(lbm::Q * fb);
            c->gd[i] = dalloc_bytes(lbm::Q * fb);
            c->h[i] = dalloc_bytes(lbm::Q * fb);
        }
        c->p[0] = dalloc(n);
        if (c->mixed) {
            c->base[0] = static_cast<float*>(dalloc_bytes(n * sizeof(float)));
            c->base[1] = static_cast<float*>(dalloc_bytes(n * sizeof(float)));
        }
        for (void** a : {&c->fei, &c->rho, &c->u, &c->v, &c->w, &c->tau, &c->src, &c->ux, &c->vy, &c->wz})
            *a = dalloc_bytes(fb);
        c->p_prev = dalloc(n);
        if (c->mixed) c->scratch1 = dalloc(n);
        const size_t nparts = std::max<size_t>(
            std::max<size_t>(std::max<size_t>(2 * static_cast<size_t>(lbm::grid_blocks(c->g.n)),
                                              2 * static_cast<size_t>(lbm::acc::pull_blocks(c->g))),
                             lbm::SOURCE_TERMS * static_cast<size_t>(lbm::stencil_blocks(c->g))),
            9 * static_cast<size_t>(lbm::RED_BLOCKS));
        c->partials = dalloc(nparts);
        c->dst = static_cast<DevState*>(dalloc_bytes(sizeof(DevState)));
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->hst), sizeof(DevState), hipHostMallocDefault));
        for (auto& e : c->ev) HIPCHK(hipEventCreate(&e));
        c->pc = c->pb = 0;
        lbm_poisson_accel_defaults(&c->accel);
        {
            // Inamuro::initializeArrays, CPU/Inamuro.cpp:519-524.
            std::vector<double> tmp(n, P.rho_G);
            put_field(*c, c->rho, tmp.data());
            sync(*c);
            std::fill(tmp.begin(), tmp.end(), P.fei_min);
            put_field(*c, c->fei, tmp.data());
            sync(*c);
            std::fill(tmp.begin(), tmp.end(), 0.0);
            put_field(*c, c->u, tmp.data());
            put_field(*c, c->v, tmp.data());
            put_field(*c, c->w, tmp.data());
            put_p(*c, tmp.data());
            sync(*c);
        }
        reset_state(*c);
        sync(*c);
        // lbm_mass_noalloc's scratch address on this device, resolved here
        // so that later calls there need no set-up.
        hipError_t se = hipSuccess;
        if (!lbm::noalloc_scratch(c->device, &se)) throw HipError{se, "hipGetSymbolAddress(scratch)"};
    } catch (...) {
        free_all(*c);
        delete c;
        throw;
    }
    return c;
}

float elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    return ms;
}

void record(Ctx& c, int i)
{
    if (c.timing) HIPCHK(hipEventRecord(c.ev[i], c.stream));
}

const char* status_message(int32_t st)
{
    switch (st) {
    case lbm::ST_SOURCE_NONFINITE: return "Poisson source/gauge contains NaN/Inf";
    case lbm::ST_GAUGE_NONFINITE: return "pressure gauge is NaN/Inf";
    case lbm::ST_PRESSURE_NONFINITE: return "Poisson pressure/residual contains NaN/Inf";
    case lbm::ST_RESIDUAL_NONFINITE: return "Poisson residual is NaN/Inf";
    case lbm::ST_FP_NONFINITE: return "kinetic fixed-point residual contains NaN/Inf";
    case lbm::ST_FP_REL_NONFINITE: return "normalized kinetic fixed-point residual contains NaN/Inf";
    default: return "unknown device status";
    }
}

int fg_mode(const Ctx& c)
{
    return c.fg == FG::First ? lbm::FG_FIRST : (c.fg == FG::Post ? lbm::FG_PULL : lbm::FG_LOCAL);
}

// Collision, source reduction and macro pass of one step.
void step_fg(Ctx& c)
{
    hipStream_t s = c.stream;
    const Grid& g = c.g;
    const int mode = fg_mode(c);
    const int nxt = 1 - c.fcur;
    record(c, 0);
    if (!c.mixed) {
        lbm::f64::collide(s, mode, g, c.ph, D<double>(c.fei), D<double>(c.rho), D<double>(c.u),
                          D<double>(c.v), D<double>(c.w), c.p[c.pc], D<double>(c.f[c.fcur]),
                          D<double>(c.gd[c.fcur]), D<double>(c.f[nxt]), D<double>(c.gd[nxt]),
                          D<double>(c.ux), D<double>(c.vy), D<double>(c.wz), c.partials);
    } else {
        lbm::mx::collide(s, mode, g, c.phf, D<float>(c.fei), D<float>(c.rho), D<float>(c.u),
                         D<float>(c.v), D<float>(c.w), c.p[c.pc], D<float>(c.f[c.fcur]),
                         D<float>(c.gd[c.fcur]), D<float>(c.f[nxt]), D<float>(c.gd[nxt]),
                         D<float>(c.ux), D<float>(c.vy), D<float>(c.wz), c.partials);
    }
    HIPCHK(hipGetLastError());
    c.fcur = nxt;
    c.fg = FG::Post;
    record(c, 1);
    lbm::f64::source_finalize(s, g, c.ph, c.partials, lbm::stencil_blocks(g), c.dst);
    if (!c.mixed) {
        lbm::f64::macro(s, 1, g, c.ph, D<double>(c.f[c.fcur]), D<double>(c.gd[c.fcur]),
                        D<double>(c.ux), D<double>(c.vy), D<double>(c.wz), c.dst, D<double>(c.fei),
                        D<double>(c.rho), D<double>(c.u), D<double>(c.v), D<double>(c.w),
                        D<double>(c.tau), D<double>(c.src), c.partials);
    } else {
        lbm::mx::macro(s, g, c.phf, D<float>(c.f[c.fcur]), D<float>(c.gd[c.fcur]), D<float>(c.ux),
                       D<float>(c.vy), D<float>(c.wz), c.dst, D<float>(c.fei), D<float>(c.rho),
                       D<float>(c.u), D<float>(c.v), D<float>(c.w), D<float>(c.tau), D<float>(c.src),
                       c.partials);
    }
    HIPCHK(hipGetLastError());
    record(c, 2);
}

// The push kernel's launch on the current buffers, in config `cfg`.
void launch_push(Ctx& c, int mode, int cfg)
{
    const Grid& g = c.g;
    if (!c.mixed) {
        // Only the split-family accelerators reach this kernel; its
        // relaxation scale is 1 on the strict path.
        lbm::f64::poisson_iter(c.stream, mode, g, D<double>(c.h[c.hcur]), D<double>(c.h[1 - c.hcur]),
                               c.p[c.pc], D<double>(c.tau), D<double>(c.src), c.dst, cfg,
                               c.accel.pressure_relax_scale);
    } else {
        // Decodes with base pb.  The first iteration of a segment (equilibrium,
        // explicit or gauged entry) writes new bases into the other buffer;
        // streamed iterations encode against base pb itself.
        lbm::mx::poisson_iter(c.stream, mode, g, D<float>(c.h[c.hcur]), D<float>(c.h[1 - c.hcur]),
                              c.base[c.pb], c.p[c.pc], c.base[1 - c.pb], D<float>(c.tau),
                              D<float>(c.src), c.dst, cfg);
    }
}

// Launch shape of the next segment.  A requested shape is used as is.
// Otherwise the context tunes on its own solves: segments that start from a
// check (the bulk of every long solve) run the candidates in turn, each timed
// with HIP events around its launch or graph replay, until every candidate
// has TUNE_ROUNDS timed segments; the candidate with the fastest segment is
// kept from then on.  The timing sees the solver's real ping-pong buffers,
// cache state and launch mode.  Every candidate computes the same values, so
// switching shapes between segments changes no result.
constexpr int TUNE_ROUNDS = 3;

int next_segment_config(Ctx& c, int hmode, bool* timed)
{
    *timed = false;
    if (c.push_request >= 0) return c.push_request;
    if (c.push_cfg >= 0) return c.push_cfg;
    const int ncfg = c.mixed ? lbm::mx::PUSH_CONFIGS : lbm::f64::PUSH_CONFIGS;
    if (c.tune_best.empty()) c.tune_best.assign(static_cast<size_t>(ncfg), std::numeric_limits<float>::max());
    *timed = hmode == lbm::H_GAUGED;
    return c.tune_done % ncfg;
}

void tune_record(Ctx& c, int cand, float ms)
{
    const int ncfg = static_cast<int>(c.tune_best.size());
    c.tune_best[static_cast<size_t>(cand)] = std::min(c.tune_best[static_cast<size_t>(cand)], ms);
    if (++c.tune_done == TUNE_ROUNDS * ncfg)
        c.push_cfg = static_cast<int>(std::min_element(c.tune_best.begin(), c.tune_best.end()) - c.tune_best.begin());
}

// One Poisson iteration; updates the buffer bookkeeping.
void poisson_iteration(Ctx& c, int mode)
{
    launch_push(c, mode, c.seg_cfg);
    if (c.mixed && mode != lbm::H_STREAMED) c.pb = 1 - c.pb;
    c.hcur = 1 - c.hcur;
}

// Captured segments hold the launch arguments by value (the grid and its
// boundary mode included); settings that change them drop the cache.
void clear_graphs(Ctx& c)
{
    for (auto& gph : c.graphs) (void)hipGraphExecDestroy(gph.exec);
    c.graphs.clear();
    for (auto& gph : c.acc_graphs) (void)hipGraphExecDestroy(gph.exec);
    c.acc_graphs.clear();
}

// check_interval iterations, the first in `mode`, the rest H_STREAMED, in
// launch shape c.seg_cfg: eagerly, or by replaying a captured graph of exactly
// these launches (buildPoissonGraphSegment, GPU/InamuroCUDA.cu:1482-1534).
// Kernel nodes only.  Returns true when it had to capture the graph first.
bool poisson_segment(Ctx& c, int mode)
{
    if (!c.use_graphs) {
        for (int k = 0; k < c.check_interval; ++k) {
            poisson_iteration(c, mode);
            mode = lbm::H_STREAMED;
        }
        return false;
    }
    for (const auto& gph : c.graphs) {
        if (gph.mode == mode && gph.hcur == c.hcur && gph.pb == c.pb && gph.ci == c.check_interval &&
            gph.cfg == c.seg_cfg) {
            HIPCHK(hipGraphLaunch(gph.exec, c.stream));
            c.hcur = gph.hcur_after;
            c.pb = gph.pb_after;
            return false;
        }
    }
    Ctx::SegGraph gph{mode, c.hcur, c.pb, c.check_interval, c.seg_cfg, nullptr, 0, 0};
    HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    hipGraph_t graph = nullptr;
    try {
        int m = mode;
        for (int k = 0; k < c.check_interval; ++k) {
            poisson_iteration(c, m);
            m = lbm::H_STREAMED;
        }
        HIPCHK(hipGetLastError());
    } catch (...) {
        (void)hipStreamEndCapture(c.stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        c.hcur = gph.hcur;
        c.pb = gph.pb;
        throw;
    }
    HIPCHK(hipStreamEndCapture(c.stream, &graph));
    const hipError_t e = hipGraphInstantiate(&gph.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        c.hcur = gph.hcur;
        c.pb = gph.pb;
        throw HipError{e, "hipGraphInstantiate"};
    }
    gph.hcur_after = c.hcur;
    gph.pb_after = c.pb;
    c.graphs.push_back(gph);
    HIPCHK(hipGraphLaunch(gph.exec, c.stream));
    return true;
}

// getp + gauge + relative change (+ dual residual) after a segment.
void poisson_check(Ctx& c)
{
    hipStream_t s = c.stream;
    const Grid& g = c.g;
    if (!c.mixed) {
        lbm::f64::check(s, g, c.ph, c.p[c.pc], c.p_prev, D<double>(c.h[c.hcur]), D<double>(c.tau),
                        D<double>(c.src), c.dual, c.tol, c.partials, c.dst);
        return;
    }
    const int nb = lbm::red_blocks(g.n);
    double* parts_sum = c.partials;
    double* parts_err = c.partials + lbm::RED_BLOCKS;
    double* parts_fp = c.partials + 3 * lbm::RED_BLOCKS;
    lbm::mx::getp_sum(s, g, D<float>(c.h[c.hcur]), c.base[c.pb], c.p[c.pc], parts_sum);
    lbm::f64::gauge_finalize(s, c.ph, parts_sum, nb, c.dst);
    lbm::f64::gauge_error(s, g, c.p[c.pc], c.p_prev, c.dst, parts_err);
    if (c.dual)
        lbm::mx::fixed_point(s, g, D<float>(c.h[c.hcur]), c.base[c.pb], c.p[c.pc], D<float>(c.tau),
                             D<float>(c.src), c.dst, parts_fp);
    lbm::f64::check_finalize(s, parts_err, parts_fp, nb, c.dual, c.tol, c.dst);
}

// ---- the reference's optional Poisson accelerators -------------------------

// Which loop an accelerator setting runs on.  Onepass (with or without
// Anderson) and scalar Jacobi need their own state (explicit h, p written by
// every iteration): the pull loop.  The split and fused variants compute the
// split map, so the strict push kernel carries them, with the relaxation
// scale as a kernel variant, the source-aware start as an explicit first h,
// and the two-grid correction and spatial diagnostics in its check.  The
// fused boundary kernel, and Anderson or the fixed-point relaxation without
// onepass, change no value (GPU/InamuroCUDA.cu:2216-2323).
bool accel_pull(const Ctx& c)
{
    return c.accel.variant == LBM_POISSON_ONEPASS || c.accel.variant == LBM_POISSON_SCALAR;
}

bool accel_push(const Ctx& c)
{
    const lbm_poisson_accel_t& a = c.accel;
    return !accel_pull(c) &&
           (a.pressure_relax_scale != 1.0 || a.source_aware_hh_init || a.two_grid || a.spatial_diagnostics);
}

bool accel_on(const Ctx& c) { return accel_pull(c) || accel_push(c); }

// Settings the reference's fixed-point residual refuses
// (GPU/InamuroCUDA.cu:1846-1857), effective or not.
bool accel_requested(const Ctx& c)
{
    const lbm_poisson_accel_t& a = c.accel;
    return a.variant != LBM_POISSON_SPLIT || a.fused_boundary || a.source_aware_hh_init || a.anderson_m1 ||
           a.two_grid || a.pressure_relax_scale != 1.0 || a.fixed_point_relax != 1.0;
}

double* acc_h(Ctx& c, int role) { return role == 2 ? c.acc_pi : D<double>(c.h[role]); }

// The 4^3 block sums of the two-grid correction and spatial diagnostics.
double* acc_blocks(Ctx& c)
{
    if (!c.acc_blocks) {
        int nbx = 0, nby = 0;
        c.acc_blocks = dalloc(static_cast<size_t>(lbm::acc::block_count(c.g, &nbx, &nby)));
    }
    return c.acc_blocks;
}
double* acc_pbuf(Ctx& c, int which) { return which == 0 ? c.p[c.pc] : c.acc_p; }
bool acc_anderson(const Ctx& c) { return c.accel.anderson_m1 && c.accel.variant == LBM_POISSON_ONEPASS; }

// Start of an accelerated solve: buffers, the explicit h the loop iterates
// (update_hh's E p, or the uploaded post-boundary h), and the source-aware
// start (sourceAwareHHInitKernel, GPU/InamuroCUDA.cu:2162-2172).
void accel_begin(Ctx& c)
{
    const lbm_poisson_accel_t& a = c.accel;
    const size_t n = c.n();
    const bool scalar = a.variant == LBM_POISSON_SCALAR;
    if (!c.acc_p) c.acc_p = dalloc(n);
    acc_blocks(c);
    if (acc_anderson(c) && !c.acc_pi) {
        c.acc_pi = dalloc(lbm::Q * n);
        c.acc_pr = dalloc(lbm::Q * n);
    }
    Ctx::AccRoles r{};
    if (c.hs == HState::Equilibrium) {
        r.h = 0;
        r.img = 1;
        if (!scalar) lbm::f64::equilibrium_h(c.stream, c.g, c.p[c.pc], D<double>(c.h[0]));
    } else {
        r.h = c.hcur;
        r.img = 1 - c.hcur;
    }
    r.pi = 2;
    r.pcur = 0;
    r.hist = false;
    if (a.source_aware_hh_init && !scalar)
        lbm::acc::source_init(c.stream, c.g, a.source_aware_hh_scale, c.p[c.pc], D<double>(c.tau),
                              D<double>(c.src), acc_h(c, r.h));
    HIPCHK(hipGetLastError());
    c.acc = r;
    c.acc_iters = 0;
}

// One accelerated iteration on the buffers named by r (GPU/InamuroCUDA.cu:
// 2216-2294); r advances to the roles after it.
void accel_iteration(Ctx& c, Ctx::AccRoles& r)
{
    const lbm_poisson_accel_t& a = c.accel;
    hipStream_t s = c.stream;
    const Grid& g = c.g;
    double* p = acc_pbuf(c, r.pcur);
    double* pn = acc_pbuf(c, 1 - r.pcur);
    if (a.variant == LBM_POISSON_SCALAR) {
        lbm::acc::scalar(s, g, a.scalar_source_scale, p, D<double>(c.rho), D<double>(c.src), pn);
        r.pcur = 1 - r.pcur;
        return;
    }
    const bool onepass = a.variant == LBM_POISSON_ONEPASS;
    double* h = acc_h(c, r.h);
    double* img = acc_h(c, r.img);
    const double* tau = D<double>(c.tau);
    const double* src = D<double>(c.src);
    if (!acc_anderson(c) || !r.hist) {
        const int mode = acc_anderson(c) ? 1 : 0;
        lbm::acc::pull(s, g, onepass, mode, a.pressure_relax_scale, a.fixed_point_relax, h, p, tau, src, img,
                       pn, c.acc_pr, mode ? acc_h(c, r.pi) : nullptr, nullptr);
        std::swap(r.h, r.img);
        r.pcur = 1 - r.pcur;
        if (mode) r.hist = true;
        return;
    }
    lbm::acc::pull(s, g, 1, 2, a.pressure_relax_scale, a.fixed_point_relax, h, p, tau, src, img, pn,
                   c.acc_pr, nullptr, c.partials);
    lbm::acc::anderson_mix(s, g, a.anderson_beta_max, c.partials, c.dst, h, img, acc_h(c, r.pi), c.acc_pr, pn);
    // The mixed state went into the previous image's buffer: it is the new h,
    // the image becomes the previous image, the old h is free for the next one.
    const int old_h = r.h;
    r.h = r.pi;
    r.pi = r.img;
    r.img = old_h;
    r.pcur = 1 - r.pcur;
}

// check_interval accelerated iterations, eagerly or as a captured graph per
// starting role assignment (the kernels' arguments, settings included, are
// fixed by it; lbm_set_poisson_accel drops the cache).
void accel_segment(Ctx& c)
{
    c.acc_iters += static_cast<uint64_t>(c.check_interval);
    if (!c.use_graphs) {
        for (int k = 0; k < c.check_interval; ++k) accel_iteration(c, c.acc);
        return;
    }
    for (const auto& gph : c.acc_graphs) {
        if (gph.start == c.acc && gph.ci == c.check_interval) {
            HIPCHK(hipGraphLaunch(gph.exec, c.stream));
            c.acc = gph.end;
            return;
        }
    }
    Ctx::AccGraph gph{c.acc, c.acc, c.check_interval, nullptr};
    HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    hipGraph_t graph = nullptr;
    try {
        for (int k = 0; k < c.check_interval; ++k) accel_iteration(c, gph.end);
        HIPCHK(hipGetLastError());
    } catch (...) {
        (void)hipStreamEndCapture(c.stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    HIPCHK(hipStreamEndCapture(c.stream, &graph));
    const hipError_t e = hipGraphInstantiate(&gph.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) throw HipError{e, "hipGraphInstantiate"};
    c.acc_graphs.push_back(gph);
    HIPCHK(hipGraphLaunch(gph.exec, c.stream));
    c.acc = gph.end;
}

// check_residual of the accelerated loop (GPU/InamuroCUDA.cu:1960-2160).
void accel_check(Ctx& c)
{
    const lbm_poisson_accel_t& a = c.accel;
    const bool scalar = a.variant == LBM_POISSON_SCALAR;
    lbm::acc::check(c.stream, c.g, c.ph, c.tol, a.two_grid || a.spatial_diagnostics, a.spatial_diagnostics,
                    a.two_grid, a.two_grid_strength, acc_pbuf(c, c.acc.pcur), c.p_prev,
                    scalar ? nullptr : acc_h(c, c.acc.h), c.acc_blocks, c.partials, c.dst);
}

// The solved pressure back into p[pc], where correct_uvw and the next step
// read it.
void accel_end(Ctx& c)
{
    if (c.acc.pcur == 1)
        HIPCHK(hipMemcpyAsync(c.p[c.pc], c.acc_p, c.n() * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
}

// Called after every completed convergence check with the iteration count
// (the adjoint's replays record raw-h snapshots through it).
struct CheckHook {
    void (*fn)(Ctx&, uint64_t, void*);
    void* arg;
};

// One performTimeStep.
void step_one(Ctx& c, const CheckHook* hook = nullptr)
{
    hipStream_t s = c.stream;
    const bool acc = accel_on(c);
    const bool pull = accel_pull(c);
    const bool blocks = acc && (c.accel.two_grid || c.accel.spatial_diagnostics);
    // doPressurePoisson's refusals (GPU/InamuroCUDA.cu:1846-1865), made
    // before the step starts so that a refused call changes nothing.
    if (c.dual != 0 && accel_requested(c))
        throw ApiError{LBM_ERR_STATE,
                       "fixed-point residual requires frozen split+split Poisson with unit relaxation and all "
                       "accelerators disabled"};
    if (c.dual != 0 && c.accel.spatial_diagnostics)
        throw ApiError{LBM_ERR_STATE,
                       "dual-residual exit requires check_interval=100, tolerance=1e-3, poisson_iteration_limit=0, "
                       "deterministic reductions, no spatial diagnostics"};
    if (acc && c.g.zwall)
        throw ApiError{LBM_ERR_STATE, "Poisson accelerators support the reference's periodic z only (no z walls)"};
    step_fg(c);

    ++c.solve_calls;
    lbm_poisson_diag_t& d = c.diag;
    const uint64_t prev = c.last_iterations;
    d = lbm_poisson_diag_t{};
    d.prev_iterations = prev;
    d.pressure_solve_call_count = c.solve_calls;
    d.dual_residual_enabled = c.dual == 1;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    d.fixed_point_h_l1 = d.fixed_point_h_scale = d.fixed_point_h_relative = nan;
    d.fixed_point_p_l1 = d.fixed_point_p_scale = d.fixed_point_p_relative = nan;
    d.fixed_point_relative = nan;

    int hmode = c.hs == HState::Equilibrium ? lbm::H_EQUILIBRIUM : lbm::H_EXPLICIT;
    const uint64_t forced = c.sched_pos < c.schedule.size() ? c.schedule[c.sched_pos] : 0;
    if (pull) {
        accel_begin(c);
    } else {
        if (acc && c.accel.source_aware_hh_init) {
            // sourceAwareHHInitKernel (GPU/InamuroCUDA.cu:2162-2172): an
            // explicit post-boundary h, read as uploaded h is.
            lbm::acc::source_init(s, c.g, c.accel.source_aware_hh_scale, c.p[c.pc], D<double>(c.tau),
                                  D<double>(c.src), D<double>(c.h[c.hcur]));
            HIPCHK(hipGetLastError());
            hmode = lbm::H_EXPLICIT;
        }
    }
    uint64_t iter = 0;
    bool first_segment = true;
    for (;;) {
        record(c, 4);
        if (iter > std::numeric_limits<uint64_t>::max() - static_cast<uint64_t>(c.check_interval))
            throw ApiError{LBM_ERR_OVERFLOW, "Poisson iteration counter overflow"};
        iter += static_cast<uint64_t>(c.check_interval);
        bool timed = false, captured = false;
        if (!pull) c.seg_cfg = next_segment_config(c, hmode, &timed);
        if (timed) HIPCHK(hipEventRecord(c.ev[10], s));
        if (pull)
            accel_segment(c);
        else
            captured = poisson_segment(c, hmode);
        if (timed) HIPCHK(hipEventRecord(c.ev[11], s));
        HIPCHK(hipGetLastError());
        record(c, 5);
        if (pull) {
            accel_check(c);
        } else if (blocks) {
            double* hs = D<double>(c.h[c.hcur]);
            lbm::acc::check_streamed(s, c.g, c.ph, c.tol, c.accel.spatial_diagnostics, c.accel.two_grid,
                                     c.accel.two_grid_strength, hs, hs, c.p[c.pc], c.p_prev, acc_blocks(c),
                                     c.partials, c.dst);
        } else {
            poisson_check(c);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(c.hst, c.dst, sizeof(DevState), hipMemcpyDeviceToHost, s));
        record(c, 6);
        sync(c);
        if (c.timing) {
            if (first_segment) {
                c.perf.collide_ms += elapsed(c.ev[0], c.ev[1]);
                c.perf.macro_ms += elapsed(c.ev[1], c.ev[2]);
            }
            c.perf.poisson_ms += elapsed(c.ev[4], c.ev[5]);
            c.perf.check_ms += elapsed(c.ev[5], c.ev[6]);
        }
        first_segment = false;
        if (timed && !captured) tune_record(c, c.seg_cfg, elapsed(c.ev[10], c.ev[11]));
        c.perf.poisson_iter_launches += static_cast<uint64_t>(c.check_interval);
        c.perf.poisson_checks += 1;
        const DevState& st = *c.hst;
        d.iterations = iter;
        d.relative_error = st.rel;
        d.pressure_l1_delta = st.e1;
        d.pressure_l1_norm = st.e2;
        d.divergence_mean = st.div_mean;
        // GPU/InamuroCUDA.cu:1891-1893, 1920.
        d.source_mean = -st.div_mean / 3.0;
        d.source_abs_mean = st.div_abs_sum / (3.0 * c.ph.cell_count);
        d.projected_source_mean = -st.proj_div_sum / (3.0 * c.ph.cell_count);
        d.pressure_gauge_target = st.gauge_target;
        d.pressure_gauge_max_shift = st.max_shift;
        d.pressure_converged = st.pressure_converged;
        if (c.dual) {
            d.fixed_point_h_l1 = st.fp[0];
            d.fixed_point_h_scale = st.fp[1] > st.fp[2] ? st.fp[1] : st.fp[2];
            d.fixed_point_h_relative = st.fp_h_rel;
            d.fixed_point_p_l1 = st.fp[3];
            d.fixed_point_p_scale = st.fp[4] > st.fp[5] ? st.fp[4] : st.fp[5];
            d.fixed_point_p_relative = st.fp_p_rel;
            d.fixed_point_relative = st.fp_rel;
            d.fixed_point_evaluated = st.fp_evaluated;
            d.fixed_point_converged = st.fp_converged;
        }
        d.converged = st.converged;
        if (st.status != lbm::ST_OK) throw ApiError{LBM_ERR_NUMERIC, status_message(st.status)};
        if (c.check_observer) {
            // check_residual's row, GPU/InamuroCUDA.cu:2136-2147.
            lbm_poisson_check_t rec{};
            rec.step = c.perf.steps + 1;
            rec.iteration = iter;
            rec.pressure_l1_delta = d.pressure_l1_delta;
            rec.pressure_l1_norm = d.pressure_l1_norm;
            rec.relative_error = d.relative_error;
            rec.fixed_point_h_l1 = d.fixed_point_h_l1;
            rec.fixed_point_h_scale = d.fixed_point_h_scale;
            rec.fixed_point_h_relative = d.fixed_point_h_relative;
            rec.fixed_point_p_l1 = d.fixed_point_p_l1;
            rec.fixed_point_p_scale = d.fixed_point_p_scale;
            rec.fixed_point_p_relative = d.fixed_point_p_relative;
            rec.fixed_point_relative = d.fixed_point_relative;
            rec.fixed_point_evaluated = d.fixed_point_evaluated;
            rec.pressure_converged = d.pressure_converged;
            rec.fixed_point_converged = d.fixed_point_converged;
            rec.dual_residual_enabled = d.dual_residual_enabled;
            rec.converged = d.converged;
            rec.block_low_frequency_fraction = std::numeric_limits<double>::quiet_NaN();
            rec.anderson_alpha = std::numeric_limits<double>::quiet_NaN();
            if (blocks) {
                int nbx = 0, nby = 0;
                rec.block_count = lbm::acc::block_count(c.g, &nbx, &nby);
                rec.block_size = 4;
                // GPU/InamuroCUDA.cu:2107-2110
                if (c.accel.spatial_diagnostics)
                    rec.block_low_frequency_fraction = st.e1 > 0.0 ? st.acc_block_abs / st.e1 : 0.0;
            }
            if (pull && acc_anderson(c) && c.acc_iters >= 2) rec.anderson_alpha = st.acc_alpha;
            c.check_observer(c.check_user, &rec);
        }
        if (hook) hook->fn(c, iter, hook->arg);
        if (forced != 0) {
            if (iter >= forced) break;
            hmode = lbm::H_GAUGED;
            continue;
        }
        if (st.converged) break;
        if (c.iter_limit != 0 && iter >= c.iter_limit) {
            d.hit_max_iterations = 1;
            d.fallback_count = 1;  // GPU/InamuroCUDA.cu:2335-2338
            break;
        }
        hmode = lbm::H_GAUGED;
    }
    if (pull) accel_end(c);
    c.last_iterations = iter;
    c.perf.poisson_iterations += iter;
    c.history.push_back(iter);
    if (forced != 0) ++c.sched_pos;

    // Velocity correction; h = E*p is implied for the next step.
    record(c, 7);
    if (!c.mixed)
        lbm::f64::correct(s, c.g, c.p[c.pc], D<double>(c.rho), D<double>(c.u), D<double>(c.v),
                          D<double>(c.w));
    else
        lbm::mx::correct(s, c.g, c.p[c.pc], D<float>(c.rho), D<float>(c.u), D<float>(c.v),
                         D<float>(c.w));
    HIPCHK(hipGetLastError());
    record(c, 8);
    c.hs = HState::Equilibrium;
    if (c.timing) {
        HIPCHK(hipEventSynchronize(c.ev[8]));
        c.perf.correct_ms += elapsed(c.ev[7], c.ev[8]);
    }
    c.perf.steps += 1;
}

template <typename Fn>
int guarded(void* ctx, Fn&& fn)
{
    if (!ctx) return LBM_ERR_ARG;
    Ctx& c = *static_cast<Ctx*>(ctx);
    try {
        HIPCHK(hipSetDevice(c.device));
        fn(c);
        c.err.clear();
        return LBM_OK;
    } catch (const ApiError& e) {
        c.err = e.msg;
        return e.code;
    } catch (const HipError& e) {
        c.err = std::string("HIP failure: ") + hipGetErrorString(e.code) + " in " + e.what;
        return LBM_ERR_HIP;
    } catch (const std::bad_alloc&) {
        c.err = "host allocation failed";
        return LBM_ERR_NOMEM;
    } catch (const std::exception& e) {
        c.err = e.what();
        return LBM_ERR_ARG;
    }
}

struct HostFields {
    std::vector<double> fei, rho, u, v, w, p;
};

HostFields download_host(Ctx& c)
{
    HostFields h;
    for (auto* a : {&h.fei, &h.rho, &h.u, &h.v, &h.w, &h.p}) a->resize(c.n());
    get_field(c, h.fei.data(), c.fei);
    get_field(c, h.rho.data(), c.rho);
    get_field(c, h.u.data(), c.u);
    get_field(c, h.v.data(), c.v);
    get_field(c, h.w.data(), c.w);
    HIPCHK(hipMemcpyAsync(h.p.data(), c.p[c.pc], c.n() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    sync(c);
    return h;
}

void upload_host(Ctx& c, const HostFields& h)
{
    reset_state(c);
    put_field(c, c.fei, h.fei.data());
    put_field(c, c.rho, h.rho.data());
    put_field(c, c.u, h.u.data());
    put_field(c, c.v, h.v.data());
    put_field(c, c.w, h.w.data());
    put_p(c, h.p.data());
    sync(c);
}

HostFields background(const Ctx& c)
{
    // Inamuro::initializeArrays, CPU/Inamuro.cpp:519-524.
    HostFields h;
    h.fei.assign(c.n(), c.P.fei_min);
    h.rho.assign(c.n(), c.P.rho_G);
    h.u.assign(c.n(), 0.0);
    h.v.assign(c.n(), 0.0);
    h.w.assign(c.n(), 0.0);
    h.p.assign(c.n(), 0.0);
    return h;
}

size_t hidx(const Ctx& c, int x, int y, int z)
{
    return (static_cast<size_t>(z) * c.P.ly + y) * c.P.lx + x;
}

void ensure_scratch(Ctx& c)
{
    if (!c.scratch) c.scratch = dalloc(lbm::Q * c.n());
}


// ---- discrete adjoint (no counterpart in the reference) ---------------------
//
// Checkpoint-and-recompute reverse sweep of the FP64 step with frozen Poisson
// iteration counts.
//   forward pass   the full state every K steps (37 doubles per cell), the
//                  solve's final pressure of every step (1 per cell), and,
//                  while a snapshot pool has room, the raw h / gauged p /
//                  shift at iterations L, 2L, ... of solves longer than L
//   per segment    replay from its checkpoint without the Poisson solves:
//                  collide + macro, the archived final p, correct_uvw give
//                  the next step's start bit for bit
//   per step,      recompute collide + macro; then, last block first,
//   reverse        recompute the block's L Poisson iterations recording every
//                  iterate and run their adjoints (kernels_adjoint.hip);
//                  then the macro, collision and stencil adjoints.
// The recorded iterate of iteration k is the raw streamed h the forward
// kernel writes anyway, so recording adds no traffic.  A solve whose
// snapshots did not fit in the pool during the forward pass is reversed with
// binomial (revolve-style) checkpointing over its blocks in the pool space
// left free: with F free snapshots and d recompute passes, C(F + d, d) blocks.

constexpr size_t STATE_PER_CELL = 7 + 2 * lbm::Q;  // fei rho u v w p p_prev, f, g

struct StepState {
    double* buf = nullptr;
    FG fg = FG::First;
    HState hs = HState::Equilibrium;
};

struct AdjWork {
    std::vector<StepState> ckpt, seg;
    StepState final_state;
    double* pfinal = nullptr;     // final p of every step, T x n
    uint64_t L = 0;               // block length (recorded iterates)
    double* slots = nullptr;      // max(L, 2) raw-h slots, 15n each
    double* pstore = nullptr;     // gauged p of the block's checks, (L/ci + 1) x n
    double* shifts = nullptr;     // their shifts
    double* snap_pool = nullptr;  // snapshots: raw h (15n), gauged p (n), shift
    double* scratch_snap = nullptr;  // one more, for solves whose snapshots overflowed
    size_t snap_stride = 0;       // doubles per snapshot
    size_t snap_cap = 0, snap_next = 0;
    std::vector<std::vector<double*>> step_snaps;  // per step; empty = not recorded
    std::vector<char> have_snaps;   // SNAPS_*
    std::vector<size_t> pool_mark;  // per step: first pool snapshot free in its reverse
    int lcur = 0;                   // carrier of the adjoint Poisson state in use
    double *lfei = nullptr, *lrho = nullptr, *lu = nullptr, *lv = nullptr, *lw = nullptr;
    double *lp = nullptr, *lps = nullptr, *lpf = nullptr, *lS = nullptr, *ltau = nullptr;
    double *lmac = nullptr, *pc = nullptr, *parts = nullptr, *target = nullptr;
    double* lS_prev = nullptr;  // lS at the previous reversed check (convergence record)
    double* conv = nullptr;     // per step: final relative change, reverse iterations to tolerance
    void* Lb[2] = {nullptr, nullptr};  // adjoint Poisson carriers, FP64 or FP32
    double* cf[2] = {nullptr, nullptr};
    double* cg[2] = {nullptr, nullptr};
    lbm::AdjState* as = nullptr;
    lbm::AdjState* has = nullptr;
    // Timing of the adjoint kernel's two block sizes (results do not depend
    // on it): best time of each over the sampled reverse iterations.
    int bs_samples = 0;
    float bs_best[2] = {3.0e38f, 3.0e38f};
    int rec_samples = 0;
    std::vector<float> rec_best;

    ~AdjWork()
    {
        auto F = [](auto*& p) {
            if (p) (void)hipFree(p);
            p = nullptr;
        };
        for (auto& s : ckpt) F(s.buf);
        for (auto& s : seg) F(s.buf);
        F(final_state.buf);
        F(pfinal); F(slots); F(pstore); F(shifts); F(snap_pool); F(scratch_snap);
        for (double** p : {&lfei, &lrho, &lu, &lv, &lw, &lp, &lps, &lpf, &lS, &ltau, &lmac, &pc, &parts, &target,
                           &lS_prev, &conv})
            F(*p);
        for (int i = 0; i < 2; ++i) { F(Lb[i]); F(cf[i]); F(cg[i]); }
        F(as);
        if (has) (void)hipHostFree(has);
    }
    double* slot(const Ctx& c, uint64_t m) const { return slots + (m - 1) * lbm::Q * c.n(); }
    double* pool(size_t i) const { return snap_pool + i * snap_stride; }
};

// A snapshot buffer: raw h at offset 0, gauged p at 15n, shift at 16n.
struct SnapView {
    double* h;
    double* p;
    double* shift;
};

SnapView snap_view(const Ctx& c, double* base)
{
    const size_t n = c.n();
    return SnapView{base, base + lbm::Q * n, base + (lbm::Q + 1) * n};
}

void d2d(Ctx& c, void* dst, const void* src, size_t bytes)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c.stream));
}

void alloc_state(Ctx& c, StepState& st) { st.buf = dalloc(STATE_PER_CELL * c.n()); }

void save_state(Ctx& c, StepState& st)
{
    const size_t n = c.n(), nb = n * sizeof(double);
    const void* src[7] = {c.fei, c.rho, c.u, c.v, c.w, c.p[c.pc], c.p_prev};
    for (int i = 0; i < 7; ++i) d2d(c, st.buf + i * n, src[i], nb);
    d2d(c, st.buf + 7 * n, c.f[c.fcur], lbm::Q * nb);
    d2d(c, st.buf + (7 + lbm::Q) * n, c.gd[c.fcur], lbm::Q * nb);
    st.fg = c.fg;
    st.hs = c.hs;
}

void restore_state(Ctx& c, const StepState& st)
{
    const size_t n = c.n(), nb = n * sizeof(double);
    void* dst[7] = {c.fei, c.rho, c.u, c.v, c.w, c.p[c.pc], c.p_prev};
    for (int i = 0; i < 7; ++i) d2d(c, dst[i], st.buf + i * n, nb);
    d2d(c, c.f[c.fcur], st.buf + 7 * n, lbm::Q * nb);
    d2d(c, c.gd[c.fcur], st.buf + (7 + lbm::Q) * n, lbm::Q * nb);
    c.fg = st.fg;
    c.hs = st.hs;
}

// Forward-pass hook: after the check at iteration L*j keep raw h, the gauged
// p and the shift in the step's next snapshot, while the pool has room.
// With `marks` (a solve of known length whose block snapshots do not all fit)
// only the checks listed there are kept: the top level of its binomial
// checkpointing.
struct SnapRecorder {
    AdjWork* a;
    std::vector<double*>* snaps;
    bool overflow = false;
    std::vector<uint64_t> marks;
    size_t next = 0;
};

void record_snapshot(Ctx& c, uint64_t iter, void* arg)
{
    SnapRecorder& r = *static_cast<SnapRecorder*>(arg);
    if (!r.marks.empty()) {
        if (r.next >= r.marks.size() || iter != r.marks[r.next]) return;
        ++r.next;
    } else if (iter % r.a->L != 0 || r.overflow) {
        return;
    }
    if (r.a->snap_next >= r.a->snap_cap) {
        r.overflow = true;
        return;
    }
    double* base = r.a->snap_pool + (r.a->snap_next++) * r.a->snap_stride;
    r.snaps->push_back(base);
    const SnapView v = snap_view(c, base);
    const size_t nb = c.n() * sizeof(double);
    d2d(c, v.h, c.h[c.hcur], lbm::Q * nb);
    d2d(c, v.p, c.p[c.pc], nb);
    d2d(c, v.shift, &c.dst->shift, sizeof(double));
}

// A step of a segment replay: collide + macro, the archived final pressure
// (also the last check's p_prev) and the velocity correction reproduce what
// step_one leaves behind without running the solve.
void step_light(Ctx& c, const double* p_final)
{
    step_fg(c);
    const size_t nb = c.n() * sizeof(double);
    d2d(c, c.p[c.pc], p_final, nb);
    d2d(c, c.p_prev, p_final, nb);
    lbm::f64::correct(c.stream, c.g, c.p[c.pc], D<double>(c.rho), D<double>(c.u), D<double>(c.v),
                      D<double>(c.w));
    HIPCHK(hipGetLastError());
    c.hs = HState::Equilibrium;
}

// Recompute Poisson iterations a0+1..a1 of the current step (collide + macro
// already run), writing iteration k's raw h into slot k - a0 and the gauged p
// and shift of each check into the block-local store.  For a0 > 0 the block
// starts from snapshot `start` (raw h, gauged p and shift of check a0).
void adj_forward_block(Ctx& c, AdjWork& a, const double* p_start, double* start, uint64_t a0, uint64_t a1)
{
    const uint64_t ci = static_cast<uint64_t>(c.check_interval);
    const size_t n = c.n();
    int mode = lbm::H_EQUILIBRIUM;
    const double* in = a.slot(c, 1);  // not read in the equilibrium mode
    const double* plast = p_start;
    if (a0 > 0) {
        const SnapView v = snap_view(c, start);
        mode = lbm::H_GAUGED;
        in = v.h;
        plast = v.p;
        d2d(c, &c.dst->shift, v.shift, sizeof(double));
    }
    for (uint64_t k = a0 + 1; k <= a1; ++k) {
        double* out = a.slot(c, k - a0);
        // Launch shape: a requested one, else the recomputation's own, timed
        // on its first streamed iterations (the same values whichever runs).
        int cfg = c.push_request >= 0 ? c.push_request : (c.rec_cfg >= 0 ? c.rec_cfg : 0);
        const bool sample = c.push_request < 0 && c.rec_cfg < 0 && mode == lbm::H_STREAMED;
        if (sample) {
            if (a.rec_best.empty()) a.rec_best.assign(lbm::f64::PUSH_CONFIGS, 3.0e38f);
            cfg = a.rec_samples % lbm::f64::PUSH_CONFIGS;
            HIPCHK(hipEventRecord(c.ev[10], c.stream));
        }
        lbm::f64::poisson_iter(c.stream, mode, c.g, in, out, plast, D<double>(c.tau), D<double>(c.src), c.dst, cfg);
        if (sample) {
            HIPCHK(hipEventRecord(c.ev[11], c.stream));
            HIPCHK(hipEventSynchronize(c.ev[11]));
            float& best = a.rec_best[static_cast<size_t>(cfg)];
            best = std::min(best, elapsed(c.ev[10], c.ev[11]));
            if (++a.rec_samples == TUNE_ROUNDS * lbm::f64::PUSH_CONFIGS)
                c.rec_cfg = static_cast<int>(std::min_element(a.rec_best.begin(), a.rec_best.end()) -
                                             a.rec_best.begin());
        }
        if (k % ci == 0) {
            double* pk = a.pstore + ((k - a0) / ci) * n;
            lbm::f64::gauge(c.stream, c.g, c.ph, out, pk, c.p_prev, c.partials, c.dst);
            d2d(c, a.shifts + (k - a0) / ci, &c.dst->shift, sizeof(double));
            plast = pk;
            mode = lbm::H_GAUGED;
        } else {
            mode = lbm::H_STREAMED;
        }
        in = out;
    }
    HIPCHK(hipGetLastError());
}

// Recompute iterations from+1..marks.back().first of the current step through
// slots 1 and 2, from snapshot `start` (check `from`; nullptr with from = 0:
// the solve's start), and leave the state of each listed check in its
// snapshot.
void adj_forward_snaps(Ctx& c, AdjWork& a, const double* p_start, const double* start, uint64_t from,
                       const std::vector<std::pair<uint64_t, double*>>& marks)
{
    const uint64_t ci = static_cast<uint64_t>(c.check_interval);
    const size_t n = c.n();
    int mode = lbm::H_EQUILIBRIUM;
    const double* in = a.slot(c, 1);  // not read in the equilibrium mode
    const double* plast = p_start;
    if (from > 0) {
        const SnapView v = snap_view(c, const_cast<double*>(start));
        mode = lbm::H_GAUGED;
        in = v.h;
        plast = v.p;
        d2d(c, &c.dst->shift, v.shift, sizeof(double));
    }
    size_t m = 0;
    for (uint64_t k = from + 1; m < marks.size(); ++k) {
        const bool mark = k == marks[m].first;
        const SnapView v = snap_view(c, mark ? marks[m].second : a.slot(c, 1));
        double* out = mark ? v.h : a.slot(c, 1 + (k & 1));
        // a ping-pong pair, like the forward solve: its launch shape
        lbm::f64::poisson_iter(c.stream, mode, c.g, in, out, plast, D<double>(c.tau), D<double>(c.src), c.dst,
                               c.push_request >= 0 ? c.push_request : (c.push_cfg >= 0 ? c.push_cfg : 0));
        if (k % ci == 0) {
            double* pk = mark ? v.p : a.pstore + ((k / ci) & 1) * n;
            lbm::f64::gauge(c.stream, c.g, c.ph, out, pk, c.p_prev, c.partials, c.dst);
            if (mark) d2d(c, v.shift, &c.dst->shift, sizeof(double));
            plast = pk;
            mode = lbm::H_GAUGED;
        } else {
            mode = lbm::H_STREAMED;
        }
        in = out;
        if (mark) ++m;
    }
    HIPCHK(hipGetLastError());
}

// C(f + d, d), saturating: the blocks d recompute passes reverse with f free
// snapshots.
uint64_t revolve_capacity(uint64_t f, uint64_t d)
{
    double r = 1.0;
    for (uint64_t i = 1; i <= d; ++i) r = r * static_cast<double>(f + i) / static_cast<double>(i);
    return r > 1e18 ? static_cast<uint64_t>(1e18) : static_cast<uint64_t>(r + 0.5);
}

// Progress report of lbm_adjoint: the stream is drained first, so the report
// follows the work it names.
void report(Ctx& c, int phase, int step, uint64_t iterations)
{
    if (!c.progress) return;
    sync(c);
    c.progress(c.progress_user, phase, step, iterations);
}

// What the forward pass kept of a solve: nothing, a snapshot at every block
// boundary, or the top level of its binomial checkpointing.
enum : char { SNAPS_NONE = 0, SNAPS_ALL = 1, SNAPS_TOP = 2 };

// One step's reverse: the context of adj_block.
struct StepRev {
    const double* p_start;
    uint64_t nit;
    int step;
    lbm_adjoint_stats_t* stats;
};

// Threads per block of the adjoint Poisson kernel, 256 or 768; which is
// faster depends on the grid's shape (as for the forward push kernel), so the
// first reverse iterations of a context's first adjoint time both (three
// interleaved samples each, every sample a plain streamed iteration that would
// run anyway) and keep the faster.  The results do not depend on it.
constexpr int ADJ_BS[2] = {lbm::BLOCK, 768};
constexpr int ADJ_BS_SAMPLES = 6;

// Block j of the solve (iterations jL+1 .. min(nit, (j+1)L)) from snapshot
// `start` of check jL (nullptr for j = 0): recompute it recording every
// iterate, then run the adjoints of its iterations, last first.
void adj_block(Ctx& c, AdjWork& a, const StepRev& r, uint64_t j, double* start)
{
    hipStream_t s = c.stream;
    const Grid& g = c.g;
    const size_t n = c.n();
    const uint64_t ci = static_cast<uint64_t>(c.check_interval);
    const uint64_t nit = r.nit;
    const uint64_t a0 = j * a.L, a1 = std::min(nit, a0 + a.L);
    report(c, LBM_PROGRESS_REVERSE_BLOCK, r.step, a0);
    adj_forward_block(c, a, r.p_start, start, a0, a1);
    r.stats->recomputed_iterations += a1 - a0;
    if (a1 == nit) {
        // correct_uvw reads the solve's final (gauged) p.
        const double* p_final = a.pstore + ((nit - a0) / ci) * n;
        lbm::adj::correct(s, g, c.ph, p_final, D<double>(c.rho), a.lu, a.lv, a.lw, a.lp, a.lrho, a.lpf,
                          a.parts, a.as);
    }
    const SnapView sv = start ? snap_view(c, start) : SnapView{nullptr, nullptr, nullptr};
    for (uint64_t k = a1; k > a0; --k) {
        const int src = k == nit ? lbm::ADJ_SRC_INIT : (k % ci == 0 ? lbm::ADJ_SRC_LAZY : lbm::ADJ_SRC_PLAIN);
        const int mode = k == 1 ? lbm::H_EQUILIBRIUM
                                : ((k - 1) % ci == 0 ? lbm::H_GAUGED : lbm::H_STREAMED);
        const double* raw = k >= a0 + 2 ? a.slot(c, k - 1 - a0) : sv.h;
        const double* pin = nullptr;
        const double* shift = nullptr;
        if (mode == lbm::H_EQUILIBRIUM) {
            pin = r.p_start;
        } else if (mode == lbm::H_GAUGED) {
            pin = k - 1 == a0 ? sv.p : a.pstore + ((k - 1 - a0) / ci) * n;
            shift = k - 1 == a0 ? sv.shift : a.shifts + (k - 1 - a0) / ci;
        }
        int& bs = c.adj_bs[c.adj_state_f32 ? 1 : 0];
        const bool sample = bs < 0 && src == lbm::ADJ_SRC_PLAIN && mode == lbm::H_STREAMED;
        const int cand = a.bs_samples % 2;
        if (sample) HIPCHK(hipEventRecord(c.ev[10], s));
        lbm::adj::poisson(s, g, c.ph, src, mode, c.adj_state_f32, a.Lb[a.lcur], a.lpf, a.as, raw, pin,
                          shift, D<double>(c.tau), a.lS, a.ltau, k == 1 ? nullptr : a.Lb[1 - a.lcur],
                          k == 1 ? a.lps : nullptr, a.parts, sample ? ADJ_BS[cand] : (bs < 0 ? lbm::BLOCK : bs));
        if (sample) {
            HIPCHK(hipEventRecord(c.ev[11], s));
            HIPCHK(hipEventSynchronize(c.ev[11]));
            a.bs_best[cand] = std::min(a.bs_best[cand], elapsed(c.ev[10], c.ev[11]));
            if (++a.bs_samples == ADJ_BS_SAMPLES) bs = ADJ_BS[a.bs_best[1] < a.bs_best[0] ? 1 : 0];
        }
        if (k > 1) a.lcur = 1 - a.lcur;
        if ((k - 1) % ci == 0)  // a check interval of the solve reversed
            lbm::adj::converge(s, g, a.lS, a.lS_prev, nit - k + 1, c.tol, k + ci - 1 == nit ? 1 : 0, a.parts,
                               a.as);
    }
    HIPCHK(hipGetLastError());
}

// Segment boundaries b[0] = lo < ... < b[s] = hi of blocks [lo, hi) with F
// free snapshots: the fewest passes d with C(F + d, d) >= hi - lo, then
// segment i of length C(F - i + d - 1, d - 1) (what d - 1 passes reverse with
// the F - i snapshots left to it), greedily from the front.
std::vector<uint64_t> revolve_plan(uint64_t lo, uint64_t hi, uint64_t F)
{
    uint64_t d = 1;
    while (revolve_capacity(F, d) < hi - lo) ++d;
    std::vector<uint64_t> b{lo};
    for (uint64_t i = 0; b.back() < hi; ++i) {
        const uint64_t len = i <= F ? revolve_capacity(F - i, d - 1) : hi - b.back();
        b.push_back(std::min(hi, b.back() + len));
    }
    return b;
}

void adj_reverse_range(Ctx& c, AdjWork& a, const StepRev& r, uint64_t lo, uint64_t hi, double* start,
                       size_t fb);

// The segments of plan b, whose starts b[1..] sit in pool snapshots fb..,
// last first; segment i reuses the snapshots after its own.
void adj_reverse_segments(Ctx& c, AdjWork& a, const StepRev& r, const std::vector<uint64_t>& b, double* start,
                          size_t fb)
{
    for (size_t i = b.size() - 1; i-- > 0;)
        adj_reverse_range(c, a, r, b[i], b[i + 1], i == 0 ? start : a.pool(fb + i - 1), fb + i);
}

// Reverse blocks [lo, hi) from snapshot `start` of block lo's start, with
// pool snapshots fb.. free: snapshot the starts of revolve_plan's segments in
// one forward pass, then reverse the segments.  No free snapshot: each
// block's start is recomputed from `start` into the scratch snapshot.
void adj_reverse_range(Ctx& c, AdjWork& a, const StepRev& r, uint64_t lo, uint64_t hi, double* start,
                       size_t fb)
{
    const uint64_t L = a.L;
    if (hi - lo == 1) {
        adj_block(c, a, r, lo, start);
        return;
    }
    const uint64_t F = fb < a.snap_cap ? a.snap_cap - fb : 0;
    if (F == 0) {
        for (uint64_t j = hi; j-- > lo;) {
            double* sj = start;
            if (j > lo) {
                adj_forward_snaps(c, a, r.p_start, start, lo * L, {{j * L, a.scratch_snap}});
                r.stats->recomputed_iterations += (j - lo) * L;
                sj = a.scratch_snap;
            }
            adj_block(c, a, r, j, sj);
        }
        return;
    }
    const std::vector<uint64_t> b = revolve_plan(lo, hi, F);
    const size_t segs = b.size() - 1;
    std::vector<std::pair<uint64_t, double*>> marks;
    for (size_t i = 1; i < segs; ++i) marks.emplace_back(b[i] * L, a.pool(fb + i - 1));
    adj_forward_snaps(c, a, r.p_start, start, lo * L, marks);
    r.stats->recomputed_iterations += (b[segs - 1] - lo) * L;
    adj_reverse_segments(c, a, r, b, start, fb);
}

// Reverse of one step whose start state is st (already restored into c).
// snaps: the forward pass's snapshots of this solve, if it recorded them.
// reverse_solve = false for the last step: J reads phi_T, which getMacro forms
// before that step's Poisson solve, so every adjoint entering the solve and
// correct_uvw is zero (lambda_p, lambda_u/v/w start at zero) and their
// transposes would only propagate zeros.  Then only the correct_uvw adjoint
// runs (it resets the gauge adjoint state), lambda_p_start is zero, and the
// solve is neither recomputed nor reversed.
void adj_step(Ctx& c, AdjWork& a, const StepState& st, const std::vector<double*>& snaps, int have_snaps,
              size_t pool_free, int step, uint64_t nit, int ccur, lbm_adjoint_stats_t& stats,
              bool reverse_solve)
{
    hipStream_t s = c.stream;
    const Grid& g = c.g;
    const size_t n = c.n();
    const uint64_t L = a.L;
    const double* s_fei = st.buf;
    const double* s_rho = st.buf + n;
    const double* s_u = st.buf + 2 * n;
    const double* s_v = st.buf + 3 * n;
    const double* s_w = st.buf + 4 * n;
    const double* p_start = st.buf + 5 * n;
    const bool first = st.fg == FG::First;

    step_fg(c);  // fei', rho', u', tau_h, S
    const uint64_t nblocks = (nit + L - 1) / L;
    HIPCHK(hipMemsetAsync(a.lS, 0, n * sizeof(double), s));
    HIPCHK(hipMemsetAsync(a.lS_prev, 0, n * sizeof(double), s));
    HIPCHK(hipMemsetAsync(a.ltau, 0, n * sizeof(double), s));
    a.lcur = 0;
    const StepRev r{p_start, nit, step, &stats};
    if (!reverse_solve) {
        lbm::adj::correct(s, g, c.ph, a.pfinal + static_cast<size_t>(step) * n, D<double>(c.rho), a.lu, a.lv,
                          a.lw, a.lp, a.lrho, a.lpf, a.parts, a.as);
        HIPCHK(hipMemsetAsync(a.lps, 0, n * sizeof(double), s));
        HIPCHK(hipMemsetAsync(a.conv + 2 * static_cast<size_t>(step), 0, 2 * sizeof(double), s));
    } else {
        if (have_snaps == SNAPS_ALL || nblocks == 1) {
            for (uint64_t j = nblocks; j-- > 0;) adj_block(c, a, r, j, j > 0 ? snaps[j - 1] : nullptr);
        } else if (have_snaps == SNAPS_TOP) {
            adj_reverse_segments(c, a, r, revolve_plan(0, nblocks, a.snap_cap - pool_free), nullptr, pool_free);
        } else {
            adj_reverse_range(c, a, r, 0, nblocks, nullptr, pool_free);
        }
        lbm::adj::converge_record(s, a.as, a.conv, step);
        stats.adjoint_iterations += nit;
    }
    lbm::adj::source(s, g, c.ph, D<double>(c.fei), D<double>(c.rho), a.ltau, a.lS, a.lfei, a.as, a.lrho,
                     a.lps, a.lmac, a.parts);
    lbm::adj::collide(s, first ? 1 : 0, g, c.ph, s_fei, s_rho, s_u, s_v, s_w, a.cf[ccur], a.cg[ccur], a.lmac,
                      a.lu, a.lv, a.lw, a.lS, a.as, a.cf[1 - ccur], a.cg[1 - ccur], a.pc, a.parts);
    lbm::adj::stencil(s, g, a.pc, a.lfei, a.lrho, a.lu, a.lv, a.lw);
    HIPCHK(hipGetLastError());
    std::swap(a.lp, a.lps);
}

// Settings lbm_adjoint changes while it runs, put back on every exit: the
// caller's iteration schedule and its position, event timing, the check
// observer (the adjoint's recomputations are not steps of the simulation);
// and, unless the forward pass completed, the performance totals and the
// step history.
struct AdjointGuard {
    Ctx& c;
    std::vector<uint64_t> schedule;
    size_t sched_pos;
    bool timing;
    lbm_check_fn observer;
    void* observer_user;
    lbm_perf_t perf;
    std::vector<uint64_t> history;

    explicit AdjointGuard(Ctx& ctx)
        : c(ctx), schedule(ctx.schedule), sched_pos(ctx.sched_pos), timing(ctx.timing),
          observer(ctx.check_observer), observer_user(ctx.check_user), perf(ctx.perf),
          history(ctx.history)
    {
        c.check_observer = nullptr;
    }
    ~AdjointGuard()
    {
        c.schedule = std::move(schedule);
        c.sched_pos = sched_pos;
        c.timing = timing;
        c.check_observer = observer;
        c.check_user = observer_user;
        c.perf = perf;
        c.history = std::move(history);
    }
    AdjointGuard(const AdjointGuard&) = delete;
    AdjointGuard& operator=(const AdjointGuard&) = delete;
};

void run_adjoint(Ctx& c, const double* phi0, const double* target, int T, int K, const uint64_t* iters,
                 double* J, double* grad_phi0, double* grad_kg)
{
    if (c.mixed) throw ApiError{LBM_ERR_ARG, "the adjoint runs on an FP64 context"};
    const bool misfit = c.adj_objective == lbm::ADJ_OBJ_MISFIT;
    if (!phi0 || (misfit && !target) || !J || T < 1 || K < 1)
        throw ApiError{LBM_ERR_ARG, "adjoint needs phi0, target, objective, steps >= 1 and checkpoint interval >= 1"};
    const uint64_t ci = static_cast<uint64_t>(c.check_interval);
    uint64_t sched_max = 0;
    if (iters)
        for (int k = 0; k < T; ++k) {
            if (iters[k] == 0 || iters[k] % ci != 0)
                throw ApiError{LBM_ERR_ARG, "scheduled iteration counts must be positive multiples of the check interval"};
            if (k + 1 < T) sched_max = std::max(sched_max, iters[k]);  // the solves the reverse sweep replays
        }
    hipStream_t s = c.stream;
    const Grid& g = c.g;
    const size_t n = c.n();
    const lbm_params_t& P = c.P;
    lbm_adjoint_stats_t stats{};
    stats.steps = T;
    stats.checkpoint_interval = K;
    AdjWork a;

    // Everything is allocated before the forward pass.
    const int nseg = (T + K - 1) / K;
    a.ckpt.resize(static_cast<size_t>(nseg));
    for (auto& st : a.ckpt) alloc_state(c, st);
    alloc_state(c, a.final_state);
    a.seg.resize(static_cast<size_t>(std::min(K, T)));
    for (auto& st : a.seg) alloc_state(c, st);
    a.pfinal = dalloc(static_cast<size_t>(T) * n);
    for (double** p : {&a.lfei, &a.lrho, &a.lu, &a.lv, &a.lw, &a.lp, &a.lps, &a.lpf, &a.lS, &a.ltau, &a.lmac, &a.target,
                       &a.lS_prev})
        *p = dalloc(n);
    a.conv = dalloc(2 * static_cast<size_t>(T));
    for (int i = 0; i < 2; ++i) {
        a.Lb[i] = dalloc_bytes(lbm::Q * n * (c.adj_state_f32 ? sizeof(float) : sizeof(double)));
        a.cf[i] = dalloc(lbm::Q * n);
        a.cg[i] = dalloc(lbm::Q * n);
    }
    a.pc = dalloc(lbm::adj::PC_SLOTS * n);
    a.parts = dalloc(2 * static_cast<size_t>(lbm::grid_blocks(g.n)));
    a.as = static_cast<lbm::AdjState*>(dalloc_bytes(sizeof(lbm::AdjState)));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&a.has), sizeof(lbm::AdjState), hipHostMallocDefault));

    // Iterate memory: three quarters for the block's recorded iterates (and
    // its check store), one quarter for the snapshot pool.
    size_t budget = static_cast<size_t>(std::max<int64_t>(c.adj_store_bytes, 0));
    if (budget == 0) {
        size_t free_b = 0, total_b = 0;
        HIPCHK(hipMemGetInfo(&free_b, &total_b));
        const size_t reserve = size_t(2) << 30;
        budget = free_b > reserve ? (free_b - reserve) / 10 * 9 : 0;
    }
    const size_t slot_d = lbm::Q * n;
    a.snap_stride = ((lbm::Q + 1) * n + 1 + 31) / 32 * 32;
    const size_t total_d = budget / sizeof(double);
    auto block_mem = [&](uint64_t l) { return std::max<uint64_t>(l, 2) * slot_d + (l / ci + 1) * (n + 1); };
    size_t pool_d = total_d / 4;
    size_t slots_d = total_d - pool_d;
    uint64_t L = (static_cast<uint64_t>(slots_d / (slot_d + n / ci + 1)) / ci) * ci;
    if (sched_max) {
        // Known solve lengths: the block length whose split of the budget
        // reverses the longest solve in the fewest recompute passes (the
        // longest block among equals).
        uint64_t best = 0, best_passes = ~uint64_t(0);
        for (uint64_t l = ci; l < sched_max + ci; l += ci) {
            const size_t mem = block_mem(l);
            if (mem + 2 * a.snap_stride > total_d) break;
            const uint64_t F = (total_d - mem - a.snap_stride) / a.snap_stride;
            const uint64_t nb = (sched_max + l - 1) / l;
            uint64_t passes = 1;
            if (nb >= 2 && nb - 1 > F) {
                passes = F == 0 ? nb : 1;
                while (F > 0 && revolve_capacity(F, passes) < nb) ++passes;
            }
            if (passes <= best_passes) {
                best_passes = passes;
                best = l;
            }
        }
        L = best;
        slots_d = L ? block_mem(L) : 0;
        pool_d = total_d - slots_d;
    }
    while (L >= ci && block_mem(L) > slots_d) L -= ci;
    if (L < ci) throw ApiError{LBM_ERR_NOMEM, "adjoint needs device memory for check_interval recorded Poisson iterates"};
    a.L = L;
    stats.store_slots = static_cast<int64_t>(L);
    a.slots = dalloc(std::max<uint64_t>(L, 2) * slot_d);
    a.pstore = dalloc((L / ci + 1) * n);
    a.shifts = dalloc(L / ci + 1);
    a.scratch_snap = dalloc(a.snap_stride);
    a.snap_cap = pool_d > a.snap_stride ? (pool_d - a.snap_stride) / a.snap_stride : 0;
    if (a.snap_cap) a.snap_pool = dalloc(a.snap_cap * a.snap_stride);
    a.step_snaps.resize(static_cast<size_t>(T));
    a.have_snaps.assign(static_cast<size_t>(T), 0);
    a.pool_mark.assign(static_cast<size_t>(T), 0);

    AdjointGuard guard(c);

    // Initial state: rho0 linear in phi0 (place_droplet's rule), rest zero.
    HostFields h0;
    h0.fei.assign(phi0, phi0 + n);
    h0.rho.resize(n);
    for (size_t i = 0; i < n; ++i)
        h0.rho[i] = P.rho_G + (phi0[i] - P.fei_min) * (P.rho_L - P.rho_G) / (P.fei_max - P.fei_min);
    h0.u.assign(n, 0.0);
    h0.v.assign(n, 0.0);
    h0.w.assign(n, 0.0);
    h0.p.assign(n, 0.0);
    upload_host(c, h0);
    c.hs = HState::Equilibrium;  // h = E p = 0: the h = 0 of a fresh solver
    if (iters) c.schedule.assign(iters, iters + T);
    else c.schedule.clear();
    c.sched_pos = 0;

    // Forward pass.
    sync(c);
    auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < T; ++k) {
        const size_t ks = static_cast<size_t>(k);
        if (k % K == 0) save_state(c, a.ckpt[static_cast<size_t>(k / K)]);
        const size_t pool_mark = a.snap_next;
        a.pool_mark[ks] = pool_mark;
        SnapRecorder rec{&a, &a.step_snaps[ks]};
        const bool reversed = k + 1 < T;  // the last step's solve is not reversed (adj_step)
        if (iters && reversed) {
            // a scheduled solve whose block snapshots do not fit keeps the
            // top level of its binomial checkpointing instead
            const uint64_t nb = (iters[k] + L - 1) / L;
            const uint64_t F = a.snap_cap - pool_mark;
            if (F > 0 && nb >= 2 && nb - 1 > F) {
                const std::vector<uint64_t> b = revolve_plan(0, nb, F);
                for (size_t i = 1; i + 1 < b.size(); ++i) rec.marks.push_back(b[i] * L);
            }
        }
        const CheckHook hook{record_snapshot, &rec};
        step_one(c, reversed ? &hook : nullptr);
        const uint64_t nit = c.history.back();
        report(c, LBM_PROGRESS_FORWARD_STEP, k, nit);
        const size_t need = static_cast<size_t>((nit + L - 1) / L - 1);
        if (!reversed) {
            a.step_snaps[ks].clear();
            a.snap_next = pool_mark;
        } else if (!rec.marks.empty()) {
            a.snap_next = pool_mark + rec.marks.size();
            a.have_snaps[ks] = SNAPS_TOP;
        } else if (rec.overflow) {
            a.step_snaps[ks].clear();  // incomplete: this step takes its own snapshot pass
            a.snap_next = pool_mark;
        } else {
            a.step_snaps[ks].resize(need);  // drop a snapshot taken at the final check
            a.snap_next = pool_mark + need;
            a.have_snaps[ks] = SNAPS_ALL;
        }
        d2d(c, a.pfinal + ks * n, c.p[c.pc], n * sizeof(double));
    }
    save_state(c, a.final_state);
    sync(c);
    auto t1 = std::chrono::steady_clock::now();
    stats.forward_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    const std::vector<uint64_t> sched(c.history.end() - T, c.history.end());
    for (uint64_t v : sched) stats.forward_iterations += v;
    // The forward pass's steps stay on the record; the replays below are not
    // steps of the simulation.
    guard.history = c.history;
    guard.perf = c.perf;

    // Seed: J and dJ/dfei_T; every other adjoint starts at zero.
    if (misfit) HIPCHK(hipMemcpyAsync(a.target, target, n * sizeof(double), hipMemcpyHostToDevice, s));
    lbm::adj::seed(s, g, c.ph, c.adj_objective, D<double>(c.fei), a.target, a.lfei, a.parts, a.as);
    for (double* p : {a.lrho, a.lu, a.lv, a.lw, a.lp}) HIPCHK(hipMemsetAsync(p, 0, n * sizeof(double), s));
    HIPCHK(hipMemsetAsync(a.cf[0], 0, slot_d * sizeof(double), s));
    HIPCHK(hipMemsetAsync(a.cg[0], 0, slot_d * sizeof(double), s));
    HIPCHK(hipGetLastError());

    c.timing = false;
    c.schedule = sched;
    int ccur = 0;
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int k0 = sg * K, k1 = std::min(T, k0 + K);
        restore_state(c, a.ckpt[static_cast<size_t>(sg)]);
        for (int k = k0; k < k1; ++k) {
            save_state(c, a.seg[static_cast<size_t>(k - k0)]);
            if (k + 1 < k1) step_light(c, a.pfinal + static_cast<size_t>(k) * n);
        }
        for (int k = k1 - 1; k >= k0; --k) {
            const size_t ks = static_cast<size_t>(k);
            const StepState& st = a.seg[static_cast<size_t>(k - k0)];
            restore_state(c, st);
            adj_step(c, a, st, a.step_snaps[ks], a.have_snaps[ks], a.pool_mark[ks], k, sched[ks], ccur, stats,
                     k + 1 < T);
            report(c, LBM_PROGRESS_REVERSE_STEP, k, sched[ks]);
            ccur = 1 - ccur;
        }
    }
    HIPCHK(hipMemcpyAsync(a.has, a.as, sizeof(lbm::AdjState), hipMemcpyDeviceToHost, s));
    std::vector<double> lf(n), lr(n), conv(2 * static_cast<size_t>(T));
    HIPCHK(hipMemcpyAsync(lf.data(), a.lfei, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(lr.data(), a.lrho, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(conv.data(), a.conv, conv.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    sync(c);
    auto t2 = std::chrono::steady_clock::now();
    stats.reverse_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    // the reverse solves' convergence: the worst step (a solve of one check
    // interval has no earlier interval to compare with)
    for (int k = 0; k < T; ++k) {
        if (sched[static_cast<size_t>(k)] <= ci) continue;
        const double change = conv[2 * static_cast<size_t>(k)], reached = conv[2 * static_cast<size_t>(k) + 1];
        stats.final_relative_change = std::max(stats.final_relative_change, change);
        if (reached < 0.0 || stats.iterations_to_tolerance < 0)
            stats.iterations_to_tolerance = -1;
        else
            stats.iterations_to_tolerance = std::max(stats.iterations_to_tolerance, static_cast<int64_t>(reached));
    }
    // phi0 enters as fei and through rho0.
    const double scale = (P.rho_L - P.rho_G) / (P.fei_max - P.fei_min);
    if (grad_phi0)
        for (size_t i = 0; i < n; ++i) grad_phi0[i] = lf[i] + lr[i] * scale;
    *J = a.has->J;
    if (grad_kg) *grad_kg = a.has->kg;

    // Leave the context at the end of the forward pass (the guard puts back
    // the settings).
    restore_state(c, a.final_state);
    sync(c);
    c.adj_stats = stats;
}

} // namespace

namespace lbm {
void set_global_error(const std::string& msg) { g_create_error = msg; }
} // namespace lbm

extern "C" {

int lbm_create(const lbm_params_t* params, int precision, void** out_ctx)
{
    if (!params || !out_ctx) {
        g_create_error = "null argument";
        return LBM_ERR_ARG;
    }
    *out_ctx = nullptr;
    try {
        *out_ctx = create(*params, precision);
        g_create_error.clear();
        return LBM_OK;
    } catch (const ApiError& e) {
        g_create_error = e.msg;
        return e.code;
    } catch (const HipError& e) {
        g_create_error = std::string("HIP failure: ") + hipGetErrorString(e.code) + " in " + e.what;
        return LBM_ERR_HIP;
    } catch (const std::bad_alloc&) {
        g_create_error = "host allocation failed";
        return LBM_ERR_NOMEM;
    }
}

int lbm_destroy(void* ctx)
{
    if (!ctx) return LBM_ERR_ARG;
    Ctx* c = static_cast<Ctx*>(ctx);
    free_all(*c);
    delete c;
    return LBM_OK;
}

const char* lbm_last_error(const void* ctx)
{
    if (!ctx) return g_create_error.c_str();
    return static_cast<const Ctx*>(ctx)->err.c_str();
}

int lbm_reset_fields(void* ctx)
{
    return guarded(ctx, [&](Ctx& c) { upload_host(c, background(c)); });
}

int lbm_init_reference(void* ctx)
{
    return guarded(ctx, [&](Ctx& c) {
        // Inamuro::initializeDroplets, CPU/Inamuro.cpp:550-587.
        HostFields h = background(c);
        const double radius = c.P.DD / 2.0;
        const double cy = c.P.ly / 2.0;
        const double cz1 = c.P.lz / 2.0 + c.P.DD;
        const double cz2 = c.P.lz / 2.0 - c.P.DD;
        const double velocity = 0.035;
        for (int x = 0; x < c.P.lx; ++x)
            for (int y = 0; y < c.P.ly; ++y)
                for (int z = 0; z < c.P.lz; ++z) {
                    const double dx = x, dy = y - cy;
                    const double dz1 = z - cz1, dz2 = z - cz2;
                    const double r1 = std::sqrt(dx * dx + dy * dy + dz1 * dz1);
                    const double r2 = std::sqrt(dx * dx + dy * dy + dz2 * dz2);
                    const size_t i = hidx(c, x, y, z);
                    if (r1 < radius) {
                        h.fei[i] = c.P.fei_max;
                        h.rho[i] = c.P.rho_L;
                        h.w[i] = -velocity;
                    }
                    if (r2 < radius) {
                        h.fei[i] = c.P.fei_max;
                        h.rho[i] = c.P.rho_L;
                        h.w[i] = velocity;
                    }
                }
        upload_host(c, h);
    });
}

int lbm_place_droplet(void* ctx, double cx, double cy, double cz, double vx, double vy, double vz,
                      int profile, double interface_width)
{
    return guarded(ctx, [&](Ctx& c) {
        // place_droplet, GPU/Inamuro.cpp:499-533.
        HostFields h = download_host(c);
        const double radius = c.P.DD / 2.0;
        const double width = interface_width > 0.0 ? interface_width : 2.0;
        for (int x = 0; x < c.P.lx; ++x)
            for (int y = 0; y < c.P.ly; ++y)
                for (int z = 0; z < c.P.lz; ++z) {
                    const double dx = x - cx;
                    const double dy = y - cy;
                    const double dz = z - cz;
                    const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                    double lf = 0.0;
                    if (profile == 1)
                        lf = 0.5 * (1.0 - std::tanh((r - radius) / width));
                    else if (r < radius)
                        lf = 1.0;
                    if (lf > 0.0) {
                        const double local_fei = c.P.fei_min + lf * (c.P.fei_max - c.P.fei_min);
                        const size_t i = hidx(c, x, y, z);
                        if (local_fei > h.fei[i]) {
                            h.fei[i] = local_fei;
                            h.rho[i] = c.P.rho_G + lf * (c.P.rho_L - c.P.rho_G);
                            if (vx != 0.0) h.u[i] = vx * lf;
                            if (vy != 0.0) h.v[i] = vy * lf;
                            h.w[i] = vz * lf;
                        }
                    }
                }
        upload_host(c, h);
    });
}

int lbm_upload(void* ctx, const double* fei, const double* rho, const double* u, const double* v,
               const double* w, const double* p)
{
    return guarded(ctx, [&](Ctx& c) {
        reset_state(c);
        put_field(c, c.fei, fei);
        put_field(c, c.rho, rho);
        put_field(c, c.u, u);
        put_field(c, c.v, v);
        put_field(c, c.w, w);
        put_p(c, p);
        sync(c);
    });
}

int lbm_upload_dists(void* ctx, const double* f, const double* g, const double* h, int first_step)
{
    return guarded(ctx, [&](Ctx& c) {
        const size_t n = c.n();
        const size_t bytes = lbm::Q * n * sizeof(double);
        if (!c.mixed) {
            if (f) HIPCHK(hipMemcpyAsync(c.f[c.fcur], f, bytes, hipMemcpyHostToDevice, c.stream));
            if (g) HIPCHK(hipMemcpyAsync(c.gd[c.fcur], g, bytes, hipMemcpyHostToDevice, c.stream));
            if (h) HIPCHK(hipMemcpyAsync(c.h[c.hcur], h, bytes, hipMemcpyHostToDevice, c.stream));
        } else {
            ensure_scratch(c);
            auto enc = [&](const double* src, int is_f, void* dst) {
                HIPCHK(hipMemcpyAsync(c.scratch, src, bytes, hipMemcpyHostToDevice, c.stream));
                lbm::mx::fg_encode(c.stream, c.g, is_f, c.scratch, D<float>(c.fei), D<float>(dst));
                HIPCHK(hipGetLastError());
                sync(c);
            };
            if (f) enc(f, 1, c.f[c.fcur]);
            if (g) enc(g, 0, c.gd[c.fcur]);
            if (h) {
                HIPCHK(hipMemcpyAsync(c.scratch, h, bytes, hipMemcpyHostToDevice, c.stream));
                lbm::mx::to_f32(c.stream, static_cast<int64_t>(lbm::Q * n), c.scratch, D<float>(c.h[c.hcur]));
                HIPCHK(hipGetLastError());
            }
        }
        if (h) c.hs = HState::Explicit;
        if (f || g) c.fg = first_step ? FG::First : FG::Local;
        sync(c);
    });
}

int lbm_set_poisson(void* ctx, int32_t check_interval, double tolerance, uint64_t iteration_limit,
                    int dual_residual)
{
    return guarded(ctx, [&](Ctx& c) {
        // setPoissonConvergence / setPoissonIterationLimit, GPU/InamuroCUDA.cu:2582-2608.
        if (check_interval <= 0) throw ApiError{LBM_ERR_ARG, "poisson check interval must be positive"};
        if (!(tolerance > 0.0)) throw ApiError{LBM_ERR_ARG, "poisson tolerance must be positive"};
        if (iteration_limit != 0 && iteration_limit % static_cast<uint64_t>(check_interval) != 0)
            throw ApiError{LBM_ERR_ARG, "Poisson iteration limit must be a multiple of check interval"};
        if (dual_residual < 0 || dual_residual > 2)
            throw ApiError{LBM_ERR_ARG, "dual_residual must be 0 (off), 1 (dual exit) or 2 (shadow)"};
        if (dual_residual != 0 && c.g.zwall)
            throw ApiError{LBM_ERR_ARG, "the dual residual covers the reference boundaries only (no z walls)"};
        for (uint64_t it : c.schedule)
            if (it % static_cast<uint64_t>(check_interval) != 0)
                throw ApiError{LBM_ERR_ARG,
                               "the iteration schedule holds counts that are not multiples of this check interval"};
        // Every argument is valid: only now does the context change.
        c.check_interval = check_interval;
        c.tol = tolerance;
        c.iter_limit = iteration_limit;
        c.dual = dual_residual;
        clear_graphs(c);
    });
}

void lbm_poisson_accel_defaults(lbm_poisson_accel_t* out)
{
    if (!out) return;
    *out = lbm_poisson_accel_t{};
    // RunConfig / InamuroCUDA defaults, GPU/main.cpp:27-35, GPU/InamuroCUDA.hpp:152-157.
    out->variant = LBM_POISSON_SPLIT;
    out->pressure_relax_scale = 1.0;
    out->fixed_point_relax = 1.0;
    out->scalar_source_scale = 2.0;
    out->source_aware_hh_scale = 1.0;
    out->anderson_beta_max = 1.0;
    out->two_grid_strength = 0.5;
}

int lbm_set_poisson_accel(void* ctx, const lbm_poisson_accel_t* settings)
{
    return guarded(ctx, [&](Ctx& c) {
        if (!settings) throw ApiError{LBM_ERR_ARG, "settings must not be NULL"};
        const lbm_poisson_accel_t& a = *settings;
        // GPU/main.cpp:205-227 and the setters, GPU/InamuroCUDA.cu:2488-2540.
        if (a.variant < LBM_POISSON_SPLIT || a.variant > LBM_POISSON_SCALAR)
            throw ApiError{LBM_ERR_ARG, "--poisson must be split, fused, onepass, or scalar"};
        if (!(a.pressure_relax_scale > 0.0))
            throw ApiError{LBM_ERR_ARG, "pressure relaxation scale must be positive"};
        if (!(a.fixed_point_relax > 0.0))
            throw ApiError{LBM_ERR_ARG, "Poisson fixed-point relaxation must be positive"};
        if (!(a.anderson_beta_max >= 0.0))
            throw ApiError{LBM_ERR_ARG, "Poisson Anderson beta max must be non-negative"};
        if (!(a.two_grid_strength >= 0.0))
            throw ApiError{LBM_ERR_ARG, "Poisson two-grid strength must be non-negative"};
        if (!std::isfinite(a.scalar_source_scale) || !std::isfinite(a.source_aware_hh_scale) ||
            !std::isfinite(a.pressure_relax_scale) || !std::isfinite(a.fixed_point_relax) ||
            !std::isfinite(a.anderson_beta_max) || !std::isfinite(a.two_grid_strength))
            throw ApiError{LBM_ERR_ARG, "Poisson accelerator settings must be finite"};
        lbm_poisson_accel_t def{};
        lbm_poisson_accel_defaults(&def);
        const bool nondefault = std::memcmp(&a, &def, sizeof(a)) != 0;
        if (c.mixed && nondefault)
            throw ApiError{LBM_ERR_ARG,
                           "the Poisson accelerators are implemented for the FP64 path (the reference's precision)"};
        c.accel = a;
        clear_graphs(c);
    });
}

int lbm_get_poisson_accel(const void* ctx, lbm_poisson_accel_t* out)
{
    if (!ctx || !out) return LBM_ERR_ARG;
    *out = static_cast<const Ctx*>(ctx)->accel;
    return LBM_OK;
}

int lbm_set_z_walls(void* ctx, int enabled)
{
    return guarded(ctx, [&](Ctx& c) {
        if (enabled && c.P.lz < 2) throw ApiError{LBM_ERR_ARG, "z walls need lz >= 2"};
        if (enabled && c.dual != 0)
            throw ApiError{LBM_ERR_ARG, "the dual residual covers the reference boundaries only (no z walls)"};
        c.g.zwall = enabled ? 1 : 0;
        clear_graphs(c);
    });
}

int lbm_set_gravity(void* ctx, double gz)
{
    return guarded(ctx, [&](Ctx& c) {
        if (!std::isfinite(gz)) throw ApiError{LBM_ERR_ARG, "gravity must be finite"};
        c.ph.gz = gz;
        c.phf.gz = static_cast<float>(gz);
    });
}

int lbm_set_poisson_launch(void* ctx, int config)
{
    return guarded(ctx, [&](Ctx& c) {
        const int ncfg = c.mixed ? lbm::mx::PUSH_CONFIGS : lbm::f64::PUSH_CONFIGS;
        if (config < -1 || config >= ncfg)
            throw ApiError{LBM_ERR_ARG, "Poisson launch config must be -1 (automatic) or 0.." +
                                            std::to_string(ncfg - 1) + " for this precision"};
        c.push_request = config;
        c.push_cfg = config;  // -1: tune again
        c.tune_best.clear();
        c.tune_done = 0;
        c.rec_cfg = -1;
        clear_graphs(c);
    });
}

int lbm_get_poisson_launch(const void* ctx, int* config)
{
    if (!ctx || !config) return LBM_ERR_ARG;
    *config = static_cast<const Ctx*>(ctx)->push_cfg;
    return LBM_OK;
}

int lbm_set_poisson_graph(void* ctx, int enabled)
{
    return guarded(ctx, [&](Ctx& c) {
        // setUsePoissonGraph, GPU/InamuroCUDA.cu:2550-2556.
        c.use_graphs = enabled != 0;
    });
}

int lbm_set_iteration_schedule(void* ctx, const uint64_t* iterations, int64_t count)
{
    return guarded(ctx, [&](Ctx& c) {
        if (count < 0 || (count > 0 && !iterations)) throw ApiError{LBM_ERR_ARG, "bad iteration schedule"};
        for (int64_t i = 0; i < count; ++i)
            if (iterations[i] == 0 || iterations[i] % static_cast<uint64_t>(c.check_interval) != 0)
                throw ApiError{LBM_ERR_ARG, "scheduled iteration counts must be positive multiples of the check interval"};
        c.schedule.assign(iterations, iterations + count);
        c.sched_pos = 0;
    });
}

int lbm_iteration_history(const void* ctx, uint64_t* out, int64_t capacity, int64_t* count)
{
    if (!ctx || !count || capacity < 0 || (capacity > 0 && !out)) return LBM_ERR_ARG;
    const Ctx& c = *static_cast<const Ctx*>(ctx);
    *count = static_cast<int64_t>(c.history.size());
    const int64_t m = std::min<int64_t>(capacity, *count);
    for (int64_t i = 0; i < m; ++i) out[i] = c.history[static_cast<size_t>(i)];
    return LBM_OK;
}

int lbm_step(void* ctx, int32_t n)
{
    return guarded(ctx, [&](Ctx& c) {
        if (n < 0) throw ApiError{LBM_ERR_ARG, "step count must be non-negative"};
        for (int i = 0; i < n; ++i) {
            if (c.perf.steps == std::numeric_limits<int64_t>::max())
                throw ApiError{LBM_ERR_OVERFLOW, "time-step counter overflow"};
            step_one(c);
        }
    });
}

int lbm_download(void* ctx, double* fei, double* rho, double* u, double* v, double* w, double* p)
{
    return guarded(ctx, [&](Ctx& c) {
        get_field(c, fei, c.fei);
        get_field(c, rho, c.rho);
        get_field(c, u, c.u);
        get_field(c, v, c.v);
        get_field(c, w, c.w);
        if (p) HIPCHK(hipMemcpyAsync(p, c.p[c.pc], c.n() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        sync(c);
    });
}

int lbm_download_dists(void* ctx, double* f, double* g, double* h)
{
    return guarded(ctx, [&](Ctx& c) {
        const size_t bytes = lbm::Q * c.n() * sizeof(double);
        ensure_scratch(c);
        auto get_fg = [&](double* out, void* const* bufs, int is_f) {
            if (!out) return;
            const int pulled = c.fg == FG::Post;
            if (!c.mixed) {
                if (pulled) {
                    lbm::f64::materialize(c.stream, c.g, D<double>(bufs[c.fcur]), c.scratch);
                    HIPCHK(hipGetLastError());
                    HIPCHK(hipMemcpyAsync(out, c.scratch, bytes, hipMemcpyDeviceToHost, c.stream));
                } else {
                    HIPCHK(hipMemcpyAsync(out, bufs[c.fcur], bytes, hipMemcpyDeviceToHost, c.stream));
                }
            } else {
                lbm::mx::fg_full(c.stream, c.g, pulled, is_f, D<float>(bufs[c.fcur]), D<float>(c.fei),
                                 c.scratch);
                HIPCHK(hipGetLastError());
                HIPCHK(hipMemcpyAsync(out, c.scratch, bytes, hipMemcpyDeviceToHost, c.stream));
            }
            sync(c);
        };
        get_fg(f, c.f, 1);
        get_fg(g, c.gd, 0);
        if (h) {
            if (c.hs == HState::Equilibrium) {
                lbm::f64::equilibrium_h(c.stream, c.g, c.p[c.pc], c.scratch);
                HIPCHK(hipGetLastError());
                HIPCHK(hipMemcpyAsync(h, c.scratch, bytes, hipMemcpyDeviceToHost, c.stream));
            } else if (!c.mixed) {
                HIPCHK(hipMemcpyAsync(h, c.h[c.hcur], bytes, hipMemcpyDeviceToHost, c.stream));
            } else {
                lbm::mx::to_f64(c.stream, static_cast<int64_t>(lbm::Q * c.n()), D<float>(c.h[c.hcur]),
                                c.scratch);
                HIPCHK(hipGetLastError());
                HIPCHK(hipMemcpyAsync(h, c.scratch, bytes, hipMemcpyDeviceToHost, c.stream));
            }
            sync(c);
        }
    });
}

int lbm_poisson_diag(const void* ctx, lbm_poisson_diag_t* out)
{
    if (!ctx || !out) return LBM_ERR_ARG;
    *out = static_cast<const Ctx*>(ctx)->diag;
    return LBM_OK;
}

int lbm_set_check_observer(void* ctx, lbm_check_fn fn, void* user)
{
    return guarded(ctx, [&](Ctx& c) {
        c.check_observer = fn;
        c.check_user = user;
    });
}

int lbm_synchronize(void* ctx)
{
    return guarded(ctx, [&](Ctx& c) { sync(c); });
}

int lbm_set_timing(void* ctx, int enabled)
{
    return guarded(ctx, [&](Ctx& c) { c.timing = enabled != 0; });
}

int lbm_perf_get(const void* ctx, lbm_perf_t* out)
{
    if (!ctx || !out) return LBM_ERR_ARG;
    *out = static_cast<const Ctx*>(ctx)->perf;
    return LBM_OK;
}

int lbm_perf_reset(void* ctx)
{
    return guarded(ctx, [&](Ctx& c) { c.perf = lbm_perf_t{}; });
}

int lbm_set_adjoint_memory(void* ctx, int64_t iterate_bytes)
{
    return guarded(ctx, [&](Ctx& c) {
        if (iterate_bytes < 0) throw ApiError{LBM_ERR_ARG, "iterate memory must be non-negative"};
        c.adj_store_bytes = iterate_bytes;
    });
}

int lbm_adjoint(void* ctx, const double* phi0, const double* target, int32_t steps,
                int32_t checkpoint_interval, const uint64_t* iterations, double* objective,
                double* grad_phi0, double* grad_kg)
{
    return guarded(ctx, [&](Ctx& c) {
        if (accel_on(c) || accel_requested(c))
            throw ApiError{LBM_ERR_STATE,
                           "the adjoint differentiates the strict split Poisson map; turn the Poisson accelerators off"};
        run_adjoint(c, phi0, target, steps, checkpoint_interval, iterations, objective, grad_phi0, grad_kg);
    });
}

int lbm_set_adjoint_objective(void* ctx, int32_t kind)
{
    return guarded(ctx, [&](Ctx& c) {
        if (kind != LBM_OBJ_MISFIT && kind != LBM_OBJ_CENTROID_Z)
            throw ApiError{LBM_ERR_ARG, "unknown adjoint objective"};
        c.adj_objective = kind == LBM_OBJ_CENTROID_Z ? lbm::ADJ_OBJ_CENTROID_Z : lbm::ADJ_OBJ_MISFIT;
    });
}

int lbm_set_progress(void* ctx, lbm_progress_fn fn, void* user)
{
    return guarded(ctx, [&](Ctx& c) {
        c.progress = fn;
        c.progress_user = user;
    });
}

int lbm_set_adjoint_state_precision(void* ctx, int32_t precision)
{
    return guarded(ctx, [&](Ctx& c) {
        if (precision != LBM_ADJ_STATE_FP64 && precision != LBM_ADJ_STATE_FP32)
            throw ApiError{LBM_ERR_ARG, "unknown adjoint state precision"};
        c.adj_state_f32 = precision == LBM_ADJ_STATE_FP32;
    });
}

int lbm_adjoint_stats(const void* ctx, lbm_adjoint_stats_t* out)
{
    if (!ctx || !out) return LBM_ERR_ARG;
    *out = static_cast<const Ctx*>(ctx)->adj_stats;
    return LBM_OK;
}

int lbm_mass(void* ctx, double* sum_phi)
{
    return guarded(ctx, [&](Ctx& c) {
        if (!sum_phi) throw ApiError{LBM_ERR_ARG, "null argument"};
        if (!c.mixed) {
            lbm::f64::sum_field(c.stream, c.g, D<double>(c.fei), c.partials, c.dst);
        } else {
            lbm::mx::sum_f32(c.stream, c.g, D<float>(c.fei), c.partials);
            lbm::f64::mass_finalize(c.stream, c.partials, lbm::red_blocks(c.g.n), c.dst);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(c.hst, c.dst, sizeof(DevState), hipMemcpyDeviceToHost, c.stream));
        sync(c);
        *sum_phi = c.hst->mass;
    });
}

} // extern "C"
