// InamuroHIP.cpp -- C++ host side over the C ABI (include/lbm_hip.h).
#include "InamuroHIP.hpp"

#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace lbm {

namespace {

// Status code -> the exception type the reference throws for it.
void check(int rc, const void* ctx)
{
    if (rc == LBM_OK) return;
    const std::string msg = lbm_last_error(ctx);
    if (rc == LBM_ERR_ARG) throw std::invalid_argument(msg);
    if (rc == LBM_ERR_OVERFLOW) throw std::overflow_error(msg);
    throw std::runtime_error(msg);
}

std::string trim(const std::string& t)
{
    const auto a = t.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const auto b = t.find_last_not_of(" \t\r\n");
    return t.substr(a, b - a + 1);
}

} // namespace

// ---------------------------------------------------------------- InamuroHIP

InamuroHIP::InamuroHIP(const lbm_params_t& params, int precision)
{
    check(lbm_create(&params, precision, &ctx), nullptr);
    n = static_cast<std::size_t>(params.lx) * params.ly * params.lz;
    precision_ = precision;
    lbm_poisson_accel_defaults(&accel_);
}

InamuroHIP::~InamuroHIP()
{
    if (ctx) lbm_destroy(ctx);
}

void InamuroHIP::performTimeStepGPU()
{
    // GPU/InamuroCUDA.cu:1857-1865 (the accelerator refusal just before it
    // is the library's own).
    if (dual_ == 1 && (interval_ != 100 || tol_ != 1.0e-3 || limit_ != 0 || !deterministic_ ||
                       accel_.spatial_diagnostics))
        throw std::runtime_error(
            "dual-residual exit requires check_interval=100, tolerance=1e-3, poisson_iteration_limit=0, "
            "deterministic reductions, no spatial diagnostics");
    check(lbm_step(ctx, 1), ctx);
}
void InamuroHIP::synchronize() const { check(lbm_synchronize(ctx), ctx); }

void InamuroHIP::apply() { check(lbm_set_poisson(ctx, interval_, tol_, limit_, dual_), ctx); }

void InamuroHIP::setPoissonConvergence(int check_interval, double tolerance)
{
    const int oi = interval_;
    const double ot = tol_;
    interval_ = check_interval;
    tol_ = tolerance;
    try {
        apply();
    } catch (...) {
        interval_ = oi;
        tol_ = ot;
        throw;
    }
}

void InamuroHIP::setPoissonIterationLimit(std::uint64_t max_iterations)
{
    const std::uint64_t ol = limit_;
    limit_ = max_iterations;
    try {
        apply();
    } catch (...) {
        limit_ = ol;
        throw;
    }
}

void InamuroHIP::setUsePoissonDualResidual(bool enabled)
{
    dual_ = enabled ? 1 : (dual_ == 1 ? 0 : dual_);
    apply();
    if (enabled) deterministic_ = true;  // GPU/InamuroCUDA.cu:2573-2580
}

void InamuroHIP::setEnablePoissonFixedPointShadow(bool enabled)
{
    if (dual_ != 1) dual_ = enabled ? 2 : 0;
    apply();
}

void InamuroHIP::setUsePoissonGraph(bool enabled)
{
    check(lbm_set_poisson_graph(ctx, enabled ? 1 : 0), ctx);
    graph_ = enabled;
}

void InamuroHIP::setEnablePoissonDetailTiming(bool enabled)
{
    if (enabled) setTiming(true);
    detail_ = enabled;
}

void InamuroHIP::setUsePoissonDeterministicReductions(bool enabled) { deterministic_ = enabled; }

void InamuroHIP::resetPerformanceMetrics() { check(lbm_perf_reset(ctx), ctx); }

namespace {

const char* variant_name(int v)
{
    return v == LBM_POISSON_SCALAR ? "scalar" : v == LBM_POISSON_ONEPASS ? "onepass"
         : v == LBM_POISSON_FUSED ? "fused" : "split";
}

} // namespace

void InamuroHIP::printPerformanceMetrics(std::ostream& os) const
{
    const lbm_perf_t p = perf();
    if (p.steps <= 0) {
        os << "performance: no steps since the last reset" << std::endl;
        return;
    }
    const double steps = static_cast<double>(p.steps);
    const auto on = [](bool b) { return b ? "on" : "off"; };
    os << "---- MI355X performance over " << p.steps << " steps ----\n"
       << "precision " << (precision_ == LBM_PRECISION_MIXED ? "mixed" : "fp64") << ", Poisson "
       << variant_name(accel_.variant) << ", pressure boundary " << (accel_.fused_boundary ? "fused" : "split")
       << ", graph replay " << on(graph_) << "\n"
       << "convergence: check every " << interval_ << " iterations, tolerance " << tol_ << ", iteration limit "
       << limit_ << (limit_ == 0 ? " (unbounded)" : "") << "\n"
       << "kinetic fixed-point residual: " << (dual_ == 1 ? "dual exit" : dual_ == 2 ? "evaluated" : "off")
       << "; reductions: fixed-shape trees (deterministic)\n"
       << "accelerators: relaxation scale " << accel_.pressure_relax_scale << ", fixed-point relaxation "
       << accel_.fixed_point_relax << ", Anderson(1) " << on(accel_.anderson_m1 != 0) << " (beta max "
       << accel_.anderson_beta_max << "), two-grid " << on(accel_.two_grid != 0) << " (strength "
       << accel_.two_grid_strength << "), source-aware start " << on(accel_.source_aware_hh_init != 0)
       << " (scale " << accel_.source_aware_hh_scale << "), scalar source scale " << accel_.scalar_source_scale
       << ", spatial diagnostics " << on(accel_.spatial_diagnostics != 0) << "\n";
    const double total = p.collide_ms + p.macro_ms + p.poisson_ms + p.check_ms + p.correct_ms;
    if (!(total > 0.0)) {
        // the steps ran with event timing off (setTiming / setEnablePoissonDetailTiming)
        os << "Poisson iterations per step " << static_cast<double>(p.poisson_iterations) / steps
           << "; no phase times (event timing was off)" << std::endl;
        return;
    }
    os << "ms per step: collide+stream " << p.collide_ms / steps << ", macro " << p.macro_ms / steps
       << ", Poisson iterations " << p.poisson_ms / steps << ", checks " << p.check_ms / steps << ", correct "
       << p.correct_ms / steps << ", total " << total / steps << "\n"
       << "Poisson iterations per step " << static_cast<double>(p.poisson_iterations) / steps << ", MLUPS "
       << static_cast<double>(n) / (total / steps * 1000.0) << "\n";
    if (detail_ && p.poisson_iter_launches > 0 && p.poisson_checks > 0)
        os << "Poisson detail: " << p.poisson_ms * 1000.0 / static_cast<double>(p.poisson_iter_launches)
           << " us per iteration, " << p.check_ms * 1000.0 / static_cast<double>(p.poisson_checks)
           << " us per convergence check (" << static_cast<double>(p.poisson_checks) / steps << " per step)\n";
    if (total > 0.0)
        os << "share of step time: collide+stream " << 100.0 * p.collide_ms / total << "%, macro "
           << 100.0 * p.macro_ms / total << "%, Poisson " << 100.0 * p.poisson_ms / total << "%, checks "
           << 100.0 * p.check_ms / total << "%, correct " << 100.0 * p.correct_ms / total << "%\n";
    os << std::flush;
}

void InamuroHIP::printRooflineSummary(std::ostream& os) const
{
    // Algorithmic bytes per cell: the step outside the solve (f and g read
    // and written once, the macro fields, the source, the correction), one
    // Poisson iteration (h read and written, p, tau_h and the source; the
    // scalar sweep moves p, rho and the source only; an Anderson(1)
    // iteration also reads the previous residual in its map and reads three
    // and writes two distributions to mix), and one convergence check.
    const bool mixed = precision_ == LBM_PRECISION_MIXED;
    const double step_b = mixed ? 320.0 : 632.0;
    double iter_b = mixed ? 144.0 : 272.0;
    if (accel_.variant == LBM_POISSON_SCALAR)
        iter_b = 32.0;
    else if (accel_.variant == LBM_POISSON_ONEPASS && accel_.anderson_m1)
        iter_b = 992.0;
    const double cells = static_cast<double>(n);
    os << "---- byte model, " << n << " cells ----\n"
       << "step outside the Poisson solve: " << static_cast<int>(step_b) << " B/cell (" << step_b * cells / 1e9
       << " GB)\n"
       << "Poisson iteration: " << static_cast<int>(iter_b) << " B/cell (" << iter_b * cells / 1e9 << " GB)\n"
       << "convergence check: 24 B/cell\n";
    const lbm_perf_t p = perf();
    if (p.poisson_iter_launches > 0 && p.poisson_ms > 0.0) {
        const double ms = p.poisson_ms / static_cast<double>(p.poisson_iter_launches);
        const double gbs = iter_b * cells / (ms * 1e-3) / 1e9;
        os << "measured Poisson iteration: " << ms << " ms, " << gbs << " GB/s = " << 100.0 * gbs / 8000.0
           << "% of 8 TB/s\n";
    }
    os << std::flush;
}

// The whole accelerator set goes through the library, which validates it
// with the reference setters' messages; only an accepted set is kept.
void InamuroHIP::applyAccel(const lbm_poisson_accel_t& next)
{
    check(lbm_set_poisson_accel(ctx, &next), ctx);
    accel_ = next;
}

namespace {
// doPressurePoisson tests scalar, then onepass, then fused
// (GPU/InamuroCUDA.cu:2216-2296).
int variant_of(bool fused, bool onepass, bool scalar)
{
    return scalar ? LBM_POISSON_SCALAR : (onepass ? LBM_POISSON_ONEPASS : (fused ? LBM_POISSON_FUSED : LBM_POISSON_SPLIT));
}
} // namespace

void InamuroHIP::setUseFusedPoisson(bool enabled)
{
    lbm_poisson_accel_t a = accel_;
    a.variant = variant_of(enabled, onepass_, scalar_);
    applyAccel(a);
    fused_ = enabled;
}

void InamuroHIP::setUseOnePassPoisson(bool enabled)
{
    lbm_poisson_accel_t a = accel_;
    a.variant = variant_of(fused_, enabled, scalar_);
    applyAccel(a);
    onepass_ = enabled;
}

void InamuroHIP::setUseScalarPoisson(bool enabled)
{
    lbm_poisson_accel_t a = accel_;
    a.variant = variant_of(fused_, onepass_, enabled);
    applyAccel(a);
    scalar_ = enabled;
}

#define LBM_ACCEL_SETTER(name, type, field, conv) \
    void InamuroHIP::name(type value)             \
    {                                             \
        lbm_poisson_accel_t a = accel_;           \
        a.field = conv;                           \
        applyAccel(a);                            \
    }
LBM_ACCEL_SETTER(setScalarPoissonSourceScale, double, scalar_source_scale, value)
LBM_ACCEL_SETTER(setUseSourceAwareHHInit, bool, source_aware_hh_init, value ? 1 : 0)
LBM_ACCEL_SETTER(setSourceAwareHHScale, double, source_aware_hh_scale, value)
LBM_ACCEL_SETTER(setPressureRelaxScale, double, pressure_relax_scale, value)
LBM_ACCEL_SETTER(setPoissonFixedPointRelax, double, fixed_point_relax, value)
LBM_ACCEL_SETTER(setUsePoissonAndersonM1, bool, anderson_m1, value ? 1 : 0)
LBM_ACCEL_SETTER(setPoissonAndersonBetaMax, double, anderson_beta_max, value)
LBM_ACCEL_SETTER(setUsePoissonTwoGridCorrection, bool, two_grid, value ? 1 : 0)
LBM_ACCEL_SETTER(setPoissonTwoGridStrength, double, two_grid_strength, value)
LBM_ACCEL_SETTER(setUseFusedBoundaryPressure, bool, fused_boundary, value ? 1 : 0)
LBM_ACCEL_SETTER(setUsePoissonSpatialDiagnostics, bool, spatial_diagnostics, value ? 1 : 0)
#undef LBM_ACCEL_SETTER

namespace {

// One row of writePoissonDiagnostic (GPU/InamuroCUDA.cu:2652-2681): the
// block columns are empty / 0 / 0 unless the check formed 4^3 block sums
// (two-grid correction or spatial diagnostics).
void write_check_row(void* user, const lbm_poisson_check_t* r)
{
    std::ofstream& out = *static_cast<std::ofstream*>(user);
    out << r->step << ',' << r->iteration << ',' << std::setprecision(17) << r->pressure_l1_delta << ','
        << r->pressure_l1_norm << ',' << r->relative_error << ',';
    auto opt = [&](double v) {
        if (std::isfinite(v)) out << v;
        out << ',';
    };
    opt(r->fixed_point_h_l1);
    opt(r->fixed_point_h_scale);
    opt(r->fixed_point_h_relative);
    opt(r->fixed_point_p_l1);
    opt(r->fixed_point_p_scale);
    opt(r->fixed_point_p_relative);
    opt(r->fixed_point_relative);
    out << (r->fixed_point_evaluated ? 1 : 0) << ',' << (r->pressure_converged ? 1 : 0) << ','
        << (r->fixed_point_converged ? 1 : 0) << ',' << (r->dual_residual_enabled ? 1 : 0) << ','
        << (r->converged ? 1 : 0) << ',';
    if (std::isfinite(r->block_low_frequency_fraction)) out << r->block_low_frequency_fraction;
    out << ',' << (r->block_count > 0 ? r->block_size : 0) << ',' << r->block_count << '\n';
    out.flush();
}

} // namespace

void InamuroHIP::setPoissonDiagnosticsPath(const std::string& path)
{
    check(lbm_set_check_observer(ctx, nullptr, nullptr), ctx);
    pdiag_.reset();
    if (path.empty()) return;
    auto out = std::make_unique<std::ofstream>(path);
    if (!*out) throw std::runtime_error("failed to open Poisson diagnostics CSV: " + path);
    *out << "step,iteration,pressure_l1_delta,pressure_l1_norm,relative_error,"
         << "fixed_point_h_l1,fixed_point_h_scale,fixed_point_h_relative,"
         << "fixed_point_p_l1,fixed_point_p_scale,fixed_point_p_relative,"
         << "fixed_point_relative,fixed_point_evaluated,pressure_converged,fixed_point_converged,"
         << "dual_residual_enabled,converged,block_low_frequency_fraction,"
         << "block_size,block_count\n";
    out->flush();
    check(lbm_set_check_observer(ctx, write_check_row, out.get()), ctx);
    pdiag_ = std::move(out);
}

lbm_poisson_diag_t InamuroHIP::getLastPoissonDiagnostics() const
{
    lbm_poisson_diag_t d;
    check(lbm_poisson_diag(ctx, &d), ctx);
    return d;
}

void InamuroHIP::downloadHH(std::vector<double>& hh) const
{
    hh.resize(15 * n);
    check(lbm_download_dists(ctx, nullptr, nullptr, hh.data()), ctx);
}

void InamuroHIP::downloadFields(std::vector<double>* f) const
{
    for (int i = 0; i < 6; ++i) f[i].resize(n);
    check(lbm_download(ctx, f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(),
                       f[5].data()),
          ctx);
}

void InamuroHIP::uploadFields(const std::vector<double>* f)
{
    check(lbm_upload(ctx, f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(),
                     f[5].data()),
          ctx);
}

void InamuroHIP::initializeReferenceDroplets() { check(lbm_init_reference(ctx), ctx); }

void InamuroHIP::placeDroplet(double cx, double cy, double cz, double vx, double vy, double vz,
                              bool tanh_profile, double width)
{
    check(lbm_place_droplet(ctx, cx, cy, cz, vx, vy, vz, tanh_profile ? 1 : 0, width), ctx);
}

void InamuroHIP::resetFields() { check(lbm_reset_fields(ctx), ctx); }
void InamuroHIP::setTiming(bool on) { check(lbm_set_timing(ctx, on ? 1 : 0), ctx); }

lbm_perf_t InamuroHIP::perf() const
{
    lbm_perf_t p;
    check(lbm_perf_get(ctx, &p), ctx);
    return p;
}

double InamuroHIP::mass() const
{
    double m = 0.0;
    check(lbm_mass(ctx, &m), ctx);
    return m;
}

// ------------------------------------------------------- Inamuro::Parameters

// params.in: 14 positional lines (CPU/Inamuro.cpp:37-53), then optional
// key=value lines; '#' starts a comment (CPU/InamuroSolver.cpp:143-175).
Inamuro::Parameters::Parameters(const std::string& filename)
{
    std::ifstream file(filename);
    if (!file) throw std::runtime_error("cannot open parameter file " + filename);
    if (!(file >> lx >> ly >> lz >> period >> max_time_steps >> output_frequency >> rho_L >> rho_G >>
          tauf >> taug >> mu_L >> mu_G >> k_f >> k_g >> T >> a >> b >> fei_max >> fei_min >> fei_L >>
          fei_G >> DD))
        throw std::runtime_error("malformed parameter file " + filename);
    std::string line;
    std::getline(file, line);
    while (std::getline(file, line)) {
        const auto hash = line.find('#');
        if (hash != std::string::npos) line = line.substr(0, hash);
        line = trim(line);
        const auto eq = line.find('=');
        if (line.empty() || eq == std::string::npos) continue;
        const std::string key = trim(line.substr(0, eq));
        const std::string value = trim(line.substr(eq + 1));
        auto num = [&](double& dst) {
            dst = std::stod(value);
            gpu_tree_init = true;
        };
        if (key == "init_mode") {
            init_mode = value;
            gpu_tree_init = true;
        } else if (key == "init_profile") {
            init_profile = value;
            gpu_tree_init = true;
        } else if (key == "init_velocity") {
            num(init_velocity);
        } else if (key == "interface_width") {
            num(interface_width);
        } else if (key == "init_center_x") {
            num(init_center_x);
        } else if (key == "init_center_y") {
            num(init_center_y);
        } else if (key == "init_center_z") {
            num(init_center_z);
        } else if (key == "init_separation") {
            num(init_separation);
        } else if (key == "init_offset_x") {
            num(init_offset_x);
        } else if (key == "precision") {
            if (value != "fp64" && value != "mixed")
                throw std::invalid_argument("precision must be fp64 or mixed");
            precision = value;
        }
    }
}

lbm_params_t Inamuro::Parameters::to_abi() const
{
    return lbm_params_t{lx, ly, lz, rho_L, rho_G, mu_L, mu_G, tauf, taug, k_f, k_g, T, a, b,
                        fei_max, fei_min, fei_L, fei_G, DD};
}

// ------------------------------------------------------------------ Inamuro

Inamuro::Inamuro(const std::string& filename) : Inamuro(Parameters(filename)) {}

Inamuro::Inamuro(const Parameters& p)
    : params(p),
      hip(p.to_abi(), p.precision == "mixed" ? LBM_PRECISION_MIXED : LBM_PRECISION_FP64)
{
    initialize();
}

void Inamuro::initialize()
{
    if (!params.gpu_tree_init) {
        // CPU-tree initializeDroplets (CPU/Inamuro.cpp:550-587).
        hip.initializeReferenceDroplets();
        return;
    }
    // GPU-tree initializeDroplets (GPU/Inamuro.cpp:491-566).
    hip.resetFields();
    const bool use_tanh = params.init_profile == "tanh";
    if (params.init_profile != "sharp" && params.init_profile != "tanh")
        std::cerr << "warning: unknown init_profile=" << params.init_profile << ", using sharp"
                  << std::endl;
    const double cy_default = params.ly / 2.0;
    const double cz_default = params.lz / 2.0;
    const double width = params.interface_width > 0.0 ? params.interface_width : 2.0;
    if (params.init_mode == "single_droplet") {
        const double cx = params.init_center_x >= 0.0 ? params.init_center_x : params.lx / 2.0;
        const double cy = params.init_center_y >= 0.0 ? params.init_center_y : cy_default;
        const double cz = params.init_center_z >= 0.0 ? params.init_center_z : cz_default;
        hip.placeDroplet(cx, cy, cz, 0.0, 0.0, params.init_velocity, use_tanh, width);
        return;
    }
    if (params.init_mode != "two_droplets")
        std::cerr << "warning: unknown init_mode=" << params.init_mode << ", using two_droplets"
                  << std::endl;
    const double cx = params.init_center_x >= 0.0 ? params.init_center_x : 0.0;
    const double cy = params.init_center_y >= 0.0 ? params.init_center_y : cy_default;
    const double sep = params.init_separation >= 0.0 ? params.init_separation : params.DD;
    const double cz0 = params.init_center_z >= 0.0 ? params.init_center_z : cz_default;
    hip.placeDroplet(cx - 0.5 * params.init_offset_x, cy, cz0 + sep, 0.0, 0.0, -params.init_velocity,
                     use_tanh, width);
    hip.placeDroplet(cx + 0.5 * params.init_offset_x, cy, cz0 - sep, 0.0, 0.0, params.init_velocity,
                     use_tanh, width);
}

void Inamuro::collision()
{
    throw std::logic_error("collision runs fused inside performTimeStep on the device");
}

void Inamuro::stream(Vector4D&)
{
    throw std::logic_error("streaming runs fused inside performTimeStep on the device");
}

void Inamuro::applyBoundaryConditions(Vector4D&)
{
    throw std::logic_error("the slip boundary runs fused inside performTimeStep on the device");
}

void Inamuro::getMacro()
{
    throw std::logic_error("getMacro runs fused inside performTimeStep on the device");
}

void Inamuro::performTimeStep()
{
    hip.performTimeStepGPU();
    const lbm_poisson_diag_t d = hip.getLastPoissonDiagnostics();
    last_diag.pressure_solve_call_count = d.pressure_solve_call_count;
    last_diag.prev_iteration_count = d.prev_iterations;
    last_diag.iteration_count = d.iterations;
    last_diag.hit_max_iterations = d.hit_max_iterations != 0;
}

void Inamuro::getGridSize(int& nx, int& ny, int& nz) const
{
    nx = params.lx;
    ny = params.ly;
    nz = params.lz;
}

namespace {

void update_stats(Inamuro::FieldStatistics& s, double v, std::size_t& valid, double& acc)
{
    // CPU/Inamuro.cpp:1050-1077
    if (std::isnan(v)) {
        ++s.nan_count;
        return;
    }
    if (!std::isfinite(v)) {
        ++s.inf_count;
        return;
    }
    if (v < s.min) s.min = v;
    if (v > s.max) s.max = v;
    if (std::abs(v) > s.abs_max) s.abs_max = std::abs(v);
    acc += v;
    ++valid;
}

void finalize_stats(Inamuro::FieldStatistics& s, std::size_t valid, double acc)
{
    if (valid == 0) {
        const double nan = std::numeric_limits<double>::quiet_NaN();
        s.min = s.max = s.mean = s.abs_max = nan;
        return;
    }
    s.mean = acc / static_cast<double>(valid);
}

} // namespace

// collectDiagnostics, CPU/Inamuro.cpp:412-507: statistics of the downloaded
// fields in the reference's x-outer order.
Inamuro::DiagnosticsSnapshot Inamuro::collectDiagnostics() const
{
    std::vector<double> f[6];
    hip.downloadFields(f);
    DiagnosticsSnapshot s;
    FieldStatistics* st[7] = {&s.rho_stats, &s.fei_stats, &s.u_stats, &s.v_stats, &s.w_stats,
                              &s.velocity_magnitude_stats, &s.p_stats};
    std::size_t valid[7] = {0, 0, 0, 0, 0, 0, 0};
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    const int lx = params.lx, ly = params.ly, lz = params.lz;
    for (int x = 0; x < lx; ++x)
        for (int y = 0; y < ly; ++y)
            for (int z = 0; z < lz; ++z) {
                const std::size_t i = (static_cast<std::size_t>(z) * ly + y) * lx + x;
                const double u = f[2][i], v = f[3][i], w = f[4][i];
                const double vals[7] = {f[1][i], f[0][i], u, v, w, std::sqrt(u * u + v * v + w * w), f[5][i]};
                for (int k = 0; k < 7; ++k) update_stats(*st[k], vals[k], valid[k], acc[k]);
            }
    for (int k = 0; k < 7; ++k) finalize_stats(*st[k], valid[k], acc[k]);
    return s;
}

namespace {

void put_i32(std::ofstream& o, int v) { o.write(reinterpret_cast<const char*>(&v), sizeof(v)); }
void put_f32(std::ofstream& o, float v) { o.write(reinterpret_cast<const char*>(&v), sizeof(v)); }
void put_str(std::ofstream& o, const std::string& s)
{
    for (char ch : s) put_i32(o, static_cast<int>(ch));
    put_i32(o, 0);
}

} // namespace

// writeTecplotBinary, CPU/Inamuro.cpp:865-1013: the #!TDV101 point-packed
// layout, variables X Y Z u v w rho fei press, coordinates as 1-based int32,
// values as float32, x fastest.
void writeTecplotBinary(const std::string& output_dir, int timeStep, int lx, int ly, int lz,
                        const std::vector<double>* f)
{
    std::ostringstream name;
    name << output_dir << "3D" << std::setfill('0') << std::setw(9) << timeStep << ".plt";
    std::ofstream o(name.str(), std::ios::binary);
    if (!o) {
        std::cerr << "cannot open " << name.str() << std::endl;
        return;
    }
    o.write("#!TDV101", 8);
    put_i32(o, 1);
    put_str(o, "Inamuro");
    const char* vars[9] = {"X", "Y", "Z", "u", "v", "w", "rho", "fei", "press"};
    put_i32(o, 9);
    for (const char* v : vars) put_str(o, v);
    put_f32(o, 299.0f);  // zone marker
    put_str(o, "ZONE 001");
    for (int v : {-1, 0, 1, 0, 0}) put_i32(o, v);  // colour, type, point packing, location, faces
    put_i32(o, lx);
    put_i32(o, ly);
    put_i32(o, lz);
    put_i32(o, 0);
    put_f32(o, 357.0f);  // end of header
    put_f32(o, 299.0f);
    for (int v : {3, 3, 3, 1, 1, 1, 1, 1, 1}) put_i32(o, v);  // int32 coordinates, float32 values
    put_i32(o, 0);
    put_i32(o, -1);
    for (int z = 0; z < lz; ++z)
        for (int y = 0; y < ly; ++y)
            for (int x = 0; x < lx; ++x) {
                const std::size_t i = (static_cast<std::size_t>(z) * ly + y) * lx + x;
                put_i32(o, x + 1);
                put_i32(o, y + 1);
                put_i32(o, z + 1);
                for (int k : {2, 3, 4, 1, 0, 5}) put_f32(o, static_cast<float>(f[k][i]));
            }
}

// writeDebugOutput, CPU/Inamuro.cpp:1015-1048: x,y,z (1-based), then
// u, v, w, rho, fei, p in scientific notation with 6 digits.
void writeDebugFieldCSV(const std::string& output_dir, int timeStep, int lx, int ly, int lz,
                        const std::vector<double>* f)
{
    std::ostringstream name;
    name << output_dir << "debug_" << std::setfill('0') << std::setw(9) << timeStep << ".csv";
    std::ofstream o(name.str());
    if (!o) {
        std::cerr << "cannot open " << name.str() << std::endl;
        return;
    }
    o << "x,y,z,u,v,w,rho,fei,p\n" << std::scientific << std::setprecision(6);
    for (int z = 0; z < lz; ++z)
        for (int y = 0; y < ly; ++y)
            for (int x = 0; x < lx; ++x) {
                const std::size_t i = (static_cast<std::size_t>(z) * ly + y) * lx + x;
                o << (x + 1) << "," << (y + 1) << "," << (z + 1) << "," << f[2][i] << "," << f[3][i] << ","
                  << f[4][i] << "," << f[1][i] << "," << f[0][i] << "," << f[5][i] << "\n";
            }
}

// writeResults, CPU/Inamuro.cpp:365-381.
void Inamuro::writeResults(int timeStep)
{
    if (!output_config.enable_tecplot && !output_config.enable_debug_field_csv) return;
    if (!output_config.output_dir.empty()) std::filesystem::create_directories(output_config.output_dir);
    std::vector<double> f[6];
    hip.downloadFields(f);
    if (output_config.enable_tecplot)
        writeTecplotBinary(output_config.output_dir, timeStep, params.lx, params.ly, params.lz, f);
    if (output_config.enable_debug_field_csv)
        writeDebugFieldCSV(output_config.output_dir, timeStep, params.lx, params.ly, params.lz, f);
}

} // namespace lbm
