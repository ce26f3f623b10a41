// lbm_gpu.cpp -- command-line runner with the reference runner's flags and
// strict Poisson-contract validation (GPU/main.cpp:18-381).  The time step
// runs on the MI355X through lbm::Inamuro, with the reference's optional
// Poisson accelerators (--poisson fused|onepass|scalar, --pressure-boundary
// fused, relaxation scales, Anderson, two-grid, source-aware start, spatial
// diagnostics) applied as GPU/main.cpp:330-353 applies them.  Graph replay of
// Poisson segments (--poisson-graph) is always on here.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "InamuroHIP.hpp"

namespace {

struct RunConfig {
    std::string params = "params.in";
    std::string mode = "gpu";
    std::string poisson = "split";
    std::string pressure_boundary = "split";
    std::string precision;
    int steps = -1;
    int poisson_check_interval = 100;
    double poisson_tolerance = 0.001;
    std::uint64_t poisson_iteration_limit = 0;
    double scalar_source_scale = 2.0;
    bool source_aware_hh_init = false;
    double source_aware_hh_scale = 1.0;
    double pressure_relax_scale = 1.0;
    double poisson_fixed_point_relax = 1.0;
    bool poisson_anderson_m1 = false;
    double poisson_anderson_beta_max = 1.0;
    bool poisson_two_grid_correction = false;
    double poisson_two_grid_strength = 0.5;
    bool poisson_spatial_diagnostics = false;
    int output_frequency = 100;
    std::set<int> output_steps;
    bool write_output = false;
    bool print_roofline = true;
    bool poisson_fixed_point_shadow = false;
    bool poisson_graph = false;
    bool poisson_detail = false;
    bool poisson_deterministic_reductions = false;
    bool poisson_dual_residual = false;
    std::string poisson_diagnostics;
    std::string dump_fields;
};

void print_usage(const char* argv0)
{
    std::cout << "Usage: " << argv0 << " [--mode gpu] [--params FILE] [--steps N] [--precision fp64|mixed]\n"
              << "       [--poisson split|fused|onepass|scalar] [--pressure-boundary split|fused]\n"
              << "       [--poisson-check-interval 100] [--poisson-tolerance 1e-3]\n"
              << "       [--poisson-iteration-limit 0] (strict unbounded production)\n"
              << "       [--scalar-source-scale X]\n"
              << "       [--source-aware-hh-init] [--source-aware-hh-scale X]\n"
              << "       [--pressure-relax-scale X] [--poisson-fixed-point-relax X]\n"
              << "       [--poisson-anderson-m1] [--poisson-anderson-beta-max X]\n"
              << "       [--poisson-two-grid-correction] [--poisson-two-grid-strength X]\n"
              << "       [--output-frequency N] [--output-steps A,B,C] [--write-output] [--no-roofline]\n"
              << "       [--poisson-graph] [--poisson-detail] [--poisson-diagnostics CSV]\n"
              << "       [--poisson-deterministic-reductions] [--poisson-fixed-point-shadow]\n"
              << "       [--poisson-dual-residual] [--poisson-spatial-diagnostics] [--dump-fields FILE]\n";
}

std::set<int> parse_output_steps(const std::string& text)
{
    std::set<int> steps;
    std::stringstream ss(text);
    std::string item;
    while (std::getline(ss, item, ',')) {
        if (item.empty()) continue;
        const int step = std::stoi(item);
        if (step < 0) throw std::runtime_error("--output-steps values must be non-negative");
        steps.insert(step);
    }
    return steps;
}

RunConfig parse_args(int argc, char** argv)
{
    RunConfig cfg;
    bool steps_set = false, freq_set = false;
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        auto value = [&]() -> std::string {
            if (i + 1 >= argc) throw std::runtime_error("missing value for " + arg);
            return argv[++i];
        };
        if (arg == "--help" || arg == "-h") {
            print_usage(argv[0]);
            std::exit(0);
        } else if (arg == "--mode") cfg.mode = value();
        else if (arg == "--poisson") cfg.poisson = value();
        else if (arg == "--pressure-boundary") cfg.pressure_boundary = value();
        else if (arg == "--params") cfg.params = value();
        else if (arg == "--precision") cfg.precision = value();
        else if (arg == "--steps") { cfg.steps = std::stoi(value()); steps_set = true; }
        else if (arg == "--poisson-check-interval") cfg.poisson_check_interval = std::stoi(value());
        else if (arg == "--poisson-tolerance") cfg.poisson_tolerance = std::stod(value());
        else if (arg == "--poisson-iteration-limit") cfg.poisson_iteration_limit = std::stoull(value());
        else if (arg == "--scalar-source-scale") cfg.scalar_source_scale = std::stod(value());
        else if (arg == "--source-aware-hh-init") cfg.source_aware_hh_init = true;
        else if (arg == "--source-aware-hh-scale") cfg.source_aware_hh_scale = std::stod(value());
        else if (arg == "--pressure-relax-scale") cfg.pressure_relax_scale = std::stod(value());
        else if (arg == "--poisson-fixed-point-relax") cfg.poisson_fixed_point_relax = std::stod(value());
        else if (arg == "--poisson-anderson-m1") cfg.poisson_anderson_m1 = true;
        else if (arg == "--poisson-anderson-beta-max") cfg.poisson_anderson_beta_max = std::stod(value());
        else if (arg == "--poisson-two-grid-correction") cfg.poisson_two_grid_correction = true;
        else if (arg == "--poisson-two-grid-strength") cfg.poisson_two_grid_strength = std::stod(value());
        else if (arg == "--output-frequency") { cfg.output_frequency = std::stoi(value()); freq_set = true; }
        else if (arg == "--output-steps") cfg.output_steps = parse_output_steps(value());
        else if (arg == "--write-output") cfg.write_output = true;
        else if (arg == "--poisson-detail") cfg.poisson_detail = true;
        else if (arg == "--poisson-graph") cfg.poisson_graph = true;
        else if (arg == "--poisson-deterministic-reductions") cfg.poisson_deterministic_reductions = true;
        else if (arg == "--poisson-fixed-point-shadow") cfg.poisson_fixed_point_shadow = true;
        else if (arg == "--poisson-dual-residual") {
            cfg.poisson_dual_residual = true;
            cfg.poisson_fixed_point_shadow = true;
            cfg.poisson_deterministic_reductions = true;  // GPU/main.cpp:186-191
        } else if (arg == "--poisson-diagnostics") cfg.poisson_diagnostics = value();
        else if (arg == "--poisson-spatial-diagnostics") cfg.poisson_spatial_diagnostics = true;
        else if (arg == "--no-roofline") cfg.print_roofline = false;
        else if (arg == "--dump-fields") cfg.dump_fields = value();
        else throw std::runtime_error("unknown argument: " + arg);
    }
    // Contract validation, GPU/main.cpp:202-247.
    if (cfg.mode != "cpu" && cfg.mode != "gpu") throw std::runtime_error("--mode must be cpu or gpu");
    if (cfg.mode == "cpu")
        throw std::runtime_error("--mode cpu is the reference CPU solver; this runner executes the step on the GPU");
    if (cfg.poisson != "split" && cfg.poisson != "fused" && cfg.poisson != "onepass" && cfg.poisson != "scalar")
        throw std::runtime_error("--poisson must be split, fused, onepass, or scalar");
    if (cfg.pressure_boundary != "split" && cfg.pressure_boundary != "fused")
        throw std::runtime_error("--pressure-boundary must be split or fused");
    if (cfg.poisson_check_interval != 100)
        throw std::runtime_error("strict production requires --poisson-check-interval 100");
    if (cfg.poisson_tolerance != 1.0e-3)
        throw std::runtime_error("strict production requires --poisson-tolerance 1e-3");
    if (cfg.poisson_iteration_limit != 0)
        throw std::runtime_error("strict production requires --poisson-iteration-limit 0 (unbounded)");
    if (!(cfg.poisson_fixed_point_relax > 0.0))
        throw std::runtime_error("--poisson-fixed-point-relax must be positive");
    if (!(cfg.poisson_anderson_beta_max >= 0.0))
        throw std::runtime_error("--poisson-anderson-beta-max must be non-negative");
    if (!(cfg.poisson_two_grid_strength >= 0.0))
        throw std::runtime_error("--poisson-two-grid-strength must be non-negative");
    // GPU/main.cpp:229-247 (graph replay itself is always on here; the flag
    // combination is rejected as the reference rejects it).
    if (cfg.poisson_fixed_point_shadow &&
        (cfg.poisson != "split" || cfg.pressure_boundary != "split" || cfg.source_aware_hh_init ||
         cfg.poisson_graph || cfg.poisson_anderson_m1 || cfg.poisson_two_grid_correction ||
         cfg.pressure_relax_scale != 1.0 || cfg.poisson_fixed_point_relax != 1.0))
        throw std::runtime_error(
            "fixed-point residual requires GPU split+split, unit relaxation, "
            "and source-aware/graph/Anderson/two-grid disabled");
    if (cfg.poisson_dual_residual && cfg.poisson_spatial_diagnostics)
        throw std::runtime_error(
            "dual-residual exit requires check_interval=100, tolerance=1e-3, "
            "poisson_iteration_limit=0, deterministic reductions, and no "
            "spatial diagnostics");
    if (!cfg.precision.empty() && cfg.precision != "fp64" && cfg.precision != "mixed")
        throw std::runtime_error("--precision must be fp64 or mixed");
    // Steps and output frequency from params.in lines 3-4, GPU/main.cpp:98-118.
    std::ifstream file(cfg.params);
    int lx, ly, lz, period, steps, freq;
    if (file >> lx >> ly >> lz >> period >> steps >> freq) {
        if (!steps_set && cfg.steps < 0) cfg.steps = steps;
        if (!freq_set) cfg.output_frequency = freq;
    }
    if (cfg.steps < 0) cfg.steps = 100;
    if (cfg.output_frequency <= 0) cfg.output_frequency = 100;
    return cfg;
}

bool should_write(const RunConfig& cfg, int completed)
{
    if (!cfg.write_output) return false;
    if (!cfg.output_steps.empty()) return cfg.output_steps.count(completed) > 0;
    return completed % cfg.output_frequency == 0;
}

} // namespace

int main(int argc, char** argv)
{
    try {
        const RunConfig cfg = parse_args(argc, argv);
        lbm::Inamuro::Parameters P(cfg.params);
        P.gpu_tree_init = true;  // the GPU runner uses the GPU-tree constructor (GPU/Inamuro.cpp:491-566)
        if (!cfg.precision.empty()) P.precision = cfg.precision;
        std::cout << "LBM runner mode=gpu device=MI355X precision=" << P.precision
                  << " poisson=" << cfg.poisson << " pressure_boundary=" << cfg.pressure_boundary
                  << " poisson_check_interval=100"
                  << " poisson_tolerance=" << cfg.poisson_tolerance << " poisson_iteration_limit=0 (unbounded)"
                  << " poisson_fixed_point_shadow=" << (cfg.poisson_fixed_point_shadow ? "yes" : "no")
                  << " poisson_dual_residual=" << (cfg.poisson_dual_residual ? "yes" : "no")
                  << " poisson_fixed_point_relax=" << cfg.poisson_fixed_point_relax
                  << " poisson_anderson_m1=" << (cfg.poisson_anderson_m1 ? "yes" : "no")
                  << " poisson_anderson_beta_max=" << cfg.poisson_anderson_beta_max
                  << " poisson_two_grid_correction=" << (cfg.poisson_two_grid_correction ? "yes" : "no")
                  << " poisson_two_grid_strength=" << cfg.poisson_two_grid_strength
                  << " poisson_spatial_diagnostics=" << (cfg.poisson_spatial_diagnostics ? "yes" : "no")
                  << " params=" << cfg.params << " steps=" << cfg.steps << std::endl;
        lbm::Inamuro solver(P);
        lbm::InamuroHIP& gpu = solver.gpu();
        gpu.setTiming(true);
        gpu.setEnablePoissonDetailTiming(cfg.poisson_detail);
        gpu.setUsePoissonDeterministicReductions(cfg.poisson_deterministic_reductions);
        gpu.setPoissonConvergence(cfg.poisson_check_interval, cfg.poisson_tolerance);
        gpu.setPoissonIterationLimit(cfg.poisson_iteration_limit);
        gpu.setEnablePoissonFixedPointShadow(cfg.poisson_fixed_point_shadow);
        gpu.setUsePoissonDualResidual(cfg.poisson_dual_residual);
        // GPU/main.cpp:330-351.
        gpu.setUseFusedPoisson(cfg.poisson == "fused");
        gpu.setUseOnePassPoisson(cfg.poisson == "onepass");
        gpu.setUseScalarPoisson(cfg.poisson == "scalar");
        gpu.setScalarPoissonSourceScale(cfg.scalar_source_scale);
        gpu.setUseSourceAwareHHInit(cfg.source_aware_hh_init);
        gpu.setSourceAwareHHScale(cfg.source_aware_hh_scale);
        gpu.setPressureRelaxScale(cfg.pressure_relax_scale);
        gpu.setPoissonFixedPointRelax(cfg.poisson_fixed_point_relax);
        gpu.setUsePoissonAndersonM1(cfg.poisson_anderson_m1);
        gpu.setPoissonAndersonBetaMax(cfg.poisson_anderson_beta_max);
        gpu.setUsePoissonTwoGridCorrection(cfg.poisson_two_grid_correction);
        gpu.setPoissonTwoGridStrength(cfg.poisson_two_grid_strength);
        gpu.setUseFusedBoundaryPressure(cfg.pressure_boundary == "fused");
        gpu.setUsePoissonSpatialDiagnostics(cfg.poisson_spatial_diagnostics);
        lbm::Inamuro::OutputConfig oc;
        oc.enable_tecplot = true;
        oc.output_dir = "out/";
        solver.setOutputConfig(oc);

        // GPU/main.cpp:355: one row per convergence check.
        gpu.setPoissonDiagnosticsPath(cfg.poisson_diagnostics);
        if (cfg.write_output && (cfg.output_steps.empty() || cfg.output_steps.count(0) > 0))
            solver.writeResults(0);
        const auto t0 = std::chrono::steady_clock::now();
        for (int step = 0; step < cfg.steps; ++step) {
            solver.performTimeStep();
            if (should_write(cfg, step + 1)) solver.writeResults(step + 1);
        }
        gpu.synchronize();
        const auto t1 = std::chrono::steady_clock::now();
        const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        const double cells = static_cast<double>(P.lx) * P.ly * P.lz;
        std::cout << std::fixed << std::setprecision(3) << "GPU wall_total_ms=" << ms
                  << " wall_avg_ms_per_step=" << ms / std::max(cfg.steps, 1)
                  << " wall_MLUPS=" << cells * cfg.steps / (ms * 1000.0) << std::endl;
        // GPU/main.cpp:373-376
        gpu.printPerformanceMetrics();
        if (cfg.print_roofline) gpu.printRooflineSummary();
        if (!cfg.dump_fields.empty()) {
            std::vector<double> f[6];
            gpu.downloadFields(f);
            std::ofstream o(cfg.dump_fields, std::ios::binary);
            for (auto& a : f) o.write(reinterpret_cast<const char*>(a.data()), a.size() * sizeof(double));
        }
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << std::endl;
        print_usage(argv[0]);
        return 1;
    }
    return 0;
}
