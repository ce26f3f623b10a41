// InamuroSolver.cpp -- driver of lbm::Inamuro (CPU/InamuroSolver.cpp:27-457).
#include "InamuroSolver.hpp"

#include <iomanip>
#include <iostream>
#include <sstream>

namespace lbm {

namespace {

std::string trim(const std::string& t)
{
    const auto a = t.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const auto b = t.find_last_not_of(" \t\r\n");
    return t.substr(a, b - a + 1);
}

bool to_bool(const std::string& v)
{
    return v == "1" || v == "true" || v == "TRUE" || v == "on" || v == "ON";
}

} // namespace

InamuroSolver::InamuroSolver(const std::string& filename)
{
    inamuro = std::make_unique<Inamuro>(filename);
    inamuro->gpu().setTiming(true);
    loadConfiguration(filename);
    applyOutputConfiguration();
    openDiagnosticsFile();
}

InamuroSolver::~InamuroSolver() { releaseDiagnosticsLock(); }

void InamuroSolver::setSolverConfig(const SolverConfig& config)
{
    solver_config = config;
    applyOutputConfiguration();
}

// Lines 3-4 and key=value switches, CPU/InamuroSolver.cpp:116-219.
void InamuroSolver::loadConfiguration(const std::string& filename)
{
    const Inamuro::Parameters& P = inamuro->parameters();
    max_time_steps = P.max_time_steps;
    output_frequency = P.output_frequency > 0 ? P.output_frequency : 10;
    std::ifstream file(filename);
    std::string line;
    for (int i = 0; i < 14 && std::getline(file, line); ++i) {
    }
    while (std::getline(file, line)) {
        const auto hash = line.find('#');
        if (hash != std::string::npos) line = line.substr(0, hash);
        line = trim(line);
        const auto eq = line.find('=');
        if (line.empty() || eq == std::string::npos) continue;
        const std::string key = trim(line.substr(0, eq));
        const std::string value = trim(line.substr(eq + 1));
        if (key == "enable_timing") solver_config.enable_timing = to_bool(value);
        else if (key == "enable_progress") solver_config.enable_progress = to_bool(value);
        else if (key == "enable_tecplot_output") solver_config.enable_tecplot_output = to_bool(value);
        else if (key == "enable_debug_field_csv") solver_config.enable_debug_field_csv = to_bool(value);
        else if (key == "enable_step_summary_csv") solver_config.enable_step_summary_csv = to_bool(value);
        else if (key == "enable_basic_warnings") solver_config.enable_basic_warnings = to_bool(value);
        else if (key == "output_dir") solver_config.output_dir = value;
        else if (key == "step_summary_filename") solver_config.step_summary_filename = value;
    }
}

void InamuroSolver::applyOutputConfiguration()
{
    Inamuro::OutputConfig oc;
    oc.enable_tecplot = solver_config.enable_tecplot_output;
    oc.enable_debug_field_csv = solver_config.enable_debug_field_csv;
    oc.enable_step_summary_csv = solver_config.enable_step_summary_csv;
    oc.output_dir = solver_config.output_dir;
    oc.summary_filename = solver_config.step_summary_filename;
    inamuro->setOutputConfig(oc);
}

// CPU/InamuroSolver.cpp:234-288.
void InamuroSolver::openDiagnosticsFile()
{
    if (!solver_config.enable_step_summary_csv) return;
    std::filesystem::create_directories(solver_config.output_dir);
    const std::filesystem::path path =
        std::filesystem::path(solver_config.output_dir) / solver_config.step_summary_filename;
    diagnostics_lock_dir = path;
    diagnostics_lock_dir += ".lock";
    std::error_code ec;
    if (!std::filesystem::create_directory(diagnostics_lock_dir, ec)) {
        std::cerr << "warning: diagnostics file " << path
                  << " is held by another run; step_diagnostics.csv is disabled for this one" << std::endl;
        diagnostics_lock_dir.clear();
        return;
    }
    diagnostics_stream.open(path, std::ios::out | std::ios::trunc);
    if (!diagnostics_stream) {
        std::cerr << "warning: cannot open diagnostics file " << path << std::endl;
        releaseDiagnosticsLock();
        return;
    }
    diagnostics_stream << std::scientific << std::setprecision(10);
    diagnostics_stream
        << "time_step"
        << ",step_time_ms,pressure_solve_call_count,prev_pressure_iterations,pressure_iterations,pressure_hit_max_iter,pressure_solve_ms";
    for (const char* f : {"rho", "fei", "u", "v", "w", "vel_mag", "p"})
        diagnostics_stream << "," << f << "_min," << f << "_max," << f << "_mean," << f << "_abs_max," << f
                           << "_nan_count," << f << "_inf_count";
    diagnostics_stream << "\n";
}

// CPU/InamuroSolver.cpp:290-370.
void InamuroSolver::recordStepDiagnostics(int timeStep)
{
    if (!diagnostics_stream) return;
    const Inamuro::DiagnosticsSnapshot s = inamuro->collectDiagnostics();
    const Inamuro::PressureSolveDiagnostics& d = inamuro->getLastPressureSolveDiagnostics();
    diagnostics_stream << timeStep;
    if (timeStep == 0) {
        diagnostics_stream << ",0,0,0,0,0,0";
    } else {
        diagnostics_stream << "," << last_step_time_ms << "," << d.pressure_solve_call_count << ","
                           << d.prev_iteration_count << "," << d.iteration_count << ","
                           << (d.hit_max_iterations ? 1 : 0) << "," << last_solve_ms;
    }
    for (const Inamuro::FieldStatistics* st :
         {&s.rho_stats, &s.fei_stats, &s.u_stats, &s.v_stats, &s.w_stats, &s.velocity_magnitude_stats,
          &s.p_stats})
        diagnostics_stream << "," << st->min << "," << st->max << "," << st->mean << "," << st->abs_max << ","
                           << st->nan_count << "," << st->inf_count;
    diagnostics_stream << "\n";
    diagnostics_stream.flush();
    if (!solver_config.enable_basic_warnings) return;
    bool non_finite = false;
    for (const Inamuro::FieldStatistics* st :
         {&s.rho_stats, &s.fei_stats, &s.u_stats, &s.v_stats, &s.w_stats, &s.velocity_magnitude_stats,
          &s.p_stats})
        non_finite = non_finite || st->nan_count > 0 || st->inf_count > 0;
    if (non_finite && !warned_non_finite) {
        warned_non_finite = true;
        std::cerr << "\n[warning] step " << timeStep << ": first NaN/Inf, see step_diagnostics.csv" << std::endl;
    }
    if (s.rho_stats.min < 0.0 && !warned_negative_rho) {
        warned_negative_rho = true;
        std::cerr << "\n[warning] step " << timeStep << ": first negative density, rho_min=" << s.rho_stats.min
                  << std::endl;
    }
}

void InamuroSolver::releaseDiagnosticsLock()
{
    if (diagnostics_lock_dir.empty()) return;
    std::error_code ec;
    std::filesystem::remove(diagnostics_lock_dir, ec);
    diagnostics_lock_dir.clear();
}

// CPU/InamuroSolver.cpp:56-114.
void InamuroSolver::run()
{
    int lx, ly, lz;
    inamuro->getGridSize(lx, ly, lz);
    std::cout << "Inamuro LBM D3Q15 on MI355X: " << lx << " x " << ly << " x " << lz
              << ", steps " << max_time_steps << ", output every " << output_frequency
              << ", Poisson strict unbounded (check 100, tol 1e-3)" << std::endl;
    recordStepDiagnostics(0);
    for (int step = 0; step < max_time_steps; ++step) {
        if (step % output_frequency == 0) inamuro->writeResults(step);
        const lbm_perf_t before = inamuro->gpu().perf();
        const auto t0 = std::chrono::steady_clock::now();
        inamuro->performTimeStep();
        const auto t1 = std::chrono::steady_clock::now();
        const lbm_perf_t after = inamuro->gpu().perf();
        last_step_time_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        last_solve_ms = (after.poisson_ms + after.check_ms) - (before.poisson_ms + before.check_ms);
        total_step_time_ms += last_step_time_ms;
        recordStepDiagnostics(step + 1);
        if (solver_config.enable_progress) {
            std::cout << "  step " << std::setw(4) << step + 1 << " | " << std::fixed << std::setprecision(2)
                      << last_step_time_ms << " ms | Poisson iterations "
                      << inamuro->getLastPressureSolveDiagnostics().iteration_count << std::endl;
        }
    }
    if (solver_config.enable_timing) {
        std::cout << "total " << std::fixed << std::setprecision(2) << total_step_time_ms / 1000.0 << " s, "
                  << total_step_time_ms / std::max(1, max_time_steps) << " ms/step" << std::endl;
    }
}

} // namespace lbm
