// InamuroSolver.hpp -- run loop over lbm::Inamuro with the reference driver's
// interface and outputs (CPU/InamuroSolver.hpp:9-80): params.in lines 3-4
// (steps, output frequency) plus key=value switches, per-step
// step_diagnostics.csv with the reference's 49 columns, and a lock directory
// so two runs never write one CSV.
#pragma once

#include <chrono>
#include <filesystem>
#include <fstream>
#include <memory>
#include <string>

#include "InamuroHIP.hpp"

namespace lbm {

class InamuroSolver {
public:
    struct SolverConfig {
        bool enable_timing = true;
        bool enable_progress = true;
        bool enable_tecplot_output = false;
        bool enable_debug_field_csv = false;
        bool enable_step_summary_csv = true;
        bool enable_basic_warnings = true;
        std::string output_dir = "out/";
        std::string step_summary_filename = "step_diagnostics.csv";
    };

    explicit InamuroSolver(const std::string& filename);
    ~InamuroSolver();
    InamuroSolver(const InamuroSolver&) = delete;
    InamuroSolver& operator=(const InamuroSolver&) = delete;

    void run();
    void setSolverConfig(const SolverConfig& config);
    Inamuro& solver() { return *inamuro; }
    int maxTimeSteps() const { return max_time_steps; }

private:
    void loadConfiguration(const std::string& filename);
    void applyOutputConfiguration();
    void openDiagnosticsFile();
    void recordStepDiagnostics(int timeStep);
    void releaseDiagnosticsLock();

    std::unique_ptr<Inamuro> inamuro;
    SolverConfig solver_config;
    int max_time_steps = 20;
    int output_frequency = 10;
    double last_step_time_ms = 0.0;
    double total_step_time_ms = 0.0;
    double last_solve_ms = 0.0;
    bool warned_negative_rho = false;
    bool warned_non_finite = false;
    std::ofstream diagnostics_stream;
    std::filesystem::path diagnostics_lock_dir;
};

} // namespace lbm
