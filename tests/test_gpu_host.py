"""The C++ host layer on the GPU: the lbm_gpu runner (GPU-tree
initialisation, GPU/Inamuro.cpp:491-566) and lbm_sim / InamuroSolver
(CPU-tree initialisation, per-step step_diagnostics.csv with the reference's
49 columns, CPU/InamuroSolver.cpp:270-337) against the oracle."""
import csv
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LBM_GPU = os.path.join(ROOT, "lbm_amd", "lbm_gpu")
LBM_SIM = os.path.join(ROOT, "lbm_amd", "lbm_sim")
FIELDS = ("fei", "rho", "u", "v", "w", "p")


def rel_linf(a, b):
    return float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)


def test_runner_matches_oracle_with_gpu_tree_initialisation(tmp_path):
    P = po.Params(lx=16, ly=32, lz=32, DD=8.0, max_time_steps=3)
    pfile = tmp_path / "params.in"
    pfile.write_text(P.params_in())
    dump, diag = tmp_path / "fields.bin", tmp_path / "poisson.csv"
    r = subprocess.run([LBM_GPU, "--params", str(pfile), "--dump-fields", str(dump),
                        "--poisson-diagnostics", str(diag)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "wall_MLUPS=" in r.stdout
    # printPerformanceMetrics and printRooflineSummary (GPU/main.cpp:373-376)
    assert "---- MI355X performance over 3 steps ----" in r.stdout
    assert "Poisson iteration: 272 B/cell" in r.stdout and "% of 8 TB/s" in r.stdout
    # GPU-tree default: two sharp droplets through place_droplet at x = 0.
    o = po.Oracle(P)
    o.place_droplet(0.0, P.ly / 2.0, P.lz / 2.0 + P.DD, vz=-0.035, profile="sharp")
    o.place_droplet(0.0, P.ly / 2.0, P.lz / 2.0 - P.DD, vz=0.035, profile="sharp")
    iters, finals = [], []
    for _ in range(3):
        iters += o.step(1)
        finals.append(o.diag())
    # The Poisson diagnostics CSV: the reference's 20-column header
    # (GPU/InamuroCUDA.cu:2621-2626) and one row per convergence check.
    with open(diag) as fh:
        header = fh.readline().strip().split(",")
        lines = fh.read().splitlines()
    assert header == ["step", "iteration", "pressure_l1_delta", "pressure_l1_norm", "relative_error",
                      "fixed_point_h_l1", "fixed_point_h_scale", "fixed_point_h_relative",
                      "fixed_point_p_l1", "fixed_point_p_scale", "fixed_point_p_relative",
                      "fixed_point_relative", "fixed_point_evaluated", "pressure_converged",
                      "fixed_point_converged", "dual_residual_enabled", "converged",
                      "block_low_frequency_fraction", "block_size", "block_count"]
    rows = [dict(zip(header, ln.split(","))) for ln in lines]
    assert all(len(ln.split(",")) == 20 for ln in lines)
    want = [(s + 1, k) for s, n in enumerate(iters) for k in range(100, n + 1, 100)]
    assert [(int(r_["step"]), int(r_["iteration"])) for r_ in rows] == want
    for r_ in rows:
        last = int(r_["iteration"]) == iters[int(r_["step"]) - 1]
        assert r_["converged"] == r_["pressure_converged"] == ("1" if last else "0")
        assert (float(r_["relative_error"]) < 1e-3) == last
        # no dual residual and no spatial diagnostics: empty optional columns
        assert r_["fixed_point_h_l1"] == r_["fixed_point_relative"] == ""
        assert r_["fixed_point_evaluated"] == r_["dual_residual_enabled"] == "0"
        assert (r_["block_low_frequency_fraction"], r_["block_size"], r_["block_count"]) == ("", "0", "0")
        assert abs(float(r_["relative_error"]) - float(r_["pressure_l1_delta"]) / float(r_["pressure_l1_norm"])) \
            <= 1e-15 * float(r_["relative_error"])
    ends = [r_ for r_ in rows if r_["converged"] == "1"]
    for r_, d in zip(ends, finals):
        for k, ref in (("pressure_l1_delta", d.pressure_l1_delta), ("pressure_l1_norm", d.pressure_l1_norm),
                       ("relative_error", d.relative_error)):
            assert abs(float(r_[k]) - ref) <= 1e-10 * abs(ref), (k, r_[k], ref)
    got = np.fromfile(dump, dtype=np.float64).reshape(6, P.lz, P.ly, P.lx)
    for i, k in enumerate(FIELDS):
        assert rel_linf(got[i], o.get(k)) <= 1e-12, k


def test_runner_applies_the_poisson_accelerators(tmp_path):
    """lbm_gpu --poisson onepass --poisson-anderson-m1
    --poisson-two-grid-correction --poisson-spatial-diagnostics, as
    GPU/main.cpp:330-351 applies them, against the oracle's restatement: the
    same iteration counts, fields within 1e-10, and the CSV's block columns
    filled (4, 256 blocks, the spatial fraction)."""
    P = po.Params(lx=16, ly=32, lz=32, DD=8.0, max_time_steps=2)
    pfile = tmp_path / "params.in"
    pfile.write_text(P.params_in())
    dump, diag = tmp_path / "fields.bin", tmp_path / "poisson.csv"
    r = subprocess.run([LBM_GPU, "--params", str(pfile), "--poisson", "onepass", "--poisson-anderson-m1",
                        "--poisson-two-grid-correction", "--poisson-two-grid-strength", "0.3",
                        "--poisson-spatial-diagnostics", "--pressure-boundary", "fused",
                        "--dump-fields", str(dump), "--poisson-diagnostics", str(diag)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "poisson=onepass pressure_boundary=fused" in r.stdout
    assert "poisson_anderson_m1=yes" in r.stdout and "poisson_two_grid_correction=yes" in r.stdout
    o = po.Oracle(P, threads=4)
    o.place_droplet(0.0, P.ly / 2.0, P.lz / 2.0 + P.DD, vz=-0.035, profile="sharp")
    o.place_droplet(0.0, P.ly / 2.0, P.lz / 2.0 - P.DD, vz=0.035, profile="sharp")
    o.set_poisson_accel(variant="onepass", anderson_m1=1, two_grid=1, two_grid_strength=0.3,
                        spatial_diagnostics=1, fused_boundary=1)
    iters = o.step(2)
    d = o.diag()
    with open(diag) as fh:
        header = fh.readline().strip().split(",")
        rows = [dict(zip(header, ln.split(","))) for ln in fh.read().splitlines()]
    want = [(s_ + 1, k) for s_, n in enumerate(iters) for k in range(100, n + 1, 100)]
    assert [(int(r_["step"]), int(r_["iteration"])) for r_ in rows] == want
    for r_ in rows:
        assert (r_["block_size"], r_["block_count"]) == ("4", "256")
        assert 0.0 <= float(r_["block_low_frequency_fraction"]) <= 1.0
    assert float(rows[-1]["block_low_frequency_fraction"]) == pytest.approx(d.block_low_frequency_fraction,
                                                                            rel=1e-9)
    got = np.fromfile(dump, dtype=np.float64).reshape(6, P.lz, P.ly, P.lx)
    for i, k in enumerate(FIELDS):
        assert rel_linf(got[i], o.get(k)) <= 1e-10, k


def test_solver_writes_reference_step_diagnostics(tmp_path):
    P = po.Params(lx=16, ly=32, lz=32, DD=8.0, max_time_steps=3, output_frequency=2)
    out = tmp_path / "out"
    pfile = tmp_path / "params.in"
    pfile.write_text(P.params_in(f"enable_tecplot_output=1\noutput_dir={out}/\n"
                                 "step_summary_filename=steps.csv\nenable_progress=0\n"))
    r = subprocess.run([LBM_SIM, str(pfile)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    rows = list(csv.reader(open(out / "steps.csv")))
    assert len(rows[0]) == 49 and rows[0][:5] == ["time_step", "step_time_ms", "pressure_solve_call_count",
                                                   "prev_pressure_iterations", "pressure_iterations"]
    assert [r_[0] for r_ in rows[1:]] == ["0", "1", "2", "3"]
    o = po.Oracle(P)
    o.init_two_droplets_sharp()
    prev = 0
    for s in range(1, 4):
        it = o.step(1)[0]
        row = dict(zip(rows[0], rows[s + 1]))
        assert int(row["pressure_iterations"]) == it
        assert int(row["prev_pressure_iterations"]) == prev
        assert int(row["pressure_solve_call_count"]) == s
        prev = it
    last = dict(zip(rows[0], rows[4]))
    f = o.fields()
    assert abs(float(last["fei_mean"]) - f["fei"].mean()) <= 1e-9 * abs(f["fei"].mean())
    assert abs(float(last["rho_max"]) - f["rho"].max()) <= 1e-9 * f["rho"].max()
    assert abs(float(last["p_abs_max"]) - np.abs(f["p"]).max()) <= 1e-9 * np.abs(f["p"]).max()
    # Tecplot files at steps 0 and 2, point-packed: 3 int coordinates + 6 float32 per cell.
    for step in (0, 2):
        plt = out / f"3D{step:09d}.plt"
        data = plt.read_bytes()
        assert data[:8] == b"#!TDV101"
        assert len(data) > P.lx * P.ly * P.lz * 36
    assert not (out / "steps.csv.lock").exists()


WRAPPER_DRIVER = r"""
#include <iostream>
#include <stdexcept>
#include "InamuroHIP.hpp"
int main()
{
    lbm_params_t P{16, 32, 32, 50.0, 1.0, 0.008, 0.00016, 1.0, 1.0, 0.5, 0.0005, 0.035, 1.0, 6.7,
                   0.09714, 0.01134, 0.092, 0.015, 8.0};
    lbm::InamuroHIP gpu(P, LBM_PRECISION_FP64);
    gpu.initializeReferenceDroplets();
    gpu.setPoissonConvergence(50, 1e-3);
    gpu.setUsePoissonDualResidual(true);
    try { gpu.performTimeStepGPU(); return 2; } catch (const std::runtime_error& e) { std::cout << "A " << e.what() << std::endl; }
    gpu.setPoissonConvergence(100, 1e-3);
    gpu.setUsePoissonDeterministicReductions(false);
    try { gpu.performTimeStepGPU(); return 3; } catch (const std::runtime_error& e) { std::cout << "B " << e.what() << std::endl; }
    gpu.setUsePoissonDeterministicReductions(true);
    gpu.setEnablePoissonDetailTiming(true);
    gpu.performTimeStepGPU();
    gpu.performTimeStepGPU();
    std::cout << "C " << gpu.getLastPoissonDiagnostics().fixed_point_evaluated << std::endl;
    gpu.printPerformanceMetrics();
    gpu.printRooflineSummary();
    gpu.resetPerformanceMetrics();
    gpu.printPerformanceMetrics();
    return 0;
}
"""


def test_wrapper_mirrors_the_dual_residual_contract_and_reports(tmp_path):
    """lbm::InamuroHIP refuses the dual-residual exit outside its contract with
    the reference's message (GPU/InamuroCUDA.cu:1857-1865): a check interval
    of 50, then deterministic reductions switched off after the dual residual
    switched them on (:2573-2580).  Within the contract it steps, and the
    reports carry the detail timing and the byte model."""
    src = tmp_path / "driver.cpp"
    src.write_text(WRAPPER_DRIVER)
    exe = tmp_path / "driver"
    host = os.path.join(ROOT, "lbm_amd", "host")
    cmd = ["g++", "-std=c++17", "-O1", f"-I{ROOT}/include", f"-I{host}", str(src),
           os.path.join(host, "InamuroHIP.cpp"), "-o", str(exe), f"-L{ROOT}/lbm_amd", "-llbm_hip",
           f"-Wl,-rpath,{ROOT}/lbm_amd", "-Wl,-rpath,/opt/rocm/lib"]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert b.returncode == 0, b.stderr
    print("wrapper driver built; running it", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    msg = ("dual-residual exit requires check_interval=100, tolerance=1e-3, poisson_iteration_limit=0, "
           "deterministic reductions, no spatial diagnostics")
    assert "A " + msg in r.stdout and "B " + msg in r.stdout
    assert "C 1" in r.stdout
    assert "---- MI355X performance over 2 steps ----" in r.stdout
    assert "kinetic fixed-point residual: dual exit" in r.stdout
    assert "Poisson detail: " in r.stdout and " us per iteration" in r.stdout
    assert "Poisson iteration: 272 B/cell" in r.stdout and "measured Poisson iteration:" in r.stdout
    assert r.stdout.rstrip().endswith("performance: no steps since the last reset")
