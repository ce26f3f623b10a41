"""The C-ABI library loads and exports every entry point include/lbm_hip.h
declares; the declared structs match the ctypes mirrors.  No compute calls."""
import ctypes
import os
import re
import subprocess

import pytest

from lbm_amd import hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbm_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(lbm_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_binding_list():
    assert declared_functions() == sorted(hip.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = hip.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", hip.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (lbm_\w+)$", out, flags=re.M))
    assert set(declared_functions()) <= exported


def test_struct_sizes_match_header(tmp_path):
    # compile a tiny C program against the header and compare sizeof()
    src = str(tmp_path / "sizes_probe.c")
    with open(src, "w") as fh:
        fh.write('#include <stdio.h>\n#include "lbm_hip.h"\nint main(void){printf("%zu %zu %zu %zu %zu %zu\\n",'
                 ' sizeof(lbm_params_t), sizeof(lbm_poisson_diag_t), sizeof(lbm_perf_t),'
                 ' sizeof(lbm_adjoint_stats_t), sizeof(lbm_poisson_check_t),'
                 ' sizeof(lbm_poisson_accel_t)); return 0;}\n')
    exe = src[:-2]
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
    sizes = list(map(int, subprocess.run([exe], capture_output=True, text=True).stdout.split()))
    assert sizes == [ctypes.sizeof(hip.LbmParams), ctypes.sizeof(hip.PoissonStepDiagnostics),
                     ctypes.sizeof(hip.PerformanceMetrics), ctypes.sizeof(hip.AdjointStats),
                     ctypes.sizeof(hip.PoissonCheck), ctypes.sizeof(hip.PoissonAccel)]


def test_poisson_accelerator_defaults_are_the_reference_defaults():
    """RunConfig (GPU/main.cpp:27-35) and InamuroCUDA (GPU/InamuroCUDA.hpp:
    136-157): split, split boundary, unit relaxations, scalar source scale 2,
    source-aware scale 1, Anderson beta 1, two-grid strength 0.5, all off."""
    a = hip.PoissonAccel()
    hip.load_library().lbm_poisson_accel_defaults(ctypes.byref(a))
    got = {k: getattr(a, k) for k, _ in hip.PoissonAccel._fields_}
    assert got == {"variant": 0, "fused_boundary": 0, "pressure_relax_scale": 1.0,
                   "fixed_point_relax": 1.0, "scalar_source_scale": 2.0, "source_aware_hh_init": 0,
                   "anderson_m1": 0, "source_aware_hh_scale": 1.0, "anderson_beta_max": 1.0,
                   "two_grid_strength": 0.5, "two_grid": 0, "spatial_diagnostics": 0}
    assert hip.load_library().lbm_set_poisson_accel(None, ctypes.byref(a)) == hip.LBM_ERR_ARG


def test_null_arguments_are_rejected_without_touching_the_device():
    lib = hip.load_library()
    assert lib.lbm_destroy(None) == hip.LBM_ERR_ARG
    assert lib.lbm_step(None, 1) == hip.LBM_ERR_ARG
    assert lib.lbm_poisson_diag(None, None) == hip.LBM_ERR_ARG
    out = ctypes.c_void_p()
    assert lib.lbm_create(None, 0, ctypes.byref(out)) == hip.LBM_ERR_ARG
    assert lib.lbm_last_error(None) == b"null argument"


def test_invalid_grid_is_rejected_before_any_allocation():
    with pytest.raises(ValueError, match="lx >= 3"):
        hip.InamuroHIP(hip.SolverParams(lx=2, ly=4, lz=4))
    with pytest.raises(ValueError, match="precision"):
        hip.InamuroHIP(hip.SolverParams(lx=4, ly=4, lz=4), precision=7)
