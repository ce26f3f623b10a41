"""Builds lbm_amd/liblbm_hip.so for gfx950 with hipcc, in-tree.

Every HIP translation unit is compiled with packed-FP32 code generation
disabled.  The FP64 kernels and the host orchestration use
-ffp-contract=off so that no multiply-add is fused (the reference CPU build
has none); the mixed-precision kernels allow contraction.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblbm_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

# (source, extra flags)
UNITS = [
    ("kernels_f64.hip", ["-ffp-contract=off"]),
    ("kernels_mixed.hip", ["-ffp-contract=fast"]),
    ("kernels_adjoint.hip", ["-ffp-contract=off"]),
    ("kernels_accel.hip", ["-ffp-contract=off"]),
    ("capi.cpp", ["-x", "hip", "-ffp-contract=off"]),
    ("noalloc.cpp", ["-x", "hip", "-ffp-contract=off"]),
]


def _compile(src: str, extra: list) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(UNITS))) as ex:
        objs = list(ex.map(lambda u: _compile(*u), UNITS))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}"] + objs + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # Calibration program for the memory-traffic counters (not linked into
    # the library).
    probe = os.path.join(HERE, "copy_probe")
    cmd = [HIPCC] + COMMON + [os.path.join(CSRC, "copy_probe.hip"), "-o", probe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for copy_probe.hip:\n{r.stderr}")
    build_host()
    if verbose:
        print(f"built {LIB}")
    return LIB


HOST = os.path.join(HERE, "host")
HOST_SOURCES = ["InamuroHIP.cpp", "InamuroSolver.cpp"]


def build_host() -> None:
    """The C++ host layer (lbm_amd/host) and the lbm_gpu runner, built with
    g++ against the C ABI only (no HIP headers needed)."""
    cxx = os.environ.get("CXX", "g++")
    flags = ["-std=c++17", "-O2", "-Wall", "-Wextra", "-fPIC", f"-I{os.path.dirname(HERE)}/include"]
    link = [f"-L{HERE}", "-llbm_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
    srcs = [os.path.join(HOST, f) for f in HOST_SOURCES]
    lib = os.path.join(HERE, "liblbm_host.so")
    cmds = [
        [cxx] + flags + ["-shared"] + srcs + ["-o", lib] + link,
        [cxx] + flags + [os.path.join(HOST, "lbm_gpu.cpp")] + srcs + ["-o", os.path.join(HERE, "lbm_gpu")] + link,
        [cxx] + flags + [os.path.join(HOST, "lbm_sim.cpp")] + srcs + ["-o", os.path.join(HERE, "lbm_sim")] + link,
    ]
    for cmd in cmds:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host build failed:\n{' '.join(cmd)}\n{r.stderr}")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
