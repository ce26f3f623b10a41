"""The host layer's output files against the reference's own writers.

lbm::writeTecplotBinary and lbm::writeDebugFieldCSV (lbm_amd/host) write the
reference's Tecplot #!TDV101 file (CPU/Inamuro.cpp:865-1013) and debug CSV
(:1015-1048).  Here both the reference (oracle/_ref/ref_probe, compiled from
the reference's sources, calling its own writeResults) and the host layer (a
small driver linked against lbm_amd/liblbm_host.so) write the same fields;
the files must be byte-identical.  Runs where the reference tree exists; no
GPU is touched (the driver calls only the writers).
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "lbm_amd")

DRIVER = r"""
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include "InamuroHIP.hpp"
int main(int argc, char** argv)
{
    const int lx = std::atoi(argv[1]), ly = std::atoi(argv[2]), lz = std::atoi(argv[3]);
    const std::size_t n = static_cast<std::size_t>(lx) * ly * lz;
    std::vector<double> f[6];
    std::FILE* in = std::fopen(argv[4], "rb");
    for (auto& a : f) {
        a.resize(n);
        if (std::fread(a.data(), sizeof(double), n, in) != n) return 2;
    }
    std::fclose(in);
    const std::string dir = argv[5];
    lbm::writeTecplotBinary(dir, 0, lx, ly, lz, f);
    lbm::writeDebugFieldCSV(dir, 0, lx, ly, lz, f);
    return 0;
}
"""

pytestmark = pytest.mark.skipif(not po.ref_available(),
                                reason="oracle/_ref is built only where the reference tree exists")


def _fields():
    P = po.Params(lx=9, ly=7, lz=11, DD=5.0, rho_L=100.0, rho_G=1.0, mu_G=8e-5,
                  fei_max=0.093, fei_min=0.016, fei_L=0.093, fei_G=0.016)
    o = po.Oracle(P)
    o.place_droplet(4.0, 3.5, 5.0, vx=0.02, vy=-0.01, vz=0.03, profile="tanh", width=1.5)
    o.step(2)
    f = o.fields()
    o.close()
    return P, f


def test_tecplot_and_debug_csv_are_byte_identical_to_the_reference(tmp_path):
    P, f = _fields()
    assert np.abs(f["p"]).max() > 0 and np.abs(f["u"]).max() > 0
    init = tmp_path / "fields.bin"
    po.write_init(str(init), f)
    pfile = tmp_path / "params.in"
    pfile.write_text(P.params_in())
    ref_dir, our_dir = tmp_path / "ref", tmp_path / "ours"
    ref_dir.mkdir()
    our_dir.mkdir()
    subprocess.run([po.REF_PROBE, str(pfile), str(init), "0", str(tmp_path / "probe.bin"), "0",
                    str(ref_dir) + "/"], check=True, capture_output=True)
    src, exe = tmp_path / "writer_driver.cpp", tmp_path / "writer_driver"
    src.write_text(DRIVER)
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT}/include", f"-I{ROOT}/lbm_amd/host", str(src),
                    "-o", str(exe), f"-L{LIBDIR}", "-llbm_host", "-llbm_hip",
                    f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath,/opt/rocm/lib"], check=True)
    subprocess.run([str(exe), str(P.lx), str(P.ly), str(P.lz), str(init), str(our_dir) + "/"], check=True)
    for name in ("3D000000000.plt", "debug_000000000.csv"):
        a, b = (our_dir / name).read_bytes(), (ref_dir / name).read_bytes()
        assert len(a) == len(b) and a == b, name
    plt = (ref_dir / "3D000000000.plt").read_bytes()
    assert plt[:8] == b"#!TDV101"
    # header + 9 four-byte words per cell
    assert len(plt) - 36 * P.lx * P.ly * P.lz > 0
    rows = (ref_dir / "debug_000000000.csv").read_text().splitlines()
    assert rows[0] == "x,y,z,u,v,w,rho,fei,p" and len(rows) == 1 + P.lx * P.ly * P.lz
