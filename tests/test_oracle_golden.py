"""The oracle reproduces the compiled reference bit for bit.

The fixtures in tests/golden/ were written by tests/golden/make_golden.py
from oracle/_ref/ref_probe (the reference's CPU/Inamuro.cpp compiled from its
own sources).  Here the oracle replays each case and must produce the same
per-step Poisson iteration counts and the same SHA-256 of every float64
field after every step.  Where oracle/_ref is built (only where the reference
tree exists) the probe is also run live on a short case.
"""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# (case, steps replayed here) -- sized so the CPU suite stays short.
CASES = [("smoke_16x32x32", 5), ("droplet_20x16x28_offset", 2), ("droplet_24_ratio1000", 2),
         ("static_droplet_64", 1)]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def oracle_for(rec):
    P = po.Params(**rec["params"])
    o = po.Oracle(P)
    if rec["init"] == "two_droplets_sharp":
        o.init_two_droplets_sharp()
    else:
        cx, cy, cz = rec["init"]["center"]
        o.place_droplet(cx, cy, cz, profile=rec["init"]["kind"], width=rec["init"]["width"])
    return o


@pytest.mark.parametrize("case,steps", CASES)
def test_oracle_replays_reference_fixture(case, steps):
    rec = json.load(open(os.path.join(GOLDEN, case + ".json")))
    o = oracle_for(rec)
    assert {k: sha(v) for k, v in o.fields().items()} == rec["init_sha256"]
    for s in range(steps):
        assert o.step(1) == [rec["iterations"][s]], f"step {s + 1}"
        assert {k: sha(v) for k, v in o.fields().items()} == rec["step_sha256"][s], f"step {s + 1}"
    if steps == rec["steps"]:
        samples = np.load(os.path.join(GOLDEN, case + ".npz"))
        for k in po.FIELD_NAMES:
            assert np.array_equal(o.get(k).reshape(-1)[samples["index"]], samples[k])


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref is built only where the reference tree exists")
def test_oracle_matches_live_reference_probe(tmp_path):
    P = po.Params(lx=12, ly=10, lz=14, DD=6.0, rho_L=20.0)
    pfile = tmp_path / "p.in"
    pfile.write_text(P.params_in())
    out = tmp_path / "out.bin"
    subprocess.run([po.REF_PROBE, str(pfile), "-", "3", str(out), "1"], check=True)
    iters, _, ref = po.read_probe_output(str(out))
    o = po.Oracle(P)
    o.init_two_droplets_sharp()
    assert o.step(3) == list(map(int, iters))
    for k in list(po.FIELD_NAMES) + ["f", "g", "h"]:
        assert np.array_equal(o.get(k).view(np.uint64), ref[k].view(np.uint64)), k


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref is built only where the reference tree exists")
def test_tanh_initialisation_matches_gpu_tree_constructor(tmp_path):
    P = po.Params(lx=14, ly=12, lz=16, DD=9.0, fei_max=0.093, fei_min=0.016)
    cx, cy, cz, wdt, vz = 6.5, 5.25, 8.0, 1.5, 0.01
    pfile = tmp_path / "p.in"
    pfile.write_text(P.params_in(
        f"init_mode=single_droplet\ninit_profile=tanh\ninit_velocity={vz!r}\ninterface_width={wdt!r}\n"
        f"init_center_x={cx!r}\ninit_center_y={cy!r}\ninit_center_z={cz!r}\n"))
    out = tmp_path / "init.bin"
    subprocess.run([po.REF_PROBE_INIT, str(pfile), str(out)], check=True)
    ref = po.read_init_probe_output(str(out))
    o = po.Oracle(P)
    o.place_droplet(cx, cy, cz, vz=vz, profile="tanh", width=wdt)
    for k in po.FIELD_NAMES:
        assert np.array_equal(o.get(k).view(np.uint64), ref[k].view(np.uint64)), k


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref is built only where the reference tree exists")
@pytest.mark.parametrize("shape,dd,centre,steps", [((3, 66, 21), 2.0, (1.5, 33.0, 10.0), 2),
                                                   ((70, 9, 13), 6.0, (35.0, 4.5, 6.5), 1)])
def test_oracle_matches_live_reference_on_ragged_grids(tmp_path, shape, dd, centre, steps):
    """The grids of the GPU's partial-tile tests (tests/test_gpu_ragged.py),
    with a moving tanh droplet uploaded into the reference: every field and
    population bit for bit, the same iteration counts (3x66x21: 2000, 2300;
    70x9x13: 7500)."""
    lx, ly, lz = shape
    P = po.Params(lx=lx, ly=ly, lz=lz, DD=dd)
    o = po.Oracle(P)
    o.place_droplet(*centre, 0.01, -0.005, 0.002, profile="tanh", width=2.0)
    pfile, init, out = tmp_path / "p.in", tmp_path / "init.bin", tmp_path / "out.bin"
    pfile.write_text(P.params_in())
    po.write_init(str(init), o.fields())
    subprocess.run([po.REF_PROBE, str(pfile), str(init), str(steps), str(out), "1"], check=True)
    iters, _, ref = po.read_probe_output(str(out))
    assert o.step(steps) == list(map(int, iters))
    for k in list(po.FIELD_NAMES) + ["f", "g", "h"]:
        assert np.array_equal(o.get(k).view(np.uint64), ref[k].view(np.uint64)), k
