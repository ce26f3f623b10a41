"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Run in a container that has the reference tree (it builds oracle/_ref with
`make -C oracle ref`):

    python tests/golden/make_golden.py [case ...]

For every case it
  1. builds the initial fields (the reference's own constructor for the
     two-droplet case; the oracle's restated GPU-tree place_droplet for tanh
     droplets, itself checked against oracle/_ref/ref_probe_init),
  2. runs oracle/_ref/ref_probe (CPU/Inamuro.cpp compiled from the reference
     sources) for the case's steps,
  3. runs the oracle on the same input and asserts bit-identical fields and
     identical per-step Poisson iteration counts,
  4. writes <case>.json (parameters, iteration counts, SHA-256 of each final
     float64 field, sums, sampled cells as hex floats) and <case>.npz (sampled
     values, for the GPU parity tests).

Only data is committed: no reference source travels with the fixtures.
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

# Config-0 physics (64^3 static droplet): density ratio 1000, viscosity ratio
# 100, kappa_f = 0.5, kappa_g = 1e-4, phi_L = 0.093, phi_G = 0.016 used both
# as the initial bounds (fei_max/fei_min) and as the rho-map thresholds
# (fei_L/fei_G).  mu_L keeps the reference default 0.008.
C0 = dict(rho_L=1000.0, rho_G=1.0, mu_L=0.008, mu_G=8e-5, tauf=1.0, taug=1.0, k_f=0.5,
          k_g=1e-4, T=0.035, a=1.0, b=6.7, fei_max=0.093, fei_min=0.016, fei_L=0.093,
          fei_G=0.016)

CASES = {
    # GPU/params.in grid and physics, the reference's own two-droplet init
    # (CPU/Inamuro.cpp:550-587): the reference's `make run` smoke case.
    "smoke_16x32x32": dict(params=dict(lx=16, ly=32, lz=32, DD=8.0), init="two_droplets_sharp",
                           steps=5, samples=1024),
    # Config-0 physics on a small grid: one tanh droplet R=6 at the centre.
    "droplet_24_ratio1000": dict(params=dict(lx=24, ly=24, lz=24, DD=12.0, **C0),
                                 init=dict(kind="tanh", center=(12.0, 12.0, 12.0), width=2.0),
                                 steps=3, samples=1024),
    # Same physics, non-cubic grid with the droplet off-centre in x so that
    # it meets neither wall symmetrically.
    "droplet_20x16x28_offset": dict(params=dict(lx=20, ly=16, lz=28, DD=10.0, **C0),
                                    init=dict(kind="tanh", center=(8.5, 7.0, 15.0), width=2.0),
                                    steps=3, samples=1024),
    # Config 0 at full size, first three steps (100, then tens of thousands of
    # Poisson iterations per step).
    "static_droplet_64": dict(params=dict(lx=64, ly=64, lz=64, DD=32.0, **C0),
                       init=dict(kind="tanh", center=(32.0, 32.0, 32.0), width=2.0),
                       steps=3, samples=4096),
}


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def make_init(case, P: po.Params, tmp: str):
    """Initial fields and the probe's INIT argument ("-" = reference's own)."""
    spec = case["init"]
    o = po.Oracle(P)
    if spec == "two_droplets_sharp":
        o.init_two_droplets_sharp()
        return o.fields(), "-"
    cx, cy, cz = spec["center"]
    o.place_droplet(cx, cy, cz, profile=spec["kind"], width=spec["width"])
    fields = o.fields()
    # Pin the restated tanh initialisation against the GPU-tree constructor.
    pfile = os.path.join(tmp, "init.in")
    extra = (f"init_mode=single_droplet\ninit_profile={spec['kind']}\ninit_velocity=0\n"
             f"interface_width={spec['width']!r}\ninit_center_x={cx!r}\ninit_center_y={cy!r}\n"
             f"init_center_z={cz!r}\n")
    with open(pfile, "w") as fh:
        fh.write(P.params_in(extra))
    out = os.path.join(tmp, "init.bin")
    subprocess.run([po.REF_PROBE_INIT, pfile, out], check=True)
    ref_init = po.read_init_probe_output(out)
    for k in po.FIELD_NAMES:
        if not np.array_equal(ref_init[k].view(np.uint64), fields[k].view(np.uint64)):
            raise AssertionError(f"tanh init differs from the GPU-tree constructor in {k}")
    ipath = os.path.join(tmp, "fields.bin")
    po.write_init(ipath, fields)
    return fields, ipath


def run_case(name: str, case: dict, reuse: str = None) -> None:
    """reuse: an existing ref_probe output file for this case (same
    parameters, initial fields and step count), to skip a long probe run."""
    P = po.Params(**case["params"])
    steps = case["steps"]
    with tempfile.TemporaryDirectory() as tmp:
        init_fields, init_arg = make_init(case, P, tmp)
        pfile = os.path.join(tmp, "params.in")
        with open(pfile, "w") as fh:
            fh.write(P.params_in())
        out = reuse or os.path.join(tmp, "out.bin")
        t0 = time.time()
        if not reuse:
            subprocess.run([po.REF_PROBE, pfile, init_arg, str(steps), out, "0"], check=True)
        t_ref = time.time() - t0
        iters, step_ms, ref = po.read_probe_output(out)
        if reuse:
            t_ref = float(sum(step_ms)) / 1e3
            assert len(iters) == steps, "reused probe output has a different step count"

    o = po.Oracle(P, threads=1)
    o.set("fei", init_fields["fei"]); o.set("rho", init_fields["rho"])
    o.set("u", init_fields["u"]); o.set("v", init_fields["v"]); o.set("w", init_fields["w"])
    o.set("p", init_fields["p"])
    t0 = time.time()
    oit, step_sha = [], []
    for _ in range(steps):
        oit += o.step(1)
        step_sha.append({k: sha(v) for k, v in o.fields().items()})
    t_orc = time.time() - t0
    if list(map(int, iters)) != list(map(int, oit)):
        raise AssertionError(f"{name}: iteration counts differ ref={iters} oracle={oit}")
    final = o.fields()
    for k in po.FIELD_NAMES:
        if not np.array_equal(final[k].view(np.uint64), ref[k].view(np.uint64)):
            raise AssertionError(f"{name}: oracle differs from reference in {k}")

    n = P.lx * P.ly * P.lz
    idx = np.sort(np.random.default_rng(20261203).choice(n, size=min(case["samples"], n), replace=False))
    rec = {
        "case": name,
        "params": {k: getattr(P, k) for k in po.param_names()},
        "init": case["init"],
        "steps": steps,
        "iterations": [int(i) for i in iters],
        "reference_ms_per_step": [float(x) for x in step_ms],
        "oracle_seconds": t_orc,
        "reference_seconds": t_ref,
        "fields": {},
        "init_sha256": {k: sha(init_fields[k]) for k in po.FIELD_NAMES},
        "step_sha256": step_sha,
        "sample_index_seed": 20261203,
    }
    samples = {"index": idx.astype(np.int64)}
    for k in po.FIELD_NAMES:
        a = ref[k].reshape(-1)
        rec["fields"][k] = {
            "sha256": sha(ref[k]),
            "sum": float(np.sum(a)),
            "sum_abs": float(np.sum(np.abs(a))),
            "max_abs": float(np.max(np.abs(a))),
            "min": float(np.min(a)),
            "max": float(np.max(a)),
            "first_samples_hex": [float(x).hex() for x in a[idx[:8]]],
        }
        samples[k] = a[idx].copy()
    with open(os.path.join(OUT, name + ".json"), "w") as fh:
        json.dump(rec, fh, indent=1)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **samples)
    print(f"{name}: iterations {rec['iterations']} ref {t_ref:.1f}s oracle {t_orc:.1f}s -> bit-identical",
          flush=True)


def main(argv):
    if not po.ref_available():
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "all", "ref"], check=True)
    args = argv[1:]
    reuse = None
    if "--reuse" in args:
        i = args.index("--reuse")
        reuse = args[i + 1]
        del args[i:i + 2]
    names = args or [k for k in CASES if k != "static_droplet_64"]
    for name in names:
        run_case(name, CASES[name], reuse)


if __name__ == "__main__":
    main(sys.argv)
