"""Writes the long 64^3 static-droplet fixture from the oracle alone:

    python tests/golden/make_oracle_long.py [steps] [threads]

The 64^3 static droplet (tanh R=16, density ratio 1000) for `steps` steps
(default 20), the oracle's per-cell loops on `threads` workers (default 1;
its global sums stay sequential, so the output does not depend on it).  A
step whose solve fails (the model going unstable) ends the run; the fixture
then holds the steps before it and records the failure.  The compiled
reference takes over an hour for the first three
steps of this case, so only those three are reference-generated
(static_droplet_64.json, where the oracle is asserted bit-identical to it); this
fixture continues the same trajectory with the oracle, whose first three
steps it re-checks against static_droplet_64.json.  Same schema as make_golden.py's
fixtures, with "generated_by": "oracle".
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def main(argv):
    steps = int(argv[1]) if len(argv) > 1 else 20
    threads = int(argv[2]) if len(argv) > 2 else 1
    base = json.load(open(os.path.join(OUT, "static_droplet_64.json")))
    P = po.Params(**base["params"])
    o = po.Oracle(P, threads=threads)
    cx, cy, cz = base["init"]["center"]
    o.place_droplet(cx, cy, cz, profile=base["init"]["kind"], width=base["init"]["width"])
    init_sha = {k: sha(v) for k, v in o.fields().items()}
    assert init_sha == base["init_sha256"]
    iters, step_sha = [], []
    t0 = time.time()
    diverged = None
    for s in range(steps):
        try:
            iters += o.step(1)
        except po.OracleError as e:
            # the step's solve failed (the model went unstable): keep the
            # steps before it
            diverged = dict(step=s + 1, error=str(e))
            print(f"step {s + 1}: {e}", flush=True)
            break
        last = o.fields()
        step_sha.append({k: sha(v) for k, v in last.items()})
        if s < base["steps"]:
            assert iters[s] == base["iterations"][s] and step_sha[s] == base["step_sha256"][s], s
        print(f"step {s + 1}: {iters[-1]} iterations, {time.time() - t0:.0f} s", flush=True)
    steps = len(iters)
    final = last  # the fields after the last step that completed
    n = P.lx * P.ly * P.lz
    idx = np.sort(np.random.default_rng(20261203).choice(n, size=4096, replace=False))
    rec = dict(case=f"static_droplet_64_oracle{len(step_sha)}", generated_by="oracle", diverged=diverged,
               params=base["params"], init=base["init"], steps=steps, iterations=iters,
               oracle_seconds=time.time() - t0, fields={}, init_sha256=init_sha, step_sha256=step_sha,
               sample_index_seed=20261203)
    samples = {"index": idx.astype(np.int64)}
    for k in po.FIELD_NAMES:
        a = final[k].reshape(-1)
        rec["fields"][k] = dict(sha256=sha(final[k]), sum=float(a.sum()), sum_abs=float(np.abs(a).sum()),
                                max_abs=float(np.abs(a).max()), min=float(a.min()), max=float(a.max()))
        samples[k] = a[idx].copy()
    name = f"static_droplet_64_oracle{len(step_sha)}"
    json.dump(rec, open(os.path.join(OUT, name + ".json"), "w"), indent=1)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **samples)
    print(f"wrote {name}: iterations {iters}", flush=True)


if __name__ == "__main__":
    main(sys.argv)
