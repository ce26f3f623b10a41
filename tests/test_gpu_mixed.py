"""Mixed precision (FP32 deviations + FP64 pressure and sums) against the
FP64 oracle.

Bar: relative L-infinity error of phi, u and p <= 1e-5 against
the FP64 reference; the same Poisson iteration counts on these cases.
"""
import numpy as np
import pytest

from lbm_amd import hip
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL_MIXED = 1e-5
FIELDS = ("fei", "rho", "u", "v", "w", "p")
RATIO1000 = dict(rho_L=1000.0, rho_G=1.0, mu_L=0.008, mu_G=8e-5, k_g=1e-4, fei_max=0.093,
                 fei_min=0.016, fei_L=0.093, fei_G=0.016)


def rel_linf(a, b):
    return float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)


def pair(P, init=None, **poisson):
    o = po.Oracle(P)
    if init is None:
        o.init_two_droplets_sharp()
    else:
        o.place_droplet(*init, profile="tanh", width=2.0)
    sim = hip.InamuroHIP(P, precision=hip.MIXED)
    sim.upload_fields(o.fields())
    if poisson:
        o.set_poisson(**poisson)
        sim.setPoisson(poisson.get("check_interval", 100), poisson.get("tol", 1e-3),
                       poisson.get("iter_cap", 0), poisson.get("dual", False))
    return o, sim


@pytest.mark.parametrize("case", ["smoke", "ratio1000"])
def test_mixed_tracks_the_fp64_oracle(case):
    if case == "smoke":
        P, init, steps = po.Params(lx=16, ly=32, lz=32, DD=8.0), None, 3
    else:
        P, init, steps = po.Params(lx=24, ly=24, lz=24, DD=12.0, **RATIO1000), (12.0, 12.0, 12.0), 3
    o, sim = pair(P, init)
    for s in range(steps):
        gi, oi = sim.step(1), o.step(1)
        assert gi == oi, (s + 1, gi, oi)
        a, b = sim.downloadFieldsToCPU(), o.fields()
        for k in FIELDS:
            err = rel_linf(a[k], b[k])
            assert err <= TOL_MIXED, (s + 1, k, err)
    d = sim.download_dists()
    for k in ("f", "g", "h"):
        assert rel_linf(d[k], o.get(k)) <= TOL_MIXED, k


def test_mixed_mass_and_dual_residual():
    P = po.Params(lx=16, ly=32, lz=32, DD=8.0)
    o, sim = pair(P, dual=True)
    assert sim.step(2) == o.step(2)
    gd, od = sim.getLastPoissonDiagnostics(), o.diag()
    assert gd.fixed_point_evaluated == 1 and gd.converged == 1
    assert abs(gd.fixed_point_h_relative - od.fixed_point_h_relative) <= 1e-3 * od.fixed_point_h_relative
    m, ref = sim.mass(), float(np.sum(o.get("fei")))
    assert abs(m - ref) <= 1e-6 * ref


def test_mixed_upload_dists_round_trip():
    P = po.Params(lx=12, ly=10, lz=14, DD=6.0)
    o = po.Oracle(P)
    o.init_two_droplets_sharp()
    o.step(1)
    sim = hip.InamuroHIP(P, precision=hip.MIXED)
    sim.upload_fields(o.fields())
    sim.upload_dists(f=o.get("f"), g=o.get("g"), h=None, first_step=False)
    d = sim.download_dists(("f", "g"))
    for k in ("f", "g"):
        assert rel_linf(d[k], o.get(k)) <= 1e-6, k


def test_mixed_over_600_steps_against_the_oracle_fixture():
    """The 64^3 static droplet (density ratio 1000) over 600 steps, mixed
    precision against the oracle's FP64 run (static_droplet_64_oracle600,
    4,096 sampled cells plus full-grid sum|.| and max|.|).

    The FP64 trajectory is not what limits the bar: started from phi and rho
    rounded to FP32 it stays within 3e-7 of the oracle over these 600 steps.
    With the lattice weights multiplied exactly (a hi + lo pair per weight),
    the mixed path stays within 1e-5 on every field while the spurious gas
    current grows from |u| = 0.02 to 0.15 (measured against the FP64 run
    every 25 steps: phi <= 3.5e-7, p <= 1.3e-6, u, v, w <= 6.3e-6; at step
    600 u 2.2e-6); with the FP32-rounded weights, whose common bias enters
    the equilibrium moments on every step, u reached 1.8e-5.  Bars: 1e-5 on
    every field, identical Poisson iteration counts on all 600 steps."""
    import json
    import os
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    rec = json.load(open(os.path.join(gold, "static_droplet_64_oracle600.json")))
    samples = np.load(os.path.join(gold, "static_droplet_64_oracle600.npz"))
    P = po.Params(**rec["params"])
    o = po.Oracle(P)
    o.place_droplet(*rec["init"]["center"], profile=rec["init"]["kind"], width=rec["init"]["width"])
    start = o.fields()
    o.close()
    idx = samples["index"]
    bars = dict.fromkeys(FIELDS, 1e-5)

    def errors(sim):
        a = sim.downloadFieldsToCPU()
        out = {}
        for k in FIELDS:
            ref = rec["fields"][k]
            got = a[k].reshape(-1)
            out[k] = max(float(np.max(np.abs(got[idx] - samples[k]))) / ref["max_abs"],
                         abs(float(np.max(np.abs(got))) - ref["max_abs"]) / ref["max_abs"],
                         abs(float(np.sum(np.abs(got))) - ref["sum_abs"]) / ref["sum_abs"])
        return out

    print("mixed, 600 steps of the 64^3 droplet", flush=True)
    mixed = hip.InamuroHIP(P, precision=hip.MIXED)
    mixed.upload_fields(start)
    assert mixed.step(600) == rec["iterations"]
    em = errors(mixed)
    for k in FIELDS:
        assert em[k] <= bars[k], (k, em[k])

    print("fp64 from FP32-rounded phi and rho, 600 steps", flush=True)
    fp = hip.InamuroHIP(P, precision=hip.FP64)
    fp.upload_fields({k: (v.astype(np.float32).astype(np.float64) if k in ("fei", "rho") else v)
                      for k, v in start.items()})
    assert fp.step(600) == rec["iterations"]
    ef = errors(fp)
    assert max(ef.values()) <= 1e-6, ef
