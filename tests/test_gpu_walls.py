"""No-slip z walls and gravity (the falling-droplet extension) through the HIP
step against the oracle's restatement: FP64 to 1e-12 relative L-inf with
identical iteration counts, mixed precision within 1e-5 of FP64, a
boundary change after graphs were captured replays freshly captured ones,
and the adjoint with walls and gravity matches the oracle's to 1e-9."""
import dataclasses

import numpy as np
import pytest

from lbm_amd import hip
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

FIELDS = ("fei", "rho", "u", "v", "w", "p")


def params():
    return po.Params(lx=16, ly=12, lz=28, DD=8.0, rho_L=100.0, rho_G=1.0, mu_L=0.008, mu_G=8e-5,
                     k_f=0.5, k_g=1e-4, T=0.035, a=1.0, b=6.7, fei_max=0.093, fei_min=0.016,
                     fei_L=0.093, fei_G=0.016)


def rel(a, b):
    return float(np.abs(a - b).max()) / max(float(np.abs(b).max()), 1e-300)


def falling(P, gz=-2e-4):
    o = po.Oracle(P)
    o.set_z_walls(True)
    o.set_gravity(gz)
    o.place_droplet(8.0, 6.0, 18.0, profile="tanh", width=1.5)
    return o


def gpu_sim(P, fields, precision=hip.FP64, gz=-2e-4):
    sim = hip.InamuroHIP(P, precision=precision)
    sim.set_z_walls(True)
    sim.set_gravity(gz)
    sim.upload_fields(fields)
    return sim


def test_falling_droplet_fp64_and_mixed_match_the_oracle():
    P = params()
    print("walls+gravity: 16x12x28, 3 steps", flush=True)
    o = falling(P)
    init = o.fields()
    gpu = gpu_sim(P, init)
    mx = gpu_sim(P, init, precision=hip.MIXED)
    for step in range(3):
        oi, gi, mi = o.step(1), gpu.step(1), mx.step(1)
        assert gi == oi and mi == oi, (step, oi, gi, mi)
    ref, a, b = o.fields(), gpu.downloadFieldsToCPU(), mx.downloadFieldsToCPU()
    for k in FIELDS:
        assert rel(a[k], ref[k]) <= 1e-12, (k, rel(a[k], ref[k]))
    for k in ("fei", "u", "v", "w", "p"):
        assert rel(b[k], a[k]) <= 1e-5, (k, rel(b[k], a[k]))
    d = gpu.download_dists()
    for k in ("f", "g", "h"):
        assert rel(d[k], o.get(k)) <= 1e-12, k
    # the droplet is falling
    assert a["w"][6:12, 4:8, 6:10].mean() < 0


def test_boundary_change_recaptures_graphs():
    P = params()
    o = falling(P)
    init = o.fields()
    a = hip.InamuroHIP(P)
    a.upload_fields(init)
    a.step(1)  # segments captured with periodic z
    a.set_z_walls(True)
    a.set_gravity(-2e-4)
    a.upload_fields(init)
    b = gpu_sim(P, init)
    assert a.step(2) == b.step(2)
    fa, fb = a.downloadFieldsToCPU(), b.downloadFieldsToCPU()
    for k in FIELDS:
        assert np.array_equal(fa[k], fb[k]), k


def test_extension_rejects_dual_residual():
    P = params()
    sim = hip.InamuroHIP(P)
    sim.set_z_walls(True)
    with pytest.raises(ValueError, match="reference boundaries"):
        sim.setPoisson(check_interval=100, tolerance=1e-3, dual_residual=True)


def adjoint_params():
    return po.Params(lx=10, ly=8, lz=12, DD=6.0, rho_L=10.0, rho_G=1.0, mu_L=0.008, mu_G=8e-5,
                     k_f=0.5, k_g=1e-4, T=0.035, a=1.0, b=6.7, fei_max=0.093, fei_min=0.016,
                     fei_L=0.093, fei_G=0.016)


@pytest.mark.parametrize("objective", ["misfit", "centroid_z"])
def test_adjoint_with_walls_and_gravity_matches_the_oracle(objective):
    """The bounce-back, wall-stencil, wall-correct and gravity transposes of
    the HIP adjoint against the oracle's (itself pinned by central
    differences), for the misfit and the centroid height; checkpointing every
    2 of 4 steps."""
    P = adjoint_params()
    print(f"adjoint walls+gravity: {objective}", flush=True)
    o = po.Oracle(P)
    o.set_z_walls(True)
    o.set_gravity(-2e-3)
    o.set_poisson(check_interval=10)
    o.set_objective(objective)
    o.place_droplet(4.0, 4.0, 3.5, profile="tanh", width=1.5)
    phi0 = o.get("fei")
    target = phi0 + 0.002 * np.random.default_rng(3).standard_normal(phi0.shape)
    sim = hip.InamuroHIP(P)
    sim.set_z_walls(True)
    sim.set_gravity(-2e-3)
    sim.setPoisson(check_interval=10)
    sim.set_adjoint_objective(objective)
    J, g, gk = sim.adjoint(phi0, target if objective == "misfit" else None, 4, checkpoint_interval=2)
    iters = sim.iteration_history()[-4:]
    Jo, go, gko = o.adjoint(phi0, target, iters)
    assert abs(J - Jo) <= 1e-12 * abs(Jo), (J, Jo)
    assert rel(g, go) <= 1e-9, rel(g, go)
    assert abs(gk - gko) <= 1e-9 * abs(gko), (gk, gko)
    if objective == "centroid_z":
        # the GPU's own frozen forward map, central differences in k_g
        hk = 1e-6
        vals = []
        for kg in (P.k_g + hk, P.k_g - hk):
            x = hip.InamuroHIP(dataclasses.replace(P, k_g=kg))
            x.set_z_walls(True)
            x.set_gravity(-2e-3)
            x.setPoisson(check_interval=10)
            x.set_adjoint_objective(objective)
            vals.append(x.objective(phi0, None, iters))
        fd = (vals[0] - vals[1]) / (2 * hk)
        assert abs(fd - gk) <= 1e-4 * abs(fd), (fd, gk)
