"""Grids whose sizes do not fill the kernels' tiles, against the oracle.

The collision and accelerator kernels work on 64x4-cell tiles marching z
chunks, the Poisson push kernel on flat blocks, the adjoint stencils on the
same tiles.  Each shape here leaves partial tiles somewhere:

* 70x9x13 -- two x tiles, the second holding 6 columns; 9 rows (two 4-row
  tiles and one row); 13 layers (a partial 16-layer chunk);
* 3x66x21 -- three columns, so both x mirror images of every row sit in one
  tile; 66 rows (a 2-row last tile); 21 layers;
* 4x128x67 -- 32 tiles of 67 layers: the accelerators' pull kernels march
  2-layer z chunks there, the last one holding a single layer.

Bars are the suite's: FP64 within 1e-12 relative L-inf of the oracle with
identical Poisson iteration counts; mixed precision within 1e-5 of the
oracle on phi, u, v, w, p with identical counts; accelerators within 1e-10
with identical counts (their global sums differ from the oracle's in order
only); the adjoint within 1e-9.
"""
import numpy as np
import pytest

from lbm_amd import hip
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

FIELDS = ("fei", "rho", "u", "v", "w", "p")

# shape -> (DD, droplet centre, frozen iteration schedule or None to converge)
SHAPES = {
    "70x9x13": ((70, 9, 13), 6.0, (35.0, 4.5, 6.5), None),
    "3x66x21": ((3, 66, 21), 2.0, (1.5, 33.0, 10.0), None),
    "4x128x67": ((4, 128, 67), 6.0, (2.0, 64.0, 33.5), [300, 200]),
}


def rel(a, b):
    return float(np.abs(a - b).max()) / max(float(np.abs(b).max()), 1e-300)


def params(name, **kw):
    (lx, ly, lz), dd, _, _ = SHAPES[name]
    return po.Params(lx=lx, ly=ly, lz=lz, DD=dd, **kw)


def oracle(name, P, walls=False):
    _, _, centre, schedule = SHAPES[name]
    o = po.Oracle(P, threads=4)
    if walls:
        o.set_z_walls(True)
        o.set_gravity(-2e-4)
    o.place_droplet(*centre, 0.01, -0.005, 0.002, profile="tanh", width=2.0)
    if schedule:
        o.set_schedule(schedule)
    return o


def gpu(name, P, fields, precision=hip.FP64, walls=False):
    schedule = SHAPES[name][3]
    sim = hip.InamuroHIP(P, precision=precision)
    if walls:
        sim.set_z_walls(True)
        sim.set_gravity(-2e-4)
    sim.upload_fields(fields)
    if schedule:
        sim.set_iteration_schedule(schedule)
    return sim


def run_pair(name, walls=False):
    P = params(name)
    print(f"ragged {name} walls={walls}: FP64 and mixed, 2 steps", flush=True)
    o = oracle(name, P, walls)
    start = o.fields()
    f64 = gpu(name, P, start, walls=walls)
    mixed = gpu(name, P, start, precision=hip.MIXED, walls=walls)
    for step in range(2):
        oi, gi, mi = o.step(1), f64.step(1), mixed.step(1)
        assert gi == oi and mi == oi, (step + 1, oi, gi, mi)
    ref, a, b = o.fields(), f64.downloadFieldsToCPU(), mixed.downloadFieldsToCPU()
    for k in FIELDS:
        assert rel(a[k], ref[k]) <= 1e-12, (k, rel(a[k], ref[k]))
    for k in ("fei", "u", "v", "w", "p"):
        assert rel(b[k], ref[k]) <= 1e-5, (k, rel(b[k], ref[k]))
    d = f64.download_dists()
    for k in ("f", "g", "h"):
        assert rel(d[k], o.get(k)) <= 1e-12, k
    for s in (f64, mixed):
        s.close()
    o.close()


@pytest.mark.parametrize("name", sorted(SHAPES))
def test_periodic_step_on_ragged_grids(name):
    run_pair(name)


@pytest.mark.parametrize("name", ["70x9x13", "3x66x21"])
def test_walls_and_gravity_on_ragged_grids(name):
    run_pair(name, walls=True)


@pytest.mark.parametrize("name", ["70x9x13", "3x66x21"])
def test_dual_residual_on_ragged_grids(name):
    """The kinetic fixed-point residual (k_fixed_point's one-cell neighbour
    reads) with the dual exit, FP64, against the oracle."""
    P = params(name)
    print(f"ragged {name}: dual residual, 2 steps", flush=True)
    o = oracle(name, P)
    start = o.fields()
    o.set_poisson(dual=True)
    sim = gpu(name, P, start)
    sim.setPoisson(100, 1e-3, 0, True)
    for step in range(2):
        oi, gi = o.step(1), sim.step(1)
        assert gi == oi, (step + 1, oi, gi)
        gd, od = sim.getLastPoissonDiagnostics(), o.diag()
        assert gd.fixed_point_evaluated == 1 and gd.converged == 1
        for k in ("fixed_point_h_relative", "fixed_point_p_relative", "fixed_point_relative"):
            g_, o_ = getattr(gd, k), getattr(od, k)
            assert abs(g_ - o_) <= 1e-8 * abs(o_) + 1e-14, (k, g_, o_)
    ref, a = o.fields(), sim.downloadFieldsToCPU()
    for k in FIELDS:
        assert rel(a[k], ref[k]) <= 1e-12, (k, rel(a[k], ref[k]))
    sim.close()
    o.close()


ACCEL = {
    "anderson_two_grid_spatial": dict(variant="onepass", anderson_m1=1, two_grid=1, spatial_diagnostics=1),
    "onepass_w0.8_relax_source_aware": dict(variant="onepass", fixed_point_relax=0.8,
                                            pressure_relax_scale=0.9, source_aware_hh_init=1),
    "scalar_two_grid": dict(variant="scalar", two_grid=1),
    "split_relax_two_grid": dict(variant="split", pressure_relax_scale=0.9, two_grid=1),
}


def oracle_spread(name, P, kw, ref, iterations):
    """How far the oracle's own 2-step result moves when its start velocity is
    nudged by at most 2 ulps per cell (relative L-inf per field)."""
    o = oracle(name, P)
    rng = np.random.default_rng(11)
    for k in ("u", "v", "w"):
        a = o.get(k)
        o.set(k, a * (1.0 + 2.0 ** -52 * rng.integers(-2, 3, a.shape)))
    o.set_poisson_accel(**kw)
    assert o.step(2) == iterations
    out = o.fields()
    o.close()
    return {k: rel(out[k], ref[k]) for k in FIELDS}


@pytest.mark.parametrize("accel", sorted(ACCEL))
@pytest.mark.parametrize("name", sorted(SHAPES))
def test_accelerators_on_ragged_grids(name, accel):
    """Anderson(1) mixing forms alpha = -<R_prev, R - R_prev> / |R - R_prev|^2
    from nearly equal residuals, so its path amplifies rounding: on the
    3x66x21 grid a 2-ulp nudge of the start velocity moves the oracle's own
    result by 2e-9 on phi and 3e-5 on p with the same iteration counts.  The
    GPU's global sums are associated differently from the oracle's, which is
    a perturbation of that size at every iteration, so an Anderson case is
    held to 10x the oracle's own spread under such a nudge (and to 1e-10
    where that spread is smaller); every other case to 1e-10."""
    kw = ACCEL[accel]
    P = params(name)
    print(f"ragged {name} accelerator {accel}: 2 steps", flush=True)
    o = oracle(name, P)
    start = o.fields()
    o.set_poisson_accel(**kw)
    sim = gpu(name, P, start)
    sim.set_poisson_accel(**kw)
    its = []
    for step in range(2):
        oi, gi = o.step(1), sim.step(1)
        assert gi == oi, (step + 1, oi, gi)
        its += oi
    ref, a = o.fields(), sim.downloadFieldsToCPU()
    sim.close()
    o.close()
    bar = dict.fromkeys(FIELDS, 1e-10)
    if kw.get("anderson_m1"):
        spread = oracle_spread(name, P, kw, ref, its)
        print(f"  oracle spread under a 2-ulp nudge: {spread}", flush=True)
        bar = {k: max(1e-10, 10.0 * spread[k]) for k in FIELDS}
    for k in FIELDS:
        assert rel(a[k], ref[k]) <= bar[k], (k, rel(a[k], ref[k]), bar[k])


@pytest.mark.parametrize("name", ["70x9x13", "3x66x21"])
def test_adjoint_on_ragged_grids(name):
    (lx, ly, lz), dd, centre, _ = SHAPES[name]
    P = po.Params(lx=lx, ly=ly, lz=lz, rho_L=10.0, rho_G=1.0, mu_L=0.008, mu_G=8e-5, k_g=1e-4,
                  fei_max=0.093, fei_min=0.016, fei_L=0.093, fei_G=0.016, DD=dd)
    print(f"ragged {name}: adjoint over 3 steps", flush=True)
    o = po.Oracle(P)
    o.place_droplet(*centre, profile="tanh", width=1.5)
    phi0 = o.get("fei")
    target = phi0 + 0.002 * np.random.default_rng(3).standard_normal(phi0.shape)
    sim = hip.InamuroHIP(P)
    sim.setPoisson(check_interval=10)
    J, g, gk = sim.adjoint(phi0, target, 3, checkpoint_interval=2)
    iters = sim.iteration_history()[-3:]
    o.set_poisson(check_interval=10)
    Jo, go, gko = o.adjoint(phi0, target, iters)
    assert abs(J - Jo) <= 1e-12 * Jo
    assert rel(g, go) <= 1e-9, rel(g, go)
    assert abs(gk - gko) <= 1e-9 * abs(gko), (gk, gko)
    sim.close()
    o.close()
