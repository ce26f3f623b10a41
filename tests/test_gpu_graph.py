"""Graph replay of Poisson segments (setUsePoissonGraph) against eager
launches: every field, distribution and iteration count bitwise equal, over
steps that enter the solve in each mode (explicit h after an upload,
equilibrium h between steps, gauged h after a check) and for both precisions,
with and without the dual residual.  Each captured segment is replayed many
times, with host reads and eager work between replays."""
import numpy as np
import pytest

from lbm_amd import hip
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def run(P, precision, graphs, dual, steps=4):
    o = po.Oracle(P)
    o.place_droplet(P.lx / 2, P.ly / 2, P.lz / 2, vz=0.01, profile="tanh", width=2.0)
    sim = hip.InamuroHIP(P, precision=precision)
    sim.setUsePoissonGraph(graphs)
    sim.upload_fields(o.fields())
    sim.setPoisson(check_interval=100, tolerance=1e-3, dual_residual=dual)
    iters = []
    for _ in range(steps):
        iters += sim.step(1)
        sim.mass()  # eager work between replays
    out = sim.downloadFieldsToCPU()
    out.update(sim.download_dists())
    sim.close()
    return iters, out


@pytest.mark.parametrize("precision", [hip.FP64, hip.MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("dual", [False, True], ids=["plain", "dual"])
def test_graph_replay_is_bitwise_eager(precision, dual):
    P = po.Params(lx=20, ly=16, lz=24, DD=8.0, rho_L=100.0)
    print(f"graph vs eager: precision={precision} dual={dual}", flush=True)
    it_g, a = run(P, precision, True, dual)
    it_e, b = run(P, precision, False, dual)
    assert it_g == it_e
    assert max(it_g) > 100  # several segments per solve
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("precision", [hip.FP64, hip.MIXED], ids=["fp64", "mixed"])
def test_every_poisson_launch_shape_gives_the_same_bits(precision):
    """The Poisson kernel's launch shapes (block size and load hint, tuned
    per grid on the solver's own segments) compute every cell the same way:
    fields, distributions and iteration counts are bitwise equal for each
    fixed shape and for the automatic choice -- which, while tuning, runs
    different shapes in successive segments of one solve -- on a grid whose
    rows are not a multiple of any block size (lx = 20) and that enters the
    solve in each mode."""
    P = po.Params(lx=20, ly=16, lz=24, DD=8.0, rho_L=100.0)
    o = po.Oracle(P)
    o.place_droplet(P.lx / 2, P.ly / 2, P.lz / 2, vz=0.01, profile="tanh", width=2.0)
    start = o.fields()
    o.close()
    outs = {}
    ncfg = 3 if precision == hip.FP64 else 4
    for cfg in [-1] + list(range(ncfg)):
        print(f"launch shape {cfg}, precision {precision}", flush=True)
        sim = hip.InamuroHIP(P, precision=precision)
        sim.set_poisson_launch(cfg)
        assert sim.poisson_launch() == cfg  # fixed, or -1 while tuning
        sim.upload_fields(start)
        iters = sim.step(3)
        used = sim.poisson_launch()
        assert used == cfg if cfg >= 0 else -1 <= used < ncfg
        out = sim.downloadFieldsToCPU()
        out.update(sim.download_dists())
        outs[cfg] = (iters, out)
        sim.close()
    it0, ref = outs[0]
    assert max(it0) > 100
    for cfg, (its, out) in outs.items():
        assert its == it0, cfg
        for k in ref:
            assert np.array_equal(out[k], ref[k]), (cfg, k)
    with pytest.raises(ValueError):
        hip.InamuroHIP(P, precision=precision).set_poisson_launch(ncfg)
