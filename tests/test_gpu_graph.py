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
