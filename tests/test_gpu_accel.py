"""The reference GPU solver's Poisson accelerators on the HIP path against the
oracle's restatement (GPU/InamuroCUDA.cu:503-1013, 1198-1323, 1816-2349).

Every case starts the GPU and the oracle from the reference's 16x32x32
two-droplet state and runs three steps.  The bar: identical Poisson iteration
counts on every step and phi, rho, u, v, w, p within 1e-10 relative L-inf
(FP64 throughout; only the order of the global sums -- gauge mean, relative
change, Anderson dot products, 4^3 block sums -- differs, and the iteration
does not amplify it beyond that).  The spatial fraction matches to 1e-9.
The Anderson coefficient alpha = -<R_prev, R - R_prev> / |R - R_prev|^2 is
formed from differences of nearly equal residuals late in a solve, so the
order of its two sums moves it by ~1e-9 relative (measured 2.4e-9): its bar
is 1e-6, while the fields it produces stay within the 1e-10 above.
"""
import numpy as np
import pytest

from lbm_amd import hip
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

P = po.Params(lx=16, ly=32, lz=32, DD=8.0)
FIELDS = ("fei", "rho", "u", "v", "w", "p")

CASES = {
    "split_relax_0.85": dict(variant="split", pressure_relax_scale=0.85),
    "fused_boundary_relax_0.9": dict(variant="fused", fused_boundary=1, pressure_relax_scale=0.9),
    "onepass": dict(variant="onepass"),
    "onepass_w0.8_s0.9": dict(variant="onepass", fixed_point_relax=0.8, pressure_relax_scale=0.9),
    "onepass_anderson": dict(variant="onepass", anderson_m1=1),
    "onepass_anderson_w0.3_beta0.25": dict(variant="onepass", anderson_m1=1, fixed_point_relax=0.3,
                                           anderson_beta_max=0.25),
    "scalar": dict(variant="scalar"),
    "scalar_scale3_two_grid": dict(variant="scalar", scalar_source_scale=3.0, two_grid=1),
    "split_two_grid": dict(variant="split", two_grid=1),
    "onepass_two_grid_spatial": dict(variant="onepass", two_grid=1, two_grid_strength=0.3,
                                     spatial_diagnostics=1),
    "source_aware_1.5": dict(variant="split", source_aware_hh_init=1, source_aware_hh_scale=1.5),
    "anderson_source_aware_two_grid": dict(variant="onepass", anderson_m1=1, source_aware_hh_init=1,
                                           two_grid=1),
}


def _oracle_run(accel, steps):
    o = po.Oracle(P, threads=4)
    o.init_two_droplets_sharp()
    start = o.fields()
    o.set_poisson_accel(**accel)
    its, diags = [], []
    for _ in range(steps):
        its += o.step(1)
        diags.append(o.diag())
    out = o.fields()
    o.close()
    return start, its, out, diags


def _gpu_run(start, accel, steps, graphs=True):
    sim = hip.InamuroHIP(P)
    sim.upload_fields(start)
    sim.setUsePoissonGraph(graphs)
    sim.set_poisson_accel(**accel)
    records = []
    sim.set_check_observer(records.append)
    its = sim.step(steps)
    out = sim.downloadFieldsToCPU()
    sim.close()
    return its, out, records


def _rel_linf(a, b):
    scale = max(float(np.abs(b).max()), 1e-300)
    return float(np.abs(a - b).max()) / scale


@pytest.mark.parametrize("name", sorted(CASES))
def test_accelerator_matches_the_oracle(name):
    accel = CASES[name]
    print(f"accelerator case {name}: {accel}", flush=True)
    start, o_its, o_out, o_diags = _oracle_run(accel, 3)
    g_its, g_out, recs = _gpu_run(start, accel, 3)
    assert g_its == o_its, (g_its, o_its)
    for k in FIELDS:
        assert _rel_linf(g_out[k], o_out[k]) <= 1e-10, (k, _rel_linf(g_out[k], o_out[k]))
    # the last check of the last step against the oracle's last check
    last = recs[-1]
    d = o_diags[-1]
    assert last["iteration"] == d.iteration_count
    assert last["relative_error"] == pytest.approx(d.relative_error, rel=1e-9)
    if accel.get("spatial_diagnostics"):
        assert last["block_low_frequency_fraction"] == pytest.approx(d.block_low_frequency_fraction, rel=1e-9)
    else:
        assert np.isnan(last["block_low_frequency_fraction"])
    blocks = accel.get("two_grid") or accel.get("spatial_diagnostics")
    assert (last["block_size"], last["block_count"]) == ((4, 4 * 8 * 8) if blocks else (0, 0))
    if accel.get("anderson_m1") and accel["variant"] == "onepass":
        assert last["anderson_alpha"] == pytest.approx(d.anderson_alpha, rel=1e-6)
    else:
        assert np.isnan(last["anderson_alpha"])


def test_fused_and_fused_boundary_run_the_strict_kernel_bit_for_bit():
    """At unit relaxation the fused variant and the fused boundary compute the
    split values bit for bit, so they keep the strict push kernel."""
    o = po.Oracle(P)
    o.init_two_droplets_sharp()
    start = o.fields()
    o.close()
    its0, out0, _ = _gpu_run(start, {}, 2)
    its1, out1, _ = _gpu_run(start, dict(variant="fused", fused_boundary=1, anderson_m1=1,
                                         fixed_point_relax=0.5), 2)
    assert its0 == its1
    for k in FIELDS:
        assert np.array_equal(out0[k], out1[k]), k


def test_accelerated_graph_replay_equals_eager_launches():
    o = po.Oracle(P)
    o.init_two_droplets_sharp()
    start = o.fields()
    o.close()
    accel = dict(variant="onepass", anderson_m1=1, two_grid=1, spatial_diagnostics=1)
    its0, out0, rec0 = _gpu_run(start, accel, 2, graphs=False)
    its1, out1, rec1 = _gpu_run(start, accel, 2, graphs=True)
    assert its0 == its1
    for k in FIELDS:
        assert np.array_equal(out0[k], out1[k]), k
    assert [r["anderson_alpha"] for r in rec0] == [r["anderson_alpha"] for r in rec1]


def test_reference_setter_names_and_precedence():
    """setUse{Fused,OnePass,Scalar}Poisson keep three flags; scalar beats
    onepass beats fused (GPU/InamuroCUDA.cu:2216-2296)."""
    sim = hip.InamuroHIP(P)
    sim.setUseFusedPoisson(True)
    sim.setUseOnePassPoisson(True)
    assert sim.poisson_accel()["variant"] == hip.POISSON_VARIANTS["onepass"]
    sim.setUseScalarPoisson(True)
    assert sim.poisson_accel()["variant"] == hip.POISSON_VARIANTS["scalar"]
    sim.setUseScalarPoisson(False)
    sim.setUseOnePassPoisson(False)
    assert sim.poisson_accel()["variant"] == hip.POISSON_VARIANTS["fused"]
    sim.setPressureRelaxScale(0.9)
    sim.setPoissonFixedPointRelax(1.1)
    sim.setUsePoissonAndersonM1(True)
    sim.setPoissonAndersonBetaMax(0.5)
    sim.setUsePoissonTwoGridCorrection(True)
    sim.setPoissonTwoGridStrength(0.25)
    sim.setUseSourceAwareHHInit(True)
    sim.setSourceAwareHHScale(2.0)
    sim.setScalarPoissonSourceScale(3.0)
    sim.setUseFusedBoundaryPressure(True)
    sim.setUsePoissonSpatialDiagnostics(True)
    a = sim.poisson_accel()
    assert (a["pressure_relax_scale"], a["fixed_point_relax"], a["anderson_m1"], a["anderson_beta_max"],
            a["two_grid"], a["two_grid_strength"], a["source_aware_hh_init"], a["source_aware_hh_scale"],
            a["scalar_source_scale"], a["fused_boundary"], a["spatial_diagnostics"]) == \
        (0.9, 1.1, 1, 0.5, 1, 0.25, 1, 2.0, 3.0, 1, 1)
    for setter, value, msg in (
            (sim.setPressureRelaxScale, 0.0, "pressure relaxation scale must be positive"),
            (sim.setPoissonFixedPointRelax, -1.0, "Poisson fixed-point relaxation must be positive"),
            (sim.setPoissonAndersonBetaMax, -0.5, "Poisson Anderson beta max must be non-negative"),
            (sim.setPoissonTwoGridStrength, -1.0, "Poisson two-grid strength must be non-negative")):
        with pytest.raises(ValueError, match=msg):
            setter(value)
    # a rejected call changed nothing
    assert sim.poisson_accel() == a
    sim.close()


def test_refusals():
    """The dual residual with an accelerator fails the step with the
    reference's message before anything changes; mixed-precision contexts
    refuse the accelerators; so do z walls and the adjoint."""
    o = po.Oracle(P)
    o.init_two_droplets_sharp()
    start = o.fields()
    o.close()
    sim = hip.InamuroHIP(P)
    sim.upload_fields(start)
    sim.setUsePoissonDualResidual(True)
    sim.set_poisson_accel(variant="onepass")
    with pytest.raises(RuntimeError, match="fixed-point residual requires frozen split"):
        sim.step(1)
    sim.set_poisson_accel(variant="split", spatial_diagnostics=1)
    with pytest.raises(RuntimeError, match="no spatial diagnostics"):
        sim.step(1)
    sim.setUsePoissonDualResidual(False)
    assert sim.step(1) == [800]  # nothing had changed: the reference's first step
    sim.set_poisson_accel(spatial_diagnostics=0, variant="scalar")
    with pytest.raises(RuntimeError, match="strict split Poisson map"):
        sim.adjoint(start["fei"], start["fei"], 1, 1)
    sim.set_poisson_accel(variant="split")
    sim.set_z_walls(True)
    sim.set_poisson_accel(two_grid=1)
    with pytest.raises(RuntimeError, match="periodic z only"):
        sim.step(1)
    sim.close()
    mixed = hip.InamuroHIP(P, precision=hip.MIXED)
    with pytest.raises(ValueError, match="FP64 path"):
        mixed.set_poisson_accel(variant="onepass")
    mixed.set_poisson_accel(variant="split")  # the defaults are accepted
    mixed.close()
