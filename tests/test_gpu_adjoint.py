"""HIP discrete adjoint (lbm_adjoint) against the oracle's adjoint and against
central finite differences of the GPU's own frozen-schedule forward map.

Bars: J and the gradients agree with the oracle to 1e-9 relative (both are
the exact transpose of the same FP64 computation; they differ only in the
association order of global sums); the checkpoint interval and the recorded
iterate budget change how much is recomputed, never the result (bitwise);
the directional derivative matches central differences (h = 1e-6 in phi0,
1e-6 in k_g) to 1e-4 relative.
"""
import dataclasses
import sys

import numpy as np
import pytest

from lbm_amd import hip
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ADJ_TOL = 1e-9
FD_TOL = 1e-4


def params(lx=12, ly=10, lz=10, rho_L=10.0, DD=6.0):
    return po.Params(lx=lx, ly=ly, lz=lz, rho_L=rho_L, rho_G=1.0, mu_L=0.008, mu_G=8e-5, tauf=1.0,
                     taug=1.0, k_f=0.5, k_g=1e-4, T=0.035, a=1.0, b=6.7, fei_max=0.093, fei_min=0.016,
                     fei_L=0.093, fei_G=0.016, DD=DD)


def case(P, cx=None, seed=0):
    o = po.Oracle(P)
    o.place_droplet(P.lx / 2 if cx is None else cx, P.ly / 2, P.lz / 2, profile="tanh", width=1.5)
    phi0 = o.get("fei")
    rng = np.random.default_rng(seed)
    target = phi0 + 0.002 * rng.standard_normal(phi0.shape)
    return phi0, target, rng


def rel(a, b):
    return float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)


def announce(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("rho_L,ci,steps,cx,lx", [(10.0, 10, 3, None, 12), (10.0, 10, 4, 1.5, 12),
                                                (1000.0, 100, 3, None, 10)])
def test_adjoint_matches_oracle(rho_L, ci, steps, cx, lx):
    P = params(lx=lx, ly=8 if lx == 10 else 10, lz=8 if lx == 10 else 10, rho_L=rho_L)
    announce(f"adjoint vs oracle: rho_L={rho_L} check={ci} steps={steps} cx={cx}")
    phi0, target, _ = case(P, cx)
    sim = hip.InamuroHIP(P)
    sim.setPoisson(check_interval=ci)
    J, g, gk = sim.adjoint(phi0, target, steps, checkpoint_interval=2)
    iters = sim.iteration_history()[-steps:]
    o = po.Oracle(P)
    o.set_poisson(check_interval=ci)
    Jo, go, gko = o.adjoint(phi0, target, iters)
    assert abs(J - Jo) <= 1e-12 * Jo
    assert rel(g, go) <= ADJ_TOL, rel(g, go)
    assert abs(gk - gko) <= ADJ_TOL * abs(gko), (gk, gko)
    st = sim.adjoint_stats()
    assert st.steps == steps and st.forward_iterations == sum(iters)
    # the last step's solve carries no adjoint (J reads phi_T, formed before
    # it) and is neither recomputed nor reversed; the gradient above still
    # matches the oracle, which reverses it
    assert st.adjoint_iterations == sum(iters[:-1])
    # the reverse solves' convergence record is a relative change and an
    # iteration count within the solves
    assert 0.0 <= st.final_relative_change <= 1.0
    assert -1 <= st.iterations_to_tolerance <= max(iters)


def test_recomputation_does_not_change_the_result():
    """Checkpoint every step / every 2 / once, and iterate budgets that split
    long solves into blocks restarted from recorded or recomputed snapshots:
    all bitwise equal."""
    P = params()
    announce("adjoint recomputation invariance")
    phi0, target, _ = case(P)
    sim = hip.InamuroHIP(P)
    sim.setPoisson(check_interval=10)
    J0, g0, k0 = sim.adjoint(phi0, target, 4, checkpoint_interval=4)
    iters = sim.iteration_history()[-4:]
    assert max(iters) > 10
    n = P.lx * P.ly * P.lz
    m = max(iters)

    def budget(L):
        # three quarters of the budget hold L recorded iterates (15n each)
        # and the block's check store; a quarter is the snapshot pool
        return int(8 * 4 / 3 * (L + 1) * (15 * n + n // 10 + 1))

    # L ~ 10, 20: the pool holds too few snapshots for the long solves, which
    # are reversed by binomial checkpointing over their blocks (several
    # recompute passes); L ~ 50: all snapshots recorded in the forward pass.
    for K, L in ((1, 0), (2, 0), (4, 20), (3, 20), (4, 10), (1, 10), (4, 50)):
        sim.set_adjoint_memory(budget(L) if L else 0)
        J, g, k = sim.adjoint(phi0, target, 4, checkpoint_interval=K)
        st = sim.adjoint_stats()
        if L:
            assert st.store_slots < m and st.recomputed_iterations >= sum(iters[:-1]), (st.store_slots, m)
        assert J == J0 and k == k0 and np.array_equal(g, g0), (K, L)
    # Given the iteration counts, the budget split is chosen for the longest
    # solve, and a solve whose block snapshots do not fit keeps the top level
    # of its binomial checkpointing from the forward pass.
    for K, L in ((4, 10), (2, 20), (1, 12)):
        sim.set_adjoint_memory(budget(L))
        J, g, k = sim.adjoint(phi0, target, 4, checkpoint_interval=K, iterations=iters)
        assert J == J0 and k == k0 and np.array_equal(g, g0), ("scheduled", K, L)


def test_fp32_state_matches_fp64_state():
    """The adjoint Poisson state stored in FP32 (arithmetic and reductions in
    FP64): density ratio 1000 with thousands of iterations per step, and the
    walls + gravity centroid case.  Measured differences are ~1e-8."""
    announce("adjoint FP32 state vs FP64 state")
    P = params(lx=10, ly=8, lz=8, rho_L=1000.0)
    phi0, target, _ = case(P)
    res = {}
    for prec in ("fp64", "fp32"):
        sim = hip.InamuroHIP(P)
        sim.setPoisson(check_interval=100)
        sim.set_adjoint_state_precision(prec)
        res[prec] = sim.adjoint(phi0, target, 3, checkpoint_interval=2)
        assert sim.iteration_history()[-1] >= 1000
    (J64, g64, k64), (J32, g32, k32) = res["fp64"], res["fp32"]
    assert J32 == J64
    assert rel(g32, g64) <= 1e-6, rel(g32, g64)
    assert abs(k32 - k64) <= 1e-6 * abs(k64), (k32, k64)


def test_progress_reports_and_setting_validation():
    """lbm_set_progress: one report per forward-pass step and per reversed
    step, in order, and block reports inside the reverse sweep; the objective
    and state-precision setters reject unknown values."""
    announce("adjoint progress reports")
    P = params()
    phi0, target, _ = case(P)
    sim = hip.InamuroHIP(P)
    sim.setPoisson(check_interval=10)
    calls = []
    sim.set_progress(lambda phase, step, it: calls.append((phase, step, it)))
    sim.adjoint(phi0, target, 3, checkpoint_interval=2)
    iters = sim.iteration_history()[-3:]
    fwd = [c for c in calls if c[0] == "forward step"]
    rev = [c for c in calls if c[0] == "reverse step"]
    assert fwd == [("forward step", k, iters[k]) for k in range(3)]
    assert rev == [("reverse step", k, iters[k]) for k in (2, 1, 0)]
    assert any(c[0] == "reverse block" for c in calls)
    sim.set_progress(None)
    n = len(calls)
    sim.adjoint(phi0, target, 1, checkpoint_interval=1)
    assert len(calls) == n
    for bad in (lambda: load().lbm_set_adjoint_objective(sim._h, 7),
                lambda: load().lbm_set_adjoint_state_precision(sim._h, 3)):
        assert bad() == hip.LBM_ERR_ARG


def load():
    return hip.load_library()


def test_context_state_after_adjoint_is_the_forward_end():
    P = params()
    phi0, target, _ = case(P)
    sim = hip.InamuroHIP(P)
    sim.setPoisson(check_interval=10)
    sim.adjoint(phi0, target, 3, checkpoint_interval=1)
    after = sim.downloadFieldsToCPU()
    fwd = hip.InamuroHIP(P)
    fwd.setPoisson(check_interval=10)
    fwd.initFromCPU(fei=phi0, rho=fwd.initial_rho(phi0), u=0.0, v=0.0, w=0.0, p=0.0)
    fwd.step(3)
    ref = fwd.downloadFieldsToCPU()
    for k in ("fei", "rho", "u", "v", "w", "p"):
        assert np.array_equal(after[k], ref[k]), k
    assert sim.iteration_history() == fwd.iteration_history()


def test_gradient_matches_central_fd_16_directions():
    """Density ratio 1000, the reference's check interval; 16 seeded unit
    Gaussian directions over phi0 (h = 1e-6), and k_g."""
    P = params(lx=16, ly=12, lz=12, rho_L=1000.0, DD=8.0)
    announce("adjoint vs central FD, 16 directions")
    phi0, target, rng = case(P, seed=3)
    sim = hip.InamuroHIP(P)
    J, g, gk = sim.adjoint(phi0, target, 3, checkpoint_interval=2)
    iters = sim.iteration_history()[-3:]
    h = 1e-6
    worst = 0.0
    for _ in range(16):
        d = rng.standard_normal(phi0.shape)
        fd = (sim.objective(phi0 + h * d, target, iters) - sim.objective(phi0 - h * d, target, iters)) / (2 * h)
        ad = float((g * d).sum())
        worst = max(worst, abs(fd - ad) / abs(fd))
    assert worst <= FD_TOL, worst
    sp = hip.InamuroHIP(dataclasses.replace(P, k_g=P.k_g + h))
    sm = hip.InamuroHIP(dataclasses.replace(P, k_g=P.k_g - h))
    fd = (sp.objective(phi0, target, iters) - sm.objective(phi0, target, iters)) / (2 * h)
    assert abs(fd - gk) <= FD_TOL * abs(fd), (fd, gk)


def test_adjoint_rejects_mixed_context_and_bad_schedule():
    P = params()
    phi0, target, _ = case(P)
    mx = hip.InamuroHIP(P, precision=hip.MIXED)
    with pytest.raises(Exception, match="FP64"):
        mx.adjoint(phi0, target, 1)
    sim = hip.InamuroHIP(P)
    with pytest.raises(Exception, match="multiples"):
        sim.adjoint(phi0, target, 2, iterations=[100, 150])


def test_failed_adjoint_leaves_the_settings_as_they_were():
    """A forward pass that fails on the host-side numeric check (a NaN in
    phi0 makes rho, tau_h and then the pressure NaN; every kernel exits
    normally) must not leave the adjoint's frozen counts, its timing switch
    or its check observer installed: afterwards lbm_upload + lbm_step
    converge as on a fresh context, with the caller's own schedule in force
    and timing still on."""
    P = params()
    phi0, target, _ = case(P)
    sim = hip.InamuroHIP(P)
    sim.setPoisson(check_interval=10)
    sim.set_timing(True)
    sim.set_iteration_schedule([40])  # the caller's own frozen count for one step
    checks = []
    sim.set_check_observer(checks.append)
    bad = phi0.copy()
    bad[4, 5, 6] = np.nan
    announce("adjoint with a NaN in phi0 (expected to fail on the host check)")
    with pytest.raises(RuntimeError, match="NaN/Inf"):
        sim.adjoint(bad, target, 3, checkpoint_interval=1, iterations=[10, 200, 200])
    assert checks == []  # no records from inside lbm_adjoint
    ref = hip.InamuroHIP(P)
    ref.setPoisson(check_interval=10)
    for s_ in (sim, ref):
        s_.initFromCPU(fei=phi0, rho=s_.initial_rho(phi0), u=0.0, v=0.0, w=0.0, p=0.0)
    ref.set_iteration_schedule([40])
    assert sim.step(3) == ref.step(3)
    assert sim.iteration_history()[0] == 40
    assert sim.perf().collide_ms > 0.0  # timing still on
    assert len(checks) == sum(sim.iteration_history()) // 10
    a, b = sim.downloadFieldsToCPU(), ref.downloadFieldsToCPU()
    for k in ("fei", "rho", "u", "v", "w", "p"):
        assert np.array_equal(a[k], b[k]), k


def test_set_poisson_validates_every_argument_first():
    """A rejected lbm_set_poisson changes no setting: neither an invalid dual
    mode after a valid interval and tolerance, nor a new interval that the
    stored schedule's counts are not multiples of."""
    P = params()
    sim = hip.InamuroHIP(P)
    sim.set_iteration_schedule([200, 300])
    with pytest.raises(ValueError, match="dual_residual"):
        sim._check(load().lbm_set_poisson(sim._h, 50, 1e-4, 0, 7))
    with pytest.raises(ValueError, match="multiples of this check interval"):
        sim.setPoisson(200, 1e-4)
    ref = hip.InamuroHIP(P)
    ref.set_iteration_schedule([200, 300])
    phi0, target, _ = case(P)
    for s_ in (sim, ref):
        s_.initFromCPU(fei=phi0, rho=s_.initial_rho(phi0), u=0.0, v=0.0, w=0.0, p=0.0)
    got = sim.step(3)
    assert got[:2] == [200, 300]
    assert got == ref.step(3)  # the third step converges with the default interval and tolerance


def test_launch_shape_does_not_change_the_gradient():
    """The forward kernel's launch shape (which the recomputation follows)
    and the adjoint Poisson kernel's block size (256 or 768 threads: a
    context's first adjoint call times both on its first reverse iterations,
    later calls use the faster) change no bit of J or the gradients."""
    P = params(lx=16, ly=12, lz=12, rho_L=1000.0, DD=8.0)
    phi0, target, _ = case(P, seed=5)
    out = []
    for cfg in (0, 1, 2):
        announce(f"adjoint with launch shape {cfg}")
        sim = hip.InamuroHIP(P)
        sim.set_poisson_launch(cfg)
        out.append(sim.adjoint(phi0, target, 3, checkpoint_interval=2))
        out.append(sim.adjoint(phi0, target, 3, checkpoint_interval=2))
        assert sim.poisson_launch() == cfg
        sim.close()
    for J, g, gk in out[1:]:
        assert J == out[0][0] and gk == out[0][2]
        assert np.array_equal(g, out[0][1])
