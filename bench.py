#!/usr/bin/env python3
"""Throughput of the Inamuro two-phase D3Q15 time step on one MI355X.

Workload (default): the 256^3 static droplet, R = 64, density ratio 1000,
tanh profile of width 4 (interface_width 2), phi_L = 0.093, phi_G = 0.016,
a = 1, b = 6.7, T = 0.035, kappa_f = 0.5, kappa_g = 1e-4, tau_f = tau_g = 1,
mu_L = 0.008, mu_G = 8e-5, x-slip walls as in the reference, Poisson
check interval 100 and tolerance 1e-3 (the reference contract).  A step is
one full performTimeStep: collision + streaming + macro + the Poisson solve
iterated to convergence + velocity correction.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64]

Prints one JSON line.  `value` is MLUPS = cells * steps / wall seconds with
every input resident in HBM.  `roofline` is the dominant kernel (the Poisson
iteration): algorithmic bytes per launch (cells x bytes-per-cell of the
Poisson iteration byte model) over its average launch time measured with
HIP events on the library's stream.  `cpu_baseline` times the CPU oracle
(single thread) on a bounded sample of the same 256^3 state.
`hbm_copy_ceiling_gbs` is what plain copies reach on the same GPU.  The
`adjoint` record is the discrete adjoint on the 96^3 ellipsoid over 8 steps.

Other legs: --workload head_on_collision | falling_droplet | static_droplet_64
(...); --adjoint-only [--adjoint-steps T --adjoint-checkpoint K --adjoint-fd N];
--adjoint-falling T (the falling droplet's centroid-height adjoint at full
size); --traffic-probe (one step per precision, for counter runs); --accel
(the reference's Poisson accelerators: iterations to tolerance and time per
iteration); --mixed-drift N (mixed against FP64 in lockstep over N steps of
--workload).
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS & % HBM roofline at 256³, FP64 vs mixed precision; adjoint/forward cost"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per cell of one Poisson iteration (h read+write, p
# read+write, rho and source read): FP64 15*8*2 + 8*2 + 8*2, mixed
# 15*4*2 + 8*2 + 4*2.
POISSON_BYTES = {"fp64": 272, "mixed": 144}
# Bytes per cell per step outside the Poisson iterations, and per check.
STEP_BYTES = {"fp64": 632, "mixed": 320}
CHECK_BYTES = 24

PHYSICS = dict(rho_L=1000.0, rho_G=1.0, mu_L=0.008, mu_G=8e-5, tauf=1.0, taug=1.0, k_f=0.5,
               k_g=1e-4, T=0.035, a=1.0, b=6.7, fei_max=0.093, fei_min=0.016, fei_L=0.093,
               fei_G=0.016)

# Replay Poisson segments as captured HIP graphs (results identical either
# way); --no-graphs launches them one by one, for profilers that cannot trace
# graph capture.
GRAPHS = True

WORKLOADS = {
    "static_droplet_256": dict(shape=(256, 256, 256), R=64),
    "static_droplet_64": dict(shape=(64, 64, 64), R=16),
    "static_droplet_128": dict(shape=(128, 128, 128), R=32),
    # two R=24 droplets at x=80 and x=176 moving at -/+0.02 towards each other,
    # radii perturbed by seeded real spherical harmonics 2<=l<=8 (RMS 0.5)
    "head_on_collision": dict(shape=(128, 128, 256), R=24, head_on=True),
    # R=32 droplet released at z=300 in a 192x192x384 box between no-slip z
    # walls, falling under gravity 1e-5 (lattice units)
    "falling_droplet": dict(shape=(384, 192, 192), R=32, cz=300.0, zwalls=True, gz=-1e-5),
}


def workload_params(hip, wl):
    lz, ly, lx = wl["shape"]
    return hip.SolverParams(lx=lx, ly=ly, lz=lz, DD=2.0 * wl["R"], **PHYSICS)


def initial_fields(wl, P):
    """Host fields of the workload (None: the library's place_droplet)."""
    if not wl.get("head_on"):
        return None
    from lbm_amd import configs

    return configs.head_on_collision(P, shape=wl["shape"], radius=wl["R"])


def make_sim(hip, wl, precision):
    P = workload_params(hip, wl)
    sim = hip.InamuroHIP(P, precision={"fp64": hip.FP64, "mixed": hip.MIXED}[precision])
    sim.setUsePoissonGraph(GRAPHS)
    if wl.get("zwalls"):
        sim.set_z_walls(True)
    if wl.get("gz"):
        sim.set_gravity(wl["gz"])
    fields = initial_fields(wl, P)
    if fields is None:
        lz, ly, lx = wl["shape"]
        sim.reset_fields()
        sim.place_droplet(lx / 2.0, ly / 2.0, wl.get("cz", lz / 2.0), profile="tanh", width=2.0)
    else:
        sim.upload_fields(fields)
    return sim, P


def run_precision(hip, wl, precision, steps, warmup):
    sim, P = make_sim(hip, wl, precision)
    sim.set_timing(True)
    t_start = time.perf_counter()
    for k in range(warmup):
        sim.performTimeStepGPU(1)
        # progress on stderr as each step completes (the getter waits for it)
        print(f"{precision} warm-up step {k + 1}/{warmup}: {sim.getLastPoissonDiagnostics().iterations} "
              f"Poisson iterations, {time.perf_counter() - t_start:.1f} s", file=sys.stderr, flush=True)
    sim.synchronize()
    sim.reset_perf()
    iters = []
    t0 = time.perf_counter()
    for k in range(steps):
        sim.performTimeStepGPU(1)
        iters.append(int(sim.getLastPoissonDiagnostics().iterations))
        print(f"{precision} step {k + 1}/{steps}: {iters[-1]} Poisson iterations, "
              f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    sim.synchronize()
    t1 = time.perf_counter()
    pf = sim.perf()
    cells = P.lx * P.ly * P.lz
    wall = t1 - t0
    avg_iter_ms = pf.poisson_ms / max(pf.poisson_iter_launches, 1)
    bpi = POISSON_BYTES[precision] * cells
    achieved = bpi / (avg_iter_ms * 1e-3) / 1e9
    n_it = sum(iters)
    n_chk = n_it // 100
    step_bytes = cells * (STEP_BYTES[precision] * steps + POISSON_BYTES[precision] * n_it + CHECK_BYTES * n_chk)
    out = dict(
        precision=precision, wall_s=wall, mlups=cells * steps / wall / 1e6,
        ms_per_step=wall / steps * 1e3, iterations=iters,
        poisson_iter_ms=avg_iter_ms, poisson_iter_launches=int(pf.poisson_iter_launches),
        mlups_h=cells * n_it / (pf.poisson_ms * 1e-3) / 1e6 if pf.poisson_ms > 0 else None,
        achieved_gbs=achieved, bytes_per_launch=bpi,
        whole_step_gbs=step_bytes / wall / 1e9,
        phase_ms=dict(collide=pf.collide_ms, macro=pf.macro_ms, poisson=pf.poisson_ms,
                      check=pf.check_ms, correct=pf.correct_ms),
    )
    out["fields"] = sim.downloadFieldsToCPU()
    out["mass"] = sim.mass()  # device FP64 sum of phi
    out["poisson_launch"] = sim.poisson_launch()  # launch shape the context chose
    sim.close()
    return out


def field_errors(mixed, fp64):
    """Per-field error of the mixed-precision run against the FP64 run of the
    same steps: L-inf and L2 relative to the FP64 field's max / L2 norm; the
    spurious current is max |u| of a static droplet (ideally zero)."""
    import numpy as np

    err = {}
    for k in ("fei", "rho", "u", "v", "w", "p"):
        a, b = mixed[k], fp64[k]
        d = a - b
        err[k] = dict(linf_abs=float(np.abs(d).max()),
                      linf_rel=float(np.abs(d).max() / max(np.abs(b).max(), 1e-300)),
                      l2_rel=float(np.sqrt((d * d).sum()) / max(np.sqrt((b * b).sum()), 1e-300)))
    speed = {name: float(np.sqrt(f["u"] ** 2 + f["v"] ** 2 + f["w"] ** 2).max())
             for name, f in (("fp64", fp64), ("mixed", mixed))}
    return err, speed


# Adjoint workload: 96^3 oscillating ellipsoid, semi-axes 20, 16, 16, the
# static droplet's physics, J = sum (phi_T - phi_target)^2 with phi_target a
# sphere of R = 17.2 (same tanh profile, width 4).
ADJ = dict(L=96, axes=(20.0, 16.0, 16.0), R_target=17.2)


def _progress(label, t0):
    """A progress callback for lbm_adjoint that prints at most every 15 s."""
    last = [time.perf_counter()]

    def report(phase, step, it):
        now = time.perf_counter()
        if now - last[0] > 15.0:
            last[0] = now
            print(f"{label}: {phase} {step} at iteration {it}, {now - t0:.1f} s", file=sys.stderr, flush=True)

    return report


def run_adjoint(hip, steps, checkpoint, fd_directions=0, h=1e-6):
    """Plain forward of `steps` steps, then lbm_adjoint of the same map with a
    checkpoint every `checkpoint` steps; cost ratio = reverse sweep (replays
    and recomputation included) over the plain forward.  With fd_directions,
    the gradient against central differences of the same frozen-count map:
    that many seeded unit-Gaussian directions over phi0 (step h) and the k_g
    direction (step h)."""
    import numpy as np
    from lbm_amd import configs

    L = ADJ["L"]
    P = hip.SolverParams(lx=L, ly=L, lz=L, DD=2.0 * ADJ["R_target"], **PHYSICS)
    phi0 = configs.ellipsoid_phi(L, ADJ["axes"], P)
    target = configs.ellipsoid_phi(L, (ADJ["R_target"],) * 3, P)
    sim = hip.InamuroHIP(P)
    sim.setUsePoissonGraph(GRAPHS)
    sim.initFromCPU(fei=phi0, rho=sim.initial_rho(phi0), u=0.0, v=0.0, w=0.0, p=0.0)
    sim.synchronize()
    t0 = time.perf_counter()
    for k0 in range(0, steps, 100):
        sim.performTimeStepGPU(min(100, steps - k0))
        sim.synchronize()
        if steps > 100:
            print(f"adjoint leg: plain forward {min(steps, k0 + 100)} steps, {time.perf_counter() - t0:.1f} s",
                  file=sys.stderr, flush=True)
    fwd_ms = (time.perf_counter() - t0) * 1e3
    iters = sim.iteration_history()
    print(f"adjoint leg: plain forward {fwd_ms:.0f} ms, iterations {iters[:10]}{'...' if steps > 10 else ''}",
          file=sys.stderr, flush=True)
    sim.set_progress(_progress("adjoint leg", time.perf_counter()))
    J, g, gk = sim.adjoint(phi0, target, steps, checkpoint_interval=checkpoint)
    st = sim.adjoint_stats()
    rec = dict(
        workload=f"ellipsoid_96: 96^3, semi-axes 20/16/16 -> target sphere R=17.2, FP64, {steps} steps from rest, "
                 f"checkpoint every {checkpoint}",
        steps=steps, checkpoint_interval=checkpoint,
        poisson_iterations_per_step=iters if steps <= 20 else None,
        poisson_iterations_total=int(sum(iters)),
        forward_ms=fwd_ms, adjoint_forward_ms=st.forward_ms, reverse_ms=st.reverse_ms,
        adjoint_over_forward=st.reverse_ms / fwd_ms,
        total_over_forward=(st.forward_ms + st.reverse_ms) / fwd_ms,
        recomputed_iterations=int(st.recomputed_iterations), adjoint_iterations=int(st.adjoint_iterations),
        store_slots=int(st.store_slots), J=J, dJ_dkg=gk, grad_phi0_linf=float(abs(g).max()),
        reverse_solve_final_relative_change=st.final_relative_change,
        reverse_solve_iterations_to_tolerance=int(st.iterations_to_tolerance),
    )
    if fd_directions:
        its = sim.iteration_history()[-steps:]
        rng = np.random.default_rng(3)
        rows = []
        for i in range(fd_directions):
            d = rng.standard_normal(phi0.shape)
            jp = sim.objective(phi0 + h * d, target, its)
            print(f"adjoint fd direction {i}: J(+h) {jp:.12e}", file=sys.stderr, flush=True)
            jm = sim.objective(phi0 - h * d, target, its)
            fd = (jp - jm) / (2 * h)
            ad = float((g * d).sum())
            rows.append(dict(fd=fd, adjoint=ad, rel=abs(fd - ad) / abs(fd)))
            print(f"adjoint fd direction {i}: fd {fd:.10e} adjoint {ad:.10e} rel {rows[-1]['rel']:.2e}",
                  file=sys.stderr, flush=True)
        # k_g at +-h and +-2h: the two-point central difference and the
        # five-point one, (8 (J(+h) - J(-h)) - (J(+2h) - J(-2h))) / (12 h),
        # whose truncation error is O(h^4) instead of O(h^2)
        jk = {}
        for m in (1, -1, 2, -2):
            kg = P.k_g + m * h
            x = hip.InamuroHIP(dataclasses.replace(P, k_g=kg))
            x.setUsePoissonGraph(GRAPHS)
            jk[m] = x.objective(phi0, target, its)
            print(f"adjoint fd k_g = {kg:.6e}: J {jk[m]:.12e}", file=sys.stderr, flush=True)
            x.close()
        fdk = (jk[1] - jk[-1]) / (2 * h)
        fdk4 = (8.0 * (jk[1] - jk[-1]) - (jk[2] - jk[-2])) / (12.0 * h)
        rec["finite_differences"] = dict(h=h, phi0_directions=rows, max_rel_phi0=max(r["rel"] for r in rows),
                                         kg=dict(fd=fdk, adjoint=gk, rel=abs(fdk - gk) / abs(fdk)),
                                         kg_five_point=dict(fd=fdk4, adjoint=gk, rel=abs(fdk4 - gk) / abs(fdk4)))
    sim.close()
    return rec


def run_adjoint_falling(hip, steps, checkpoint, state, fd=False, h=1e-6, iterations=None):
    """The falling droplet's adjoint: dJ/dk_g (and dJ/dphi0) of its centroid
    height after `steps` steps from the release, on the full grid with the
    walls and gravity, the adjoint Poisson state in `state` precision.  Cost
    ratio as in run_adjoint; with fd, central differences of the same
    frozen-count map in k_g.  `iterations` freezes given Poisson counts (a
    short probe of the adjoint kernels at full size for counter runs)."""
    wl = WORKLOADS["falling_droplet"]
    lz, ly, lx = wl["shape"]

    def sim_for(P):
        sim = hip.InamuroHIP(P)
        sim.setUsePoissonGraph(GRAPHS)
        sim.set_z_walls(True)
        sim.set_gravity(wl["gz"])
        sim.set_adjoint_objective("centroid_z")
        return sim

    P = workload_params(hip, wl)
    sim = sim_for(P)
    sim.reset_fields()
    sim.place_droplet(lx / 2.0, ly / 2.0, wl["cz"], profile="tanh", width=2.0)
    phi0 = sim.downloadFieldsToCPU()["fei"]
    # the adjoint's start: rho linear in phi0, at rest
    sim.initFromCPU(fei=phi0, rho=sim.initial_rho(phi0), u=0.0, v=0.0, w=0.0, p=0.0)
    sim.synchronize()
    if iterations is not None:
        sim.set_iteration_schedule(iterations)
    t0 = time.perf_counter()
    for k in range(steps):
        sim.performTimeStepGPU(1)
        sim.synchronize()
        print(f"falling adjoint leg: plain forward step {k + 1}: "
              f"{sim.getLastPoissonDiagnostics().iterations} iterations, "
              f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    fwd_ms = (time.perf_counter() - t0) * 1e3
    iters = sim.iteration_history()[-steps:]
    sim.set_adjoint_state_precision(state)
    last = [time.perf_counter()]

    def progress(phase, step, it):
        now = time.perf_counter()
        if phase != "reverse block" or now - last[0] > 20.0:
            last[0] = now
            print(f"falling adjoint leg: {phase} {step} at iteration {it}, {now - t0:.1f} s",
                  file=sys.stderr, flush=True)

    sim.set_progress(progress)
    J, g, gk = sim.adjoint(phi0, None, steps, checkpoint_interval=checkpoint, iterations=iters)
    st = sim.adjoint_stats()
    sim.close()
    rec = dict(
        workload=(f"falling_droplet: {lx}x{ly}x{lz}, tanh droplet R={wl['R']} released at z={wl['cz']:g}, "
                  f"no-slip z walls, gravity {wl['gz']:g}, density ratio 1000; J = centroid height after "
                  f"{steps} steps; FP64 forward and recomputation, {state} adjoint Poisson state"),
        steps=steps, checkpoint_interval=checkpoint, poisson_iterations_per_step=iters,
        forward_ms=fwd_ms, adjoint_forward_ms=st.forward_ms, reverse_ms=st.reverse_ms,
        adjoint_over_forward=st.reverse_ms / fwd_ms,
        total_over_forward=(st.forward_ms + st.reverse_ms) / fwd_ms,
        recomputed_iterations=int(st.recomputed_iterations), adjoint_iterations=int(st.adjoint_iterations),
        store_slots=int(st.store_slots), J=J, dJ_dkg=gk, grad_phi0_linf=float(abs(g).max()),
        reverse_solve_final_relative_change=st.final_relative_change,
        reverse_solve_iterations_to_tolerance=int(st.iterations_to_tolerance),
    )
    if fd:
        vals = []
        for kg in (P.k_g + h, P.k_g - h):
            x = sim_for(dataclasses.replace(P, k_g=kg))
            vals.append(x.objective(phi0, None, iters))
            print(f"falling adjoint leg: k_g = {kg:.6e}: J = {vals[-1]:.12e}", file=sys.stderr, flush=True)
            x.close()
        fdv = (vals[0] - vals[1]) / (2 * h)
        rec["finite_difference_kg"] = dict(h=h, fd=fdv, adjoint=gk, rel_error=abs(fdv - gk) / abs(fdv))
    return rec


# The reference's optional Poisson accelerators (FP64), each as lbm_gpu would
# set it, and the algorithmic bytes per cell of one of its iterations: the
# pull/onepass map moves what the strict iteration does (h read and written,
# p read and written, tau_h and the source read); the scalar Jacobi sweep
# moves p, rho and the source; an Anderson(1) iteration also reads the
# previous residual in the map (384) and reads three / writes two
# distributions plus p to mix (608).
ACCEL_VARIANTS = [
    ("strict split", {}, 272),
    ("fused + fused boundary", dict(variant="fused", fused_boundary=1), 272),
    ("split, relax scale 0.85", dict(pressure_relax_scale=0.85), 272),
    ("onepass", dict(variant="onepass"), 272),
    ("onepass + Anderson(1)", dict(variant="onepass", anderson_m1=1), 992),
    ("scalar Jacobi", dict(variant="scalar"), 32),
    ("split + two-grid 0.5", dict(two_grid=1), 272),
    ("split + source-aware start", dict(source_aware_hh_init=1), 272),
    ("onepass + Anderson + two-grid + source-aware",
     dict(variant="onepass", anderson_m1=1, two_grid=1, source_aware_hh_init=1), 992),
]


def run_mixed_drift(hip, workload, steps):
    """Mixed precision against FP64 in lockstep on one workload, beside a
    control (FP64 started from the FP32-rounded start state): the Poisson
    iteration counts of every step and, every 10 steps (and after steps 1-3),
    the relative L-inf error of every field of both against FP64 and the
    spurious-current magnitude max|u| of the FP64 run."""
    import numpy as np

    wl = WORKLOADS[workload]
    f64, P = make_sim(hip, wl, "fp64")
    mx, _ = make_sim(hip, wl, "mixed")
    # control: FP64 from the start state rounded to FP32, i.e. how far the
    # FP64 trajectory itself carries differences of FP32 size
    start = f64.downloadFieldsToCPU()
    ctl = hip.InamuroHIP(P, precision=hip.FP64)
    ctl.setUsePoissonGraph(GRAPHS)
    if wl.get("zwalls"):
        ctl.set_z_walls(True)
    if wl.get("gz"):
        ctl.set_gravity(wl["gz"])
    ctl.upload_fields({k: v.astype(np.float32).astype(np.float64) for k, v in start.items()})
    del start
    fields = ("fei", "rho", "u", "v", "w", "p")
    rows, counts_equal = [], True
    worst = {"mixed": dict.fromkeys(fields, 0.0), "fp64_from_fp32_start": dict.fromkeys(fields, 0.0)}
    t0 = time.perf_counter()

    def rel(B, A):
        return {k: float(np.abs(B[k] - A[k]).max()) / max(float(np.abs(A[k]).max()), 1e-300) for k in fields}

    for s in range(1, steps + 1):
        a, b, c = f64.step(1)[0], mx.step(1)[0], ctl.step(1)[0]
        counts_equal &= a == b
        msg = f"mixed drift: step {s}, iterations fp64 {a} mixed {b} control {c}"
        if s % 10 == 0 or s <= 3 or s == steps:
            A = f64.downloadFieldsToCPU()
            err = {"mixed": rel(mx.downloadFieldsToCPU(), A), "fp64_from_fp32_start": rel(ctl.downloadFieldsToCPU(), A)}
            for run, e in err.items():
                for k in fields:
                    worst[run][k] = max(worst[run][k], e[k])
            rows.append(dict(step=s, iterations=a, control_iterations=c, linf_rel=err,
                             max_speed_fp64=float(np.abs(A["u"]).max())))
            msg += " | mixed " + " ".join(f"{k} {v:.2e}" for k, v in err["mixed"].items())
            msg += " | control u " + f"{max(err['fp64_from_fp32_start'][k] for k in ('u', 'v', 'w')):.2e}"
        print(f"{msg} [{time.perf_counter() - t0:.0f} s]", file=sys.stderr, flush=True)
    for sim in (f64, mx, ctl):
        sim.close()
    return {"mixed_drift": dict(workload=workload, steps=steps, iteration_counts_equal=bool(counts_equal),
                                worst_linf_rel=worst, samples=rows)}


def run_accel(hip, steps=3, bw_iterations=1000):
    """Convergence and bandwidth of each accelerator.  Convergence: the 64^3
    static droplet (FP64), `steps` steps from the cold start, iterations per
    step, Poisson time, and the fields' distance from the strict run's.
    Bandwidth: the 256^3 droplet with the iteration counts frozen at
    [100, bw_iterations], time per iteration of the second solve."""
    out = []
    strict = None
    for name, accel, bpc in ACCEL_VARIANTS:
        rec = {"variant": name, "settings": accel, "bytes_per_cell": bpc}
        sim, P = make_sim(hip, WORKLOADS["static_droplet_64"], "fp64")
        sim.set_poisson_accel(**accel)
        sim.set_timing(True)
        t0 = time.perf_counter()
        try:
            its = sim.step(steps)
        except RuntimeError as e:
            rec["error"] = str(e)
            sim.close()
            out.append(rec)
            continue
        sim.synchronize()
        pf = sim.perf()
        f = sim.downloadFieldsToCPU()
        sim.close()
        rec.update(static_droplet_64_iterations=its, static_droplet_64_poisson_ms=pf.poisson_ms,
                   static_droplet_64_wall_ms=(time.perf_counter() - t0) * 1e3)
        if strict is None:
            strict = f
        else:
            err, _ = field_errors(f, strict)
            rec["relative_linf_vs_strict"] = {k: err[k]["linf_rel"] for k in ("p", "u", "fei")}
        print(f"accel {name}: {its} iterations, Poisson {pf.poisson_ms:.1f} ms", file=sys.stderr, flush=True)

        sim, P = make_sim(hip, WORKLOADS["static_droplet_256"], "fp64")
        sim.set_poisson_accel(**accel)
        sim.set_iteration_schedule([100, bw_iterations])
        try:
            sim.performTimeStepGPU(1)
            sim.set_timing(True)
            sim.reset_perf()
            sim.performTimeStepGPU(1)
            sim.synchronize()
        except RuntimeError as e:
            rec["error_256"] = str(e)
            sim.close()
            out.append(rec)
            continue
        pf = sim.perf()
        sim.close()
        cells = P.lx * P.ly * P.lz
        it_ms = pf.poisson_ms / max(pf.poisson_iter_launches, 1)
        gbs = bpc * cells / (it_ms * 1e-3) / 1e9
        rec.update(static_droplet_256_ms_per_iteration=it_ms, achieved_gbs=gbs, frac=gbs / HBM_PEAK_GBS)
        print(f"accel {name}: 256^3 {it_ms:.3f} ms/iteration, {gbs:.0f} GB/s", file=sys.stderr, flush=True)
        out.append(rec)
    return {"accelerators": out}


def copy_ceiling():
    """HBM bandwidths (GB/s) that copies reach on this GPU, by access
    pattern, from lbm_amd/copy_probe's ceiling mode run as a child process;
    None if the probe is not built or fails."""
    import subprocess

    probe = os.path.join(ROOT, "lbm_amd", "copy_probe")
    if not os.path.exists(probe):
        return None
    try:
        out = subprocess.run([probe, "ceiling"], capture_output=True, text=True, timeout=120).stdout
    except (OSError, subprocess.SubprocessError):
        return None
    rates = {}
    for line in out.splitlines():
        name, _, rest = line.partition(":")
        if rest.strip().endswith("GB/s"):
            rates[name.strip()] = float(rest.split(",")[1].split()[0])
    return rates or None


def cpu_baseline(hip, wl, workload, n_it_per_step):
    """Single-thread CPU oracle on the same initial state: one non-Poisson step
    part plus two Poisson iterations, extrapolated with the GPU's measured
    iterations per step."""
    from oracle import pyoracle as po

    lz, ly, lx = wl["shape"]
    P = po.Params(lx=lx, ly=ly, lz=lz, DD=2.0 * wl["R"], **PHYSICS)
    o = po.Oracle(P, threads=1)
    if wl.get("zwalls"):
        o.set_z_walls(True)
    if wl.get("gz"):
        o.set_gravity(wl["gz"])
    fields = initial_fields(wl, P)
    if fields is None:
        o.place_droplet(lx / 2.0, ly / 2.0, wl.get("cz", lz / 2.0), profile="tanh", width=2.0)
    else:
        for k, v in fields.items():
            o.set(k, v)
    cells = lx * ly * lz
    print(f"cpu baseline: oracle on one core, {lx}x{ly}x{lz} state ready", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for st in ("collision", "stream_fg", "bc_fg", "macro"):
        o.stage(st)
    t1 = time.perf_counter()
    print(f"cpu baseline: step without the solve {t1 - t0:.1f} s", file=sys.stderr, flush=True)
    k = 2
    o.poisson_iterations(k)
    t2 = time.perf_counter()
    o.stage("correct")
    o.stage("update_hh")
    t3 = time.perf_counter()
    t_np = (t1 - t0) + (t3 - t2)
    t_it = (t2 - t1) / k
    t_step = t_np + n_it_per_step * t_it
    o.close()
    return dict(
        value=cells / t_step / 1e6, unit="MLUPS", cores=1, kind="port",
        sample=(f"{workload} initial state ({lx}x{ly}x{lz}): one collision+stream+boundary+macro+correction pass "
                f"({t_np:.2f} s) and {k} Poisson iterations ({t_it:.3f} s each) timed on one host core, "
                f"extrapolated to {n_it_per_step:.0f} iterations per step (the GPU's measured mean)"),
        host_cpu=host_cpu_model(), nproc=os.cpu_count(),
        # measured where the reference compiles (one core, same inputs): the
        # oracle takes 0.58-0.60x the compiled reference CPU solver's time
        reference_estimate=0.59 * cells / t_step / 1e6,
    )


def host_cpu_model() -> str:
    """The host CPU's model name from /proc/cpuinfo (the machine type where
    that file is absent)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default="both")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="static_droplet_256")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the adjoint leg and print its record")
    ap.add_argument("--adjoint-steps", type=int, default=8)
    ap.add_argument("--adjoint-fd", type=int, default=0, metavar="N",
                    help="with --adjoint-only: also check N directions against central differences")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch Poisson iterations one by one instead of replaying graphs")
    ap.add_argument("--adjoint-checkpoint", type=int, default=4)
    ap.add_argument("--adjoint-falling", type=int, default=0, metavar="T",
                    help="only the falling droplet's adjoint (centroid height, k_g) over T steps")
    ap.add_argument("--adjoint-falling-state", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--adjoint-falling-fd", action="store_true",
                    help="with --adjoint-falling: central differences in k_g")
    ap.add_argument("--accel", action="store_true",
                    help="the reference's Poisson accelerators (FP64): iterations to tolerance on the 64^3 "
                         "droplet and time per iteration at 256^3; prints one JSON line and exits")
    ap.add_argument("--mixed-drift", type=int, default=0, metavar="N",
                    help="mixed precision against FP64 in lockstep over N steps of --workload (iteration counts "
                         "and field errors every 10 steps); prints one JSON line and exits")
    ap.add_argument("--accel-probe", action="store_true",
                    help="the pull-form accelerators (onepass, onepass + Anderson(1), scalar) at 256^3 with the "
                         "Poisson counts frozen at [100, 300] (for rocprofv3 runs) and exit")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run one short solve per precision (for rocprofv3 --pmc runs) and exit")
    ap.add_argument("--adjoint-probe", action="store_true",
                    help="the falling droplet's adjoint at full size over 2 steps of 100 and 300 frozen "
                         "Poisson iterations (FP32 state; for rocprofv3 runs of the adjoint kernels) and exit")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the time step is not sharded across GPUs in this "
              f"build (asked for --gpus {args.gpus})", file=sys.stderr)
        return 2

    from lbm_amd import hip

    global GRAPHS
    GRAPHS = not args.no_graphs

    if args.accel:
        print(json.dumps(run_accel(hip)), flush=True)
        return 0
    if args.mixed_drift:
        print(json.dumps(run_mixed_drift(hip, args.workload, args.mixed_drift)), flush=True)
        return 0
    if args.accel_probe:
        for name, accel, _ in ACCEL_VARIANTS:
            if accel.get("variant") not in ("onepass", "scalar") or accel.get("two_grid"):
                continue
            sim, _ = make_sim(hip, WORKLOADS["static_droplet_256"], "fp64")
            sim.set_poisson_accel(**accel)
            sim.set_iteration_schedule([100, 300])
            sim.performTimeStepGPU(2)
            sim.synchronize()
            sim.close()
            print(f"accel probe {name}: 2 steps, 400 iterations", flush=True)
        return 0
    if args.adjoint_probe:
        rec = run_adjoint_falling(hip, 2, 2, "fp32", iterations=[100, 300])
        print(json.dumps(rec), flush=True)
        return 0
    if args.adjoint_falling:
        rec = run_adjoint_falling(hip, args.adjoint_falling, min(200, args.adjoint_falling),
                                  args.adjoint_falling_state, fd=args.adjoint_falling_fd)
        print(json.dumps(rec), flush=True)
        return 0
    if args.adjoint_only:
        rec = run_adjoint(hip, args.adjoint_steps, args.adjoint_checkpoint, fd_directions=args.adjoint_fd)
        print(json.dumps(rec), flush=True)
        return 0
    wl = WORKLOADS[args.workload]
    lz, ly, lx = wl["shape"]
    precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    if args.traffic_probe:
        # One step from the cold start per precision: the first solve runs 100
        # Poisson iterations (one check), enough dispatches for counter runs.
        for p in precs:
            sim, _ = make_sim(hip, wl, p)
            sim.performTimeStepGPU(1)
            sim.synchronize()
            print(f"traffic probe {p}: {sim.getLastPoissonDiagnostics().iterations} iterations", flush=True)
            sim.close()
        return 0
    results = {p: run_precision(hip, wl, p, args.steps, args.warmup) for p in precs}
    head = results["mixed"] if "mixed" in results else results["fp64"]
    n_it_mean = sum(head["iterations"]) / len(head["iterations"])
    line = {
        "metric": METRIC,
        "value": head["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head["precision"] == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.workload}: {lx}x{ly}x{lz} D3Q15 grid, "
                         + ("two tanh droplets R=24 at x=80/176, velocities +/-0.02, radii perturbed by seeded "
                            "spherical harmonics 2<=l<=8 (RMS 0.5)" if wl.get("head_on") else
                            f"tanh droplet R={wl['R']} released at z={wl['cz']:g} between no-slip z walls, "
                            f"gravity {wl['gz']:g}" if wl.get("zwalls") else
                            f"tanh droplet R={wl['R']} at the centre")
                         + ", density ratio 1000, Poisson tol 1e-3 / check 100, steps after a cold start"),
            "grid": [lx, ly, lz], "radius": wl["R"], "precision": head["precision"],
            "poisson_iterations_per_step": head["iterations"],
            # launch shape of the Poisson kernel each context chose by timing
            # (fp64: 0 = 256 threads non-temporal, 1 = 256, 2 = 768; mixed:
            # 0 = 1024 non-temporal, 1 = 512 non-temporal, 2 = 512, 3 = 1024)
            "poisson_launch": {p: results[p]["poisson_launch"] for p in results},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": head["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": head["achieved_gbs"] / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "k_poisson_push (one Poisson LBM iteration)",
            "algorithmic_bytes_per_launch": head["bytes_per_launch"],
            "avg_launch_ms": head["poisson_iter_ms"],
        },
        "mlups_poisson": head["mlups_h"],
        "whole_step_hbm_frac": head["whole_step_gbs"] / HBM_PEAK_GBS,
        "phase_ms": head["phase_ms"],
    }
    if "mixed" in results and "fp64" in results:
        line["fp64_mlups"] = results["fp64"]["mlups"]
        line["fp64_roofline_frac"] = results["fp64"]["achieved_gbs"] / HBM_PEAK_GBS
        line["mixed_over_fp64_mlups"] = results["mixed"]["mlups"] / results["fp64"]["mlups"]
        line["mixed_over_fp64_iteration_time"] = results["fp64"]["poisson_iter_ms"] / results["mixed"]["poisson_iter_ms"]
        err, speed = field_errors(results["mixed"]["fields"], results["fp64"]["fields"])
        line["mixed_vs_fp64_error"] = err
        line["spurious_current_max_speed"] = speed
        line["mixed_same_iterations"] = results["mixed"]["iterations"] == results["fp64"]["iterations"]
        line["mass_sum_phi"] = {p: results[p]["mass"] for p in ("fp64", "mixed")}
        line["mass_sum_phi"]["mixed_rel_diff"] = abs(results["mixed"]["mass"] - results["fp64"]["mass"]) / abs(
            results["fp64"]["mass"])
    # HBM bytes per launch of the dominant kernel, from the committed
    # rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled by a copy-kernel
    # calibration); null when no measurement exists for this workload.
    for prec in results:
        traffic = os.path.join(ROOT, "profiles", f"traffic_{prec}_{args.workload}.json")
        if os.path.exists(traffic):
            rec = json.load(open(traffic))
            # the record of the launch shape this run's context chose
            shape = rec.get("by_launch", {}).get(str(results[prec]["poisson_launch"]))
            tb = shape["bytes_per_launch"] if shape else rec.get("bytes_per_launch")
            if prec == head["precision"]:
                line["roofline"]["traffic"] = tb
            else:
                line[f"{prec}_traffic_bytes_per_launch"] = tb
    # measured on this GPU next to the spec peak: a plain 4 GiB copy, and
    # copies shaped like the Poisson iteration's 15 + 15 streams
    ceiling = copy_ceiling()
    line["hbm_copy_ceiling_gbs"] = ceiling
    if ceiling:
        best = max(ceiling.values())
        line["roofline"]["frac_of_copy_ceiling"] = head["achieved_gbs"] / best
    if not args.no_adjoint:
        line["adjoint"] = run_adjoint(hip, args.adjoint_steps, args.adjoint_checkpoint)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(hip, wl, args.workload, n_it_mean)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
