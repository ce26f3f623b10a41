"""Discrete adjoint of the oracle's frozen-schedule step.

The reference has no adjoint, so nothing in it pins these functions; they are
pinned by mathematics instead:

* central finite differences of the oracle's own forward map (the same
  primitives, run forward) along seeded random directions in phi0 and along
  k_g -- a dropped term, a wrong sign or a transposed stencil in any primitive
  moves the adjoint off the difference quotient by far more than the tolerance;
* the frozen schedule reproduces the free run when it is given the free run's
  own iteration counts (so the adjoint is the derivative of the real step);
* linearity of the phi0 gradient in the misfit: J is quadratic in phi_T, so
  dJ/dphi0 for target t is affine in t.
"""
import dataclasses

import numpy as np
import pytest

from oracle.pyoracle import Oracle, Params


def _params(lx=10, ly=8, lz=8, rho_L=10.0, k_g=1e-4):
    return Params(lx=lx, ly=ly, lz=lz, rho_L=rho_L, rho_G=1.0, mu_L=0.008, mu_G=8e-5, tauf=1.0, taug=1.0,
                  k_f=0.5, k_g=k_g, T=0.035, a=1.0, b=6.7, fei_max=0.093, fei_min=0.016, fei_L=0.093,
                  fei_G=0.016, DD=6.0)


def _case(P, check_interval, steps, cx=None, seed=0):
    o = Oracle(P)
    o.set_poisson(check_interval=check_interval)
    o.place_droplet(P.lx / 2 if cx is None else cx, P.ly / 2, P.lz / 2, profile="tanh", width=1.5)
    phi0 = o.get("fei")
    free = Oracle(P)
    free.set_poisson(check_interval=check_interval)
    free.set("fei", phi0)
    free.set("rho", P.rho_G + (phi0 - P.fei_min) * (P.rho_L - P.rho_G) / (P.fei_max - P.fei_min))
    iters = free.step(steps)
    rng = np.random.default_rng(seed)
    target = phi0 + 0.002 * rng.standard_normal(phi0.shape)
    return o, free, phi0, target, iters, rng


def _fd_phi(o, phi0, target, iters, d, h):
    return (o.objective(phi0 + h * d, target, iters) - o.objective(phi0 - h * d, target, iters)) / (2 * h)


def test_frozen_schedule_reproduces_free_run():
    P = _params()
    o, free, phi0, target, iters, _ = _case(P, 10, 3)
    J = o.objective(phi0, target, iters)
    for name in ("fei", "u", "v", "w", "p"):
        assert np.array_equal(o.get(name), free.get(name)), name
    J_free = float(((free.get("fei") - target) ** 2).sum())
    assert abs(J - J_free) <= 1e-13 * J_free


@pytest.mark.parametrize("cx", [5.0, 1.5])
def test_phi0_gradient_matches_central_fd(cx):
    """Droplet in the middle, and touching the x slip wall."""
    P = _params()
    o, _, phi0, target, iters, rng = _case(P, 10, 3, cx=cx)
    J, g, _ = o.adjoint(phi0, target, iters)
    assert J > 0
    for _ in range(8):
        d = rng.standard_normal(phi0.shape)
        fd = _fd_phi(o, phi0, target, iters, d, 1e-7)
        ad = float((g * d).sum())
        assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)


def test_kg_gradient_matches_central_fd():
    P = _params()
    o, _, phi0, target, iters, _ = _case(P, 10, 3)
    _, _, gk = o.adjoint(phi0, target, iters)
    hk = 1e-6
    op = Oracle(dataclasses.replace(P, k_g=P.k_g + hk))
    om = Oracle(dataclasses.replace(P, k_g=P.k_g - hk))
    for x in (op, om):
        x.set_poisson(check_interval=10)
    fd = (op.objective(phi0, target, iters) - om.objective(phi0, target, iters)) / (2 * hk)
    assert abs(fd - gk) <= 1e-6 * abs(fd), (fd, gk)


def test_gradient_affine_in_target():
    P = _params()
    o, _, phi0, t1, iters, rng = _case(P, 10, 2)
    t2 = phi0 + 0.002 * rng.standard_normal(phi0.shape)
    _, g1, k1 = o.adjoint(phi0, t1, iters)
    _, g2, k2 = o.adjoint(phi0, t2, iters)
    _, gm, km = o.adjoint(phi0, 0.5 * (t1 + t2), iters)
    scale = np.abs(g1).max() + np.abs(g2).max()
    assert np.abs(gm - 0.5 * (g1 + g2)).max() <= 1e-12 * scale
    assert abs(km - 0.5 * (k1 + k2)) <= 1e-12 * (abs(k1) + abs(k2))


@pytest.mark.slow
def test_density_ratio_1000_matches_central_fd():
    """The reference's density ratio with its check interval of 100: thousands
    of Poisson iterations per step, each transposed with its gauge."""
    P = _params(rho_L=1000.0)
    o, _, phi0, target, iters, rng = _case(P, 100, 3)
    assert iters[1] >= 1000
    _, g, _ = o.adjoint(phi0, target, iters)
    d = rng.standard_normal(phi0.shape)
    fd = _fd_phi(o, phi0, target, iters, d, 1e-7)
    ad = float((g * d).sum())
    assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)
