"""The falling-droplet extension of the oracle: no-slip z walls and gravity.

The reference has neither, so the restatement is pinned by what the
mathematics fixes: streaming with halfway bounce-back is a permutation that
reverses a population at the wall; the stencils see a zero normal gradient
about a wall half a cell out; the gravity term adds exactly the buoyant
acceleration to the predicted velocity; with both switched off the step is
the reference's, bit for bit (pinned by the golden fixtures)."""
import dataclasses

import numpy as np
import pytest

from oracle import pyoracle as po

Q = 15
CZ = np.array([0, 0, 0, 1, 0, 0, -1, 1, 1, 1, -1, -1, -1, -1, 1])
OPP = [0, 4, 5, 6, 1, 2, 3, 11, 12, 13, 14, 7, 8, 9, 10]


def params(lx=10, ly=8, lz=12, rho_L=100.0):
    return po.Params(lx=lx, ly=ly, lz=lz, DD=6.0, rho_L=rho_L, rho_G=1.0, mu_L=0.008, mu_G=8e-5,
                     k_f=0.5, k_g=1e-4, T=0.035, a=1.0, b=6.7, fei_max=0.093, fei_min=0.016,
                     fei_L=0.093, fei_G=0.016)


def test_streaming_with_walls_is_a_permutation_that_reverses_at_the_wall():
    P = params()
    o = po.Oracle(P)
    o.set_z_walls(True)
    rng = np.random.default_rng(1)
    f = rng.standard_normal((Q, P.lz, P.ly, P.lx))
    o.set("f", f)
    o.stage("stream_fg")
    g = o.get("f")
    assert np.array_equal(np.sort(f.ravel()), np.sort(g.ravel()))
    # at z = 0 every population arriving from below is the cell's own
    # downward population, reversed; likewise at the top
    for q in range(Q):
        if CZ[q] == 1:
            assert np.array_equal(g[q, 0], f[OPP[q], 0])
        if CZ[q] == -1:
            assert np.array_equal(g[q, -1], f[OPP[q], -1])
    # away from the walls streaming is the periodic shift
    assert np.array_equal(g[3, 5], f[3, 4])


def test_stencils_see_zero_normal_gradient_at_the_walls():
    P = params()
    o = po.Oracle(P)
    o.set_z_walls(True)
    z = np.arange(P.lz, dtype=np.float64)[:, None, None] * np.ones((P.lz, P.ly, P.lx))
    o.set("u", z)
    o.stage("gradients")
    uz = o.get("u_z")
    lap = o.get("u_lap")
    # interior: exact derivative of a linear field, zero Laplacian
    assert np.allclose(uz[1:-1], 1.0, rtol=0, atol=1e-15)
    assert np.allclose(lap[1:-1], 0.0, atol=1e-14)
    # the five stencil points beyond the wall read the wall cell itself:
    # (5*1 - 5*0)/10 at z = 0 and (5*0 - 5*(-1))/10 at the top
    assert np.allclose(uz[0], 0.5, atol=1e-15) and np.allclose(uz[-1], 0.5, atol=1e-15)
    assert np.allclose(lap[0], (5 * 1.0 - 5 * 0.0) / 5.0, atol=1e-14)


def test_gravity_gives_the_buoyant_acceleration_on_the_first_step():
    """The predicted velocity gains sum_q c_qz 3 E_q c_qz a(x - c_q), with
    a = gz (1 - rho_G/rho) of the source cell, over the reference step."""
    P = params(lx=16, ly=16, lz=16, rho_L=1000.0)
    gz = -1e-3
    o = po.Oracle(P)
    o.set_gravity(gz)
    o.place_droplet(8.0, 8.0, 8.0, profile="tanh", width=1.5)
    ref = po.Oracle(P)
    ref.place_droplet(8.0, 8.0, 8.0, profile="tanh", width=1.5)
    rho0 = ref.get("rho")
    for x in (o, ref):
        for st in ("collision", "stream_fg", "bc_fg", "macro"):
            x.stage(st)
    a = gz * (1.0 - P.rho_G / rho0)
    CX = [0, 1, 0, 0, -1, 0, 0, 1, -1, 1, 1, -1, 1, -1, -1]
    CY = [0, 0, 1, 0, 0, -1, 0, 1, 1, -1, 1, -1, -1, 1, -1]
    core = (slice(4, 12),) * 3  # away from the x slip walls
    for comp, c in (("u", CX), ("v", CY), ("w", list(CZ))):
        d = o.get(comp) - ref.get(comp)
        expect = np.zeros_like(d)
        for q in range(Q):
            if CZ[q] == 0 or c[q] == 0:
                continue
            wq = 1 / 9 if q <= 6 else 1 / 72
            expect += c[q] * 3.0 * wq * CZ[q] * np.roll(a, shift=(CZ[q], CY[q], CX[q]), axis=(0, 1, 2))
        assert np.abs(d[core] - expect[core]).max() <= 1e-15, comp
    # in the droplet's core the acceleration is the liquid's buoyant value
    dw = o.get("w") - ref.get("w")
    assert abs(dw[8, 8, 8] / (gz * (1.0 - P.rho_G / rho0[8, 8, 8])) - 1.0) < 1e-3


def test_walls_keep_a_fluid_at_rest_at_rest():
    """Uniform gas between the walls: bounce-back of an equilibrium is the
    equilibrium, so nothing moves and the first solve ends at 100."""
    P = params()
    o = po.Oracle(P)
    o.set_z_walls(True)
    assert o.step(2) == [100, 100]
    for k in ("u", "v", "w", "p"):
        assert np.abs(o.get(k)).max() < 1e-15, k


def test_mass_is_conserved_with_walls_and_gravity():
    """The z walls conserve every population (the permutation test above);
    what moves sum(phi) here is the reference's own x-slip copy acting on the
    tanh tail at the x walls, at the 1e-8 level."""
    P = params(lx=24, ly=12, lz=24)
    o = po.Oracle(P)
    o.set_z_walls(True)
    o.set_gravity(-1e-4)
    o.set_poisson(check_interval=10)
    o.place_droplet(12.0, 6.0, 14.0, profile="tanh", width=1.5)
    m0 = o.get("fei").sum()
    o.step(3)
    assert abs(o.get("fei").sum() - m0) < 1e-7 * m0


def _falling_case(objective, steps=3, seed=0):
    """A droplet low between the walls (its tanh tail reaches z = 0), with
    gravity, x slip walls and a small density ratio so a few steps move it."""
    P = params(lx=10, ly=8, lz=12, rho_L=10.0)
    o = po.Oracle(P)
    o.set_z_walls(True)
    o.set_gravity(-2e-3)
    o.set_poisson(check_interval=10)
    o.set_objective(objective)
    o.place_droplet(4.0, 4.0, 3.5, profile="tanh", width=1.5)
    phi0 = o.get("fei")
    free = po.Oracle(P)
    free.set_z_walls(True)
    free.set_gravity(-2e-3)
    free.set_poisson(check_interval=10)
    free.set("fei", phi0)
    free.set("rho", P.rho_G + (phi0 - P.fei_min) * (P.rho_L - P.rho_G) / (P.fei_max - P.fei_min))
    iters = free.step(steps)
    rng = np.random.default_rng(seed)
    target = phi0 + 0.002 * rng.standard_normal(phi0.shape)
    return P, o, phi0, target, iters, rng


@pytest.mark.parametrize("objective", ["misfit", "centroid_z"])
def test_adjoint_with_walls_and_gravity_matches_central_fd(objective):
    """The adjoint's bounce-back, wall-stencil, wall-correct and gravity
    transposes against central differences of the oracle's forward map, for
    the misfit and for the droplet's centroid height."""
    P, o, phi0, target, iters, rng = _falling_case(objective)
    J, g, gk = o.adjoint(phi0, target, iters)
    if objective == "centroid_z":
        w = phi0 - P.fei_G
        z = np.arange(P.lz)[:, None, None]
        assert 2.0 < J < 6.0 and abs(J - (z * w).sum() / w.sum()) < 1.0
    h = 1e-7
    for _ in range(6):
        d = rng.standard_normal(phi0.shape)
        fd = (o.objective(phi0 + h * d, target, iters) - o.objective(phi0 - h * d, target, iters)) / (2 * h)
        ad = float((g * d).sum())
        assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)
    hk = 1e-6
    res = []
    for kg in (P.k_g + hk, P.k_g - hk):
        x = po.Oracle(dataclasses.replace(P, k_g=kg))
        x.set_z_walls(True)
        x.set_gravity(-2e-3)
        x.set_poisson(check_interval=10)
        x.set_objective(objective)
        res.append(x.objective(phi0, target, iters))
    fd = (res[0] - res[1]) / (2 * hk)
    assert abs(fd - gk) <= 1e-5 * abs(fd), (fd, gk)


def test_centroid_objective_of_a_resting_layered_state():
    """The centroid seed: J of a field whose excess over fei_G is a single
    layer is that layer's index, and its gradient is (z - z_c)/W."""
    P = params()
    o = po.Oracle(P)
    o.set_objective("centroid_z")
    o.set_z_walls(True)
    phi0 = np.full((P.lz, P.ly, P.lx), P.fei_G)
    phi0[5] += 0.01
    o.set_poisson(check_interval=10)
    # zero steps: J is the centroid of phi0 itself
    J, g, gk = o.adjoint(phi0, None, [])
    assert abs(J - 5.0) < 1e-12
    W = 0.01 * P.ly * P.lx
    z = np.arange(P.lz, dtype=np.float64)[:, None, None]
    assert np.allclose(g, np.broadcast_to((z - 5.0) / W, g.shape), rtol=0, atol=1e-12)
    assert gk == 0.0
