"""Initial conditions of the workloads (lbm_amd/configs.py).

The spherical-harmonic perturbation of the head-on collision is new (the
reference has none); it is pinned by closed forms, by orthonormality under
Gauss-Legendre quadrature, by an independent library routine (scipy's
associated Legendre functions), by the committed coefficient fixture, and by
reduction to the oracle's own droplet placement when the perturbation is zero.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy.special import lpmv

from lbm_amd import configs
from oracle import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sphere_quadrature(n=24):
    """Exact for polynomials of degree < 2n in cos(theta) and trigonometric
    degree < 2n in phi."""
    x, wx = np.polynomial.legendre.leggauss(n)
    ph = np.arange(2 * n) * (math.pi / n)
    theta = np.arccos(x)[:, None] * np.ones((1, 2 * n))
    phi = np.ones((n, 1)) * ph[None, :]
    w = wx[:, None] * np.full((1, 2 * n), math.pi / n)
    return theta, phi, w


def test_legendre_table_matches_scipy_without_condon_shortley_phase():
    x = np.linspace(-0.999, 0.999, 101)
    P = configs.legendre_table(8, x)
    for (l, m), v in P.items():
        ref = lpmv(m, l, x) * (-1.0) ** m
        assert np.allclose(v, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max()), (l, m)


def test_low_order_harmonics_closed_forms():
    th = np.array([0.3, 1.1, 2.5])
    ph = np.array([0.2, 4.0, -1.3])
    Y = configs.real_sph_harm_table(1, th, ph)
    c = math.sqrt(3.0 / (4.0 * math.pi))
    assert np.allclose(Y[(0, 0)], 1.0 / math.sqrt(4.0 * math.pi), rtol=1e-15)
    assert np.allclose(Y[(1, 0)], c * np.cos(th), rtol=1e-14)
    assert np.allclose(Y[(1, 1)], c * np.sin(th) * np.cos(ph), rtol=1e-14)
    assert np.allclose(Y[(1, -1)], c * np.sin(th) * np.sin(ph), rtol=1e-14)


def test_real_harmonics_are_orthonormal():
    theta, phi, w = sphere_quadrature()
    Y = configs.real_sph_harm_table(8, theta, phi)
    keys = sorted(Y)
    G = np.array([[float((Y[a] * Y[b] * w).sum()) for b in keys] for a in keys])
    assert np.abs(G - np.eye(len(keys))).max() < 1e-12


def test_perturbation_rms_is_sigma_and_mean_free():
    theta, phi, w = sphere_quadrature()
    for a in configs.sh_coefficients(2):
        dR = configs.radius_perturbation(a, 0.5, theta, phi)
        assert abs(math.sqrt(float((dR * dR * w).sum()) / (4 * math.pi)) - 0.5) < 1e-12
        assert abs(float((dR * w).sum())) < 1e-12  # no l = 0 term: volume kept to first order


def test_coefficients_match_the_committed_fixture():
    rec = json.load(open(os.path.join(GOLDEN, "head_on_sh_coefficients.json")))
    assert (rec["seed"], rec["lmin"], rec["lmax"]) == (configs.SH_SEED, configs.SH_LMIN, configs.SH_LMAX)
    assert [tuple(m) for m in rec["modes"]] == configs.sh_modes()
    for a, hexes in zip(configs.sh_coefficients(2), rec["coefficients"]):
        assert [float.hex(float(x)) for x in a] == hexes


def test_unperturbed_collision_is_the_oracle_droplet_placement():
    P = po.Params(lx=48, ly=24, lz=20, DD=10.0, rho_L=1000.0, fei_max=0.093, fei_min=0.016,
                  fei_L=0.093, fei_G=0.016)
    f = configs.head_on_collision(P, shape=(20, 24, 48), radius=5.0, centers_x=(14.0, 34.0),
                                  speed=0.02, sigma=0.5,
                                  coeffs=[np.zeros(len(configs.sh_modes()))] * 2)
    o = po.Oracle(P)
    o.place_droplet(14.0, 12.0, 10.0, vx=0.02, profile="tanh", width=2.0)
    o.place_droplet(34.0, 12.0, 10.0, vx=-0.02, profile="tanh", width=2.0)
    for k in ("fei", "rho", "u", "v", "w", "p"):
        ref = o.get(k)
        assert np.abs(f[k] - ref).max() <= 4e-16 * max(np.abs(ref).max(), 1.0), k


def test_perturbed_collision_fields():
    P = po.Params(lx=64, ly=32, lz=32, DD=12.0, rho_L=1000.0, fei_max=0.093, fei_min=0.016,
                  fei_L=0.093, fei_G=0.016)
    f = configs.head_on_collision(P, shape=(32, 32, 64), radius=6.0, centers_x=(20.0, 44.0))
    assert f["fei"].max() <= P.fei_max and f["fei"].min() >= P.fei_min
    # the two droplets move towards each other, with speed lf * 0.02
    assert f["u"][16, 16, 20] > 0.019 and f["u"][16, 16, 44] < -0.019
    assert np.all(f["v"] == 0) and np.all(f["w"] == 0) and np.all(f["p"] == 0)
    # rho is the linear blend of the liquid fraction
    lf = (f["fei"] - P.fei_min) / (P.fei_max - P.fei_min)
    assert np.allclose(f["rho"], P.rho_G + lf * (P.rho_L - P.rho_G), rtol=1e-12)
    # the interface (lf = 1/2) lies within the perturbation's range of R = 6
    z, y, x = np.nonzero(np.abs(lf - 0.5) < 0.05)
    r = np.sqrt((x - np.where(x < 32, 20.0, 44.0)) ** 2 + (y - 16.0) ** 2 + (z - 16.0) ** 2)
    assert r.min() > 6.0 - 3.0 and r.max() < 6.0 + 3.0
    assert np.std(r) > 0.1  # not a sphere


CFG3 = po.Params(lx=96, ly=96, lz=96, DD=34.4, rho_L=1000.0, rho_G=1.0, fei_max=0.093, fei_min=0.016,
                 fei_L=0.093, fei_G=0.016)


def test_ellipsoid_with_equal_axes_is_the_sphere_profile():
    """s = r/R, |grad s| = r/R^2, so d = (1 - s) s / |grad s| = R - r: the
    configurations' sphere profile mean + (phi_L - phi_G)/2 tanh((R - r)/2)."""
    L, R = 32, 9.5
    phi = configs.ellipsoid_phi(L, (R, R, R), CFG3)
    c = L / 2.0
    z, y, x = np.meshgrid(*(np.arange(L, dtype=np.float64) - c for _ in range(3)), indexing="ij")
    r = np.sqrt(x * x + y * y + z * z)
    want = 0.5 * (CFG3.fei_L + CFG3.fei_G) + 0.5 * (CFG3.fei_L - CFG3.fei_G) * np.tanh((R - r) / 2.0)
    assert np.max(np.abs(phi - want)) <= 1e-16  # a few ulps of phi ~ 0.09
    # and it agrees with the oracle's tanh droplet (place_droplet, interface
    # width 2, phi_G / phi_L as fei_min / fei_max) to rounding
    P = po.Params(lx=L, ly=L, lz=L, DD=2 * R, fei_max=CFG3.fei_L, fei_min=CFG3.fei_G,
                  fei_L=CFG3.fei_L, fei_G=CFG3.fei_G)
    o = po.Oracle(P)
    o.place_droplet(c, c, c, profile="tanh", width=2.0)
    assert np.max(np.abs(phi - o.get("fei"))) <= 1e-16


def test_ellipsoid_mean_on_surface_and_distance_along_axes():
    """phi equals the mean phi exactly where s = 1 (grid points on the
    semi-axes 20, 16, 16 of the 96^3 configuration); along each axis the
    first-order distance is exact, d = a - |x|; phi is above the mean inside
    and below outside."""
    L, axes = 96, (20.0, 16.0, 16.0)
    phi = configs.ellipsoid_phi(L, axes, CFG3)
    c = L // 2
    mean = 0.5 * (CFG3.fei_L + CFG3.fei_G)
    half = 0.5 * (CFG3.fei_L - CFG3.fei_G)
    for (z, y, x) in [(c, c, c + 20), (c, c, c - 20), (c, c + 16, c), (c, c - 16, c),
                      (c + 16, c, c), (c - 16, c, c)]:
        assert phi[z, y, x] == mean, (z, y, x)
    off = np.arange(L) - c
    for axis, a in ((2, 20.0), (1, 16.0), (0, 16.0)):
        idx = [c, c, c]
        vals = []
        for k in range(L):
            idx[axis] = k
            vals.append(phi[tuple(idx)])
        vals = np.array(vals)
        nz = off != 0
        want = mean + half * np.tanh((a - np.abs(off[nz])) / 2.0)
        assert np.max(np.abs(vals[nz] - want)) <= 1e-16, axis
    z, y, x = np.meshgrid(*(np.arange(L, dtype=np.float64) - c for _ in range(3)), indexing="ij")
    s = np.sqrt((x / axes[0]) ** 2 + (y / axes[1]) ** 2 + (z / axes[2]) ** 2)
    assert np.all(phi[s < 0.999] > mean) and np.all(phi[s > 1.001] < mean)
