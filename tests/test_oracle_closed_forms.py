"""Pins the CPU oracle on what mathematics fixes, independently of its code.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign, a transposed component, a wrong weight) fails it:
exact stencils on polynomial fields, the mirror-wall edge values, the van der
Waals critical point of the EOS, the symmetry of the density map, the moments
of the equilibria (up to the Inamuro pressure tensor), the streaming
permutation, the slip pairs as the specular reflection, and the first-step
Poisson special case.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

CX = np.array([0, 1, 0, 0, -1, 0, 0, 1, -1, 1, 1, -1, 1, -1, -1])
CY = np.array([0, 0, 1, 0, 0, -1, 0, 1, 1, -1, 1, -1, -1, 1, -1])
CZ = np.array([0, 0, 0, 1, 0, 0, -1, 1, 1, 1, -1, -1, -1, -1, 1])
C = np.stack([CX, CY, CZ], axis=1)


def grid(lx=8, ly=10, lz=12, **kw):
    P = po.Params(lx=lx, ly=ly, lz=lz, **kw)
    o = po.Oracle(P)
    z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
    return P, o, x.astype(float), y.astype(float), z.astype(float)


def test_lattice_moments_are_isotropic():
    E = np.array([2 / 9] + [1 / 9] * 6 + [1 / 72] * 8)
    F = np.array([-7 / 3] + [1 / 3] * 6 + [1 / 24] * 8)
    assert abs(E.sum() - 1) < 1e-15
    assert abs(F.sum()) < 1e-15
    np.testing.assert_allclose(np.einsum("i,ia,ib->ab", E, C, C), np.eye(3) / 3, atol=1e-16)
    np.testing.assert_allclose(np.einsum("i,ia,ib->ab", F, C, C), np.eye(3), atol=1e-15)
    assert np.all(np.einsum("ia,ib->ab", C, C) == 10 * np.eye(3))


def test_gradient_of_linear_field_is_exact_in_interior():
    P, o, x, y, z = grid()
    a, b, c = 0.37, -1.25, 2.5
    o.set("u", a * x + b * y + c * z)
    o.stage("gradients")
    inner = (slice(1, -1), slice(1, -1), slice(1, -1))
    np.testing.assert_allclose(o.get("u_x")[inner], a, rtol=1e-13)
    np.testing.assert_allclose(o.get("u_y")[inner], b, rtol=1e-13)
    np.testing.assert_allclose(o.get("u_z")[inner], c, rtol=1e-13)
    np.testing.assert_allclose(o.get("u_lap")[inner], 0.0, atol=1e-12)


def test_second_order_stencil_on_quadratics():
    P, o, x, y, z = grid()
    o.set("fei", x * x)
    o.set("w", x * y)
    o.stage("gradients")
    inner = (slice(1, -1), slice(1, -1), slice(1, -1))
    np.testing.assert_allclose(o.get("fei_lap")[inner], 2.0, rtol=1e-12)
    np.testing.assert_allclose(o.get("w_lap")[inner], 0.0, atol=1e-12)
    np.testing.assert_allclose(o.get("w_x")[inner], y[inner], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(o.get("w_y")[inner], x[inner], rtol=1e-12, atol=1e-12)


def test_periodic_y_sinusoid_central_difference():
    P, o, x, y, z = grid(ly=16)
    k = 2 * np.pi / 16
    o.set("v", np.sin(k * y))
    o.stage("gradients")
    yp, ym = np.sin(k * ((y + 1) % 16)), np.sin(k * ((y - 1) % 16))
    np.testing.assert_allclose(o.get("v_y"), (yp - ym) / 2, atol=1e-15)
    np.testing.assert_allclose(o.get("v_lap"), yp + ym - 2 * np.sin(k * y), atol=1e-14)
    np.testing.assert_allclose(o.get("v_x"), 0.0, atol=1e-16)


def test_x_mirror_edge_values():
    P, o, x, y, z = grid(lx=9)
    psi = np.cos(0.3 * x) + 0.1 * x
    o.set("fei", psi)
    o.stage("gradients")
    fx, lap = o.get("fei_x"), o.get("fei_lap")
    np.testing.assert_allclose(fx[..., 0], 0.0, atol=1e-16)
    np.testing.assert_allclose(fx[..., -1], 0.0, atol=1e-16)
    np.testing.assert_allclose(lap[..., 0], 2 * (psi[..., 1] - psi[..., 0]), rtol=1e-13)
    np.testing.assert_allclose(lap[..., -1], 2 * (psi[..., -2] - psi[..., -1]), rtol=1e-13)


def test_eos_has_van_der_waals_critical_point():
    a, b = 1.0, 6.7
    Tc = 8 * a / (27 * b)
    P = po.Params(lx=4, ly=4, lz=4, a=a, b=b, T=Tc)
    phic = 1 / (3 * b)
    h = 1e-4
    p = [po.eos(P, phic + k * h) for k in (-2, -1, 0, 1, 2)]
    d1 = (p[3] - p[1]) / (2 * h)
    d2 = (p[3] - 2 * p[2] + p[1]) / h ** 2
    assert abs(d1) < 1e-6 and abs(d2) < 1e-4  # a wrong sign or term gives O(0.1)
    assert abs(p[2] - 1 / (27 * b * b) * a) < 1e-15  # p_c = a / (27 b^2)
    # and the T-derivative: dp/dT = phi / (1 - b phi)
    P2 = po.Params(lx=4, ly=4, lz=4, a=a, b=b, T=Tc + 1e-3)
    assert abs((po.eos(P2, 0.05) - po.eos(P, 0.05)) / 1e-3 - 0.05 / (1 - b * 0.05)) < 1e-9


def test_density_map_clamps_and_is_odd_about_midpoint():
    P = po.Params(lx=4, ly=4, lz=4, rho_L=1000.0, rho_G=1.0, fei_L=0.093, fei_G=0.016)
    mid = (P.fei_L + P.fei_G) / 2
    assert po.rho_of_phi(P, P.fei_G) == P.rho_G
    assert po.rho_of_phi(P, P.fei_G - 0.01) == P.rho_G
    assert po.rho_of_phi(P, P.fei_L) == P.rho_L
    assert po.rho_of_phi(P, 0.2) == P.rho_L
    assert abs(po.rho_of_phi(P, mid) - (P.rho_L + P.rho_G) / 2) < 1e-12
    for d in (0.001, 0.013, 0.03):
        s = po.rho_of_phi(P, mid + d) + po.rho_of_phi(P, mid - d)
        assert abs(s - (P.rho_L + P.rho_G)) < 1e-10
    phis = np.linspace(P.fei_G, P.fei_L, 200)
    r = np.array([po.rho_of_phi(P, f) for f in phis])
    assert np.all(np.diff(r) >= 0)
    assert abs(po.rho_of_phi(P, P.fei_G + 1e-9) - P.rho_G) < 1e-3


def _smooth_state(o, x, y, z, P):
    k = 2 * np.pi
    lx, ly, lz = P.lx, P.ly, P.lz
    fei = P.fei_G + (P.fei_L - P.fei_G) * (0.5 + 0.4 * np.cos(k * x / lx) * np.sin(k * y / ly) * np.cos(k * z / lz))
    o.set("fei", fei)
    o.set("rho", np.array([[[po.rho_of_phi(P, f) for f in row] for row in pl] for pl in fei]))
    o.set("u", 1e-3 * np.sin(k * y / ly) * np.cos(k * x / lx))
    o.set("v", 2e-3 * np.cos(k * z / lz))
    o.set("w", -1.5e-3 * np.sin(k * x / lx) * np.sin(k * z / lz))
    return fei


def test_equilibrium_moments_first_collision():
    P, o, x, y, z = grid(lx=10, ly=12, lz=14, taug=0.8, k_g=2e-3)
    fei = _smooth_state(o, x, y, z, P)
    o.stage("collision")  # first call: f = f_eq, g = g_eq
    f, g = o.get("f"), o.get("g")
    u, v, w, rho = o.get("u"), o.get("v"), o.get("w"), o.get("rho")
    U = np.stack([u, v, w])
    # zeroth and first moments
    np.testing.assert_allclose(f.sum(0), fei, rtol=1e-13)
    np.testing.assert_allclose(np.einsum("ia,i...->a...", C, f), fei * U, atol=1e-16)
    div = o.get("u_x") + o.get("v_y") + o.get("w_z")
    np.testing.assert_allclose(g.sum(0), 1 + (P.taug - 0.5) * div, rtol=1e-13)
    np.testing.assert_allclose(np.einsum("ia,i...->a...", C, g), U, atol=1e-15)
    # second moment of f_eq is the Inamuro pressure tensor
    gx, gy, gz = o.get("fei_x"), o.get("fei_y"), o.get("fei_z")
    G = np.stack([gx, gy, gz])
    lap = o.get("fei_lap")
    p0 = np.vectorize(lambda q: po.eos(P, q))(fei)
    iso = p0 - P.k_f * fei * lap - P.k_f * (G ** 2).sum(0) / 2
    Pab = np.einsum("ia,ib,i...->ab...", C, C, f)
    for a_ in range(3):
        for b_ in range(3):
            expect = P.k_f * G[a_] * G[b_] + (iso if a_ == b_ else 0.0)
            np.testing.assert_allclose(Pab[a_, b_], expect, atol=1e-14)
    # second moment of g_eq: momentum flux + viscous + capillary terms
    R = np.stack([o.get("rho_x"), o.get("rho_y"), o.get("rho_z")])
    D = np.stack([np.stack([o.get(f"{c}_{d}") for d in "xyz"]) for c in "uvw"])  # D[a,b] = d_b u_a
    tg = P.taug - 0.5
    R2 = (R ** 2).sum(0)
    Mab = np.einsum("ia,ib,i...->ab...", C, C, g)
    divu = D[0, 0] + D[1, 1] + D[2, 2]
    for a_ in range(3):
        for b_ in range(3):
            if a_ == b_:
                visc = tg * (2 / 3 * D[a_, a_] + divu / 3)
                expect = 1 / 3 + U[a_] * U[a_] + visc + P.k_g / rho * (R[a_] * R[a_] - R2 / 3) - 2 / 3 * P.k_g / rho * R2
            else:
                visc = tg * (D[a_, b_] + D[b_, a_]) / 3
                expect = U[a_] * U[b_] + visc + P.k_g / rho * R[a_] * R[b_]
            np.testing.assert_allclose(Mab[a_, b_], expect, atol=1e-14)


def test_bgk_relaxation_and_viscous_force():
    P, o, x, y, z = grid(lx=8, ly=10, lz=12, tauf=0.8, taug=0.7)
    _smooth_state(o, x, y, z, P)
    o.stage("collision")
    f_eq0 = o.get("f")
    rng = np.random.default_rng(3)
    f_old = f_eq0 * (1 + 0.01 * rng.standard_normal(f_eq0.shape))
    g_old = o.get("g") * (1 + 0.01 * rng.standard_normal(f_eq0.shape))
    o.set("f", f_old)
    o.set("g", g_old)
    o.stage("collision")  # second call: BGK with the same equilibria
    f_new, g_new = o.get("f"), o.get("g")
    np.testing.assert_allclose(f_new, f_old - (f_old - f_eq0) / P.tauf, rtol=1e-12, atol=1e-15)
    # first moment of g: (1 - 1/tau) m_old + u/tau + (mu/rho) lap(u)
    rho = o.get("rho")
    mu = (rho - P.rho_G) * (P.mu_L - P.mu_G) / (P.rho_L - P.rho_G) + P.mu_G
    U = np.stack([o.get("u"), o.get("v"), o.get("w")])
    L = np.stack([o.get("u_lap"), o.get("v_lap"), o.get("w_lap")])
    m_old = np.einsum("ia,i...->a...", C, g_old)
    m_new = np.einsum("ia,i...->a...", C, g_new)
    np.testing.assert_allclose(m_new, (1 - 1 / P.taug) * m_old + U / P.taug + mu / rho * L, atol=1e-15)


def test_streaming_moves_each_population_along_its_velocity():
    P, o, x, y, z = grid(lx=6, ly=7, lz=5)
    rng = np.random.default_rng(1)
    f = rng.standard_normal((15, 5, 7, 6))
    o.set("f", f)
    o.stage("stream_fg")
    out = o.get("f")
    for q in range(15):
        expect = np.roll(f[q], shift=(CZ[q], CY[q], CX[q]), axis=(0, 1, 2))
        assert np.array_equal(out[q], expect)


def test_slip_pairs_are_the_specular_reflection():
    P, o, x, y, z = grid(lx=6, ly=4, lz=4)
    rng = np.random.default_rng(2)
    f = rng.standard_normal((15, 4, 4, 6))
    o.set("f", f)
    o.stage("bc_fg")
    out = o.get("f")
    mirror = {q: int(np.where((CX == -CX[q]) & (CY == CY[q]) & (CZ == CZ[q]))[0][0]) for q in range(15)}
    for q in range(15):
        if CX[q] == 1:   # entering from the left wall: mirror of the outgoing one
            assert np.array_equal(out[q][..., 0], f[mirror[q]][..., 0])
        else:
            assert np.array_equal(out[q][..., 0], f[q][..., 0])
        if CX[q] == -1:
            assert np.array_equal(out[q][..., -1], f[mirror[q]][..., -1])
        else:
            assert np.array_equal(out[q][..., -1], f[q][..., -1])
        assert np.array_equal(out[q][..., 1:-1], f[q][..., 1:-1])


def test_macro_and_pressure_sums():
    P, o, x, y, z = grid(lx=5, ly=4, lz=3)
    rng = np.random.default_rng(4)
    f = 0.01 * rng.random((15, 3, 4, 5))
    g = rng.standard_normal((15, 3, 4, 5))
    o.set("f", f)
    o.set("g", g)
    o.stage("macro")
    np.testing.assert_allclose(o.get("fei"), f.sum(0), rtol=1e-14)
    for name, c in zip("uvw", (CX, CY, CZ)):
        np.testing.assert_allclose(o.get(name), np.einsum("i,i...->...", c, g), atol=1e-14)
    h = rng.standard_normal((15, 3, 4, 5))
    o.set("h", h)
    o.stage("getp")
    np.testing.assert_allclose(o.get("p"), h.sum(0), atol=1e-14)


def test_pressure_collision_conserves_p_minus_source():
    P, o, x, y, z = grid(lx=5, ly=4, lz=3)
    rng = np.random.default_rng(5)
    o.set("rho", 1 + 10 * rng.random((3, 4, 5)))
    o.set("p", rng.standard_normal((3, 4, 5)))
    h = rng.standard_normal((15, 3, 4, 5))
    o.set("h", h)
    ux, vy, wz = (rng.standard_normal((3, 4, 5)) for _ in range(3))
    o.set("u_x", ux); o.set("v_y", vy); o.set("w_z", wz)
    o.stage("collision_p")
    E = np.array([2 / 9] + [1 / 9] * 6 + [1 / 72] * 8)
    tau = 1 / o.get("rho") + 0.5
    S = ux + vy + wz
    np.testing.assert_allclose(o.get("h").sum(0), h.sum(0) - (h.sum(0) - o.get("p")) / tau - S / 3, atol=1e-13)
    # each population relaxes to E_q p
    hn = o.get("h")
    np.testing.assert_allclose(hn, h - (h - E[:, None, None, None] * o.get("p")) / tau - E[:, None, None, None] * S / 3,
                               atol=1e-14)


def test_first_solve_from_rest_converges_in_exactly_one_check():
    P = po.Params(lx=8, ly=8, lz=8, DD=4.0)
    o = po.Oracle(P)
    o.place_droplet(4.0, 4.0, 4.0, profile="tanh", width=2.0)
    its = o.step(1)
    assert its == [100]
    assert np.all(o.get("p") == 0.0)
    d = o.diag()
    assert d.relative_error == 0.0 and d.converged == 1


def test_correct_uvw_is_central_pressure_difference():
    P, o, x, y, z = grid(lx=6, ly=8, lz=5)
    k = 2 * np.pi / 8
    pfield = np.sin(k * y) + 0.2 * x * x
    o.set("p", pfield)
    rho = 2.0 + 0.1 * z
    o.set("rho", rho)
    o.stage("correct")
    v = o.get("v")
    dp = np.sin(k * ((y + 1) % 8)) - np.sin(k * ((y - 1) % 8))
    np.testing.assert_allclose(v, -dp / (2 * rho), atol=1e-15)
    u = o.get("u")
    np.testing.assert_allclose(u[..., 0], 0.0, atol=1e-16)
    np.testing.assert_allclose(u[..., -1], 0.0, atol=1e-16)
    np.testing.assert_allclose(u[..., 2], -(0.2 * 9 - 0.2 * 1) / (2 * rho[..., 2]), rtol=1e-14)
