"""Pins the oracle's restatement of the reference GPU solver's Poisson
accelerators (GPU/InamuroCUDA.cu:503-1013, 1198-1323, driven by
doPressurePoisson :1816-2349; setters :2446-2540).

The reference has them only in CUDA, which cannot be built here, and its CPU
solver has none, so nothing compiled pins them directly.  They are pinned by
* equivalences with the strict path, which the step fixtures pin bit for bit
  to the compiled reference: the pull ("fused") form equals the split form
  bit for bit at any relaxation scale, the accelerated loop at its defaults
  equals the strict loop bit for bit, and onepass with unit relaxation equals
  split to rounding;
* closed forms on uniform states (relaxation scale, fixed-point relaxation)
  and on fields that vary along one axis (the scalar Jacobi sweep's
  neighbour sum, x-mirror and source term);
* rebuilds from those pinned primitives: Anderson(1) from two onepass images,
  the two-grid correction and the spatial fraction from the strict loop's
  pressure change, the source-aware start from its formula -- each with a
  plausible mistake shown to fail.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

EW = np.array([2 / 9] + [1 / 9] * 6 + [1 / 72] * 8)
CX = np.array([0, 1, 0, 0, -1, 0, 0, 1, -1, 1, 1, -1, 1, -1, -1])
CY = np.array([0, 0, 1, 0, 0, -1, 0, 1, 1, -1, 1, -1, -1, 1, -1])
CZ = np.array([0, 0, 0, 1, 0, 0, -1, 1, 1, 1, -1, -1, -1, -1, 1])
SMOKE = po.Params(lx=16, ly=32, lz=32, DD=8.0)
SMALL = po.Params(lx=12, ly=10, lz=14, rho_L=1000.0, rho_G=1.0, mu_L=0.008, mu_G=8e-5, DD=6.0,
                  fei_max=0.093, fei_min=0.016, fei_L=0.093, fei_G=0.016, k_g=1e-4)
STATE_ARRAYS = ("fei", "rho", "u", "v", "w", "p", "u_x", "v_y", "w_z", "p_prev", "h")


def _smoke():
    o = po.Oracle(SMOKE)
    o.init_two_droplets_sharp()
    return o


def _small():
    o = po.Oracle(SMALL)
    o.place_droplet(5.5, 4.5, 7.0, 0.01, 0.0, -0.005, "tanh", 2.0)
    return o


def _run(make, steps, **accel):
    o = make()
    if accel:
        o.set_poisson_accel(**accel)
    it = o.step(steps)
    out = {k: o.get(k) for k in ("fei", "rho", "u", "v", "w", "p", "h", "p_prev")}
    d = o.diag()
    o.close()
    return it, out, d


def _mid_solve_state(seed=3):
    """A droplet state after one strict step, with h pushed off equilibrium
    by seeded noise so that no term of the iteration vanishes."""
    o = _small()
    o.step(1)
    st = {k: o.get(k) for k in STATE_ARRAYS}
    o.close()
    rng = np.random.default_rng(seed)
    st["h"] = st["h"] + 1e-3 * rng.standard_normal(st["h"].shape) * EW[:, None, None, None]
    return st


def _load(params, st, **accel):
    o = po.Oracle(params)
    for k, v in st.items():
        o.set(k, v)
    if accel:
        o.set_poisson_accel(**accel)
    return o


def _bitwise_equal(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k]), k


# ---------------------------------------------------------------------------
# equivalences with the strict (reference-pinned) path
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("scale", [1.0, 0.85])
def test_pull_form_equals_split_form_bitwise(scale):
    """collisionPressureStreamKernel (collision at the source, pull) with or
    without the fused boundary/pressure kernel gives the split sequence's
    values bit for bit; the split sequence runs the oracle's literal
    collision_p / stream / slip / getp."""
    it_s, out_s, _ = _run(_smoke, 2, variant="split", pressure_relax_scale=scale,
                          spatial_diagnostics=1)
    it_f, out_f, _ = _run(_smoke, 2, variant="fused", pressure_relax_scale=scale)
    it_fb, out_fb, _ = _run(_smoke, 2, variant="fused", fused_boundary=1, pressure_relax_scale=scale)
    assert it_s == it_f == it_fb
    _bitwise_equal(out_s, out_f)
    _bitwise_equal(out_s, out_fb)


def test_accelerated_loop_at_defaults_equals_strict_loop():
    """With only the spatial diagnostic on, the accelerated loop (its own
    gauge, relative change and check) reproduces the strict loop -- which is
    bit-identical to the compiled reference -- bit for bit."""
    it0, out0, d0 = _run(_smoke, 3)
    it1, out1, d1 = _run(_smoke, 3, spatial_diagnostics=1)
    assert it0 == it1 == [800, 1300, 1300]
    _bitwise_equal(out0, out1)
    assert d1.relative_error == d0.relative_error
    assert d1.pressure_gauge_max_shift == d0.pressure_gauge_max_shift
    assert d1.block_size == 4 and d1.block_count == 4 * 8 * 8
    assert 0.0 <= d1.block_low_frequency_fraction <= 1.0
    assert np.isnan(d0.block_low_frequency_fraction) and d0.block_count == 0


def test_onepass_with_unit_relaxation_matches_split_to_rounding():
    """h_old + 1 (image - h_old) differs from the image only by rounding."""
    it_s, out_s, _ = _run(_small, 2)
    it_o, out_o, _ = _run(_small, 2, variant="onepass")
    assert it_s == it_o
    for k in ("fei", "rho", "u", "v", "w", "p"):
        scale = max(np.abs(out_s[k]).max(), 1e-300)
        assert np.abs(out_o[k] - out_s[k]).max() / scale < 1e-12, k


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------

def _uniform(rho, c, d, s):
    lx, ly, lz = 5, 4, 6
    o = po.Oracle(po.Params(lx=lx, ly=ly, lz=lz, DD=2.0))
    o.set("rho", rho)
    o.set("h", EW[:, None, None, None] * c * np.ones((15, lz, ly, lx)))
    o.set("p", c + d)
    o.set("u_x", s)
    o.set("v_y", 0.0)
    o.set("w_z", 0.0)
    return o


@pytest.mark.parametrize("variant", ["split", "fused", "onepass"])
@pytest.mark.parametrize("scale,omega", [(0.7, 1.0), (1.3, 0.6)])
def test_relaxation_closed_form_on_a_uniform_state(variant, scale, omega):
    """h = E c, p = c + d, source s, density rho everywhere: one iteration
    gives the image E (c + scale d / tau - s / 3) (stream and slip leave a
    uniform field alone); onepass then takes c + omega (image - c); p is the
    q-sum.  A relaxation scale on the wrong term, or omega applied the wrong
    way round, moves the result by O(d) or O(s)."""
    rho, c, d, s = 4.0, 0.3, 0.05, 0.06
    o = _uniform(rho, c, d, s)
    o.set_poisson_accel(variant=variant, pressure_relax_scale=scale, fixed_point_relax=omega)
    o.poisson_accel_iterations(1)
    tau = 1.0 / rho + 0.5
    image = c + scale * d / tau - s / 3.0
    val = c + omega * (image - c) if variant == "onepass" else image
    h, p = o.get("h"), o.get("p")
    np.testing.assert_allclose(h, EW[:, None, None, None] * val * np.ones_like(h), rtol=1e-14, atol=1e-17)
    np.testing.assert_allclose(p, val, rtol=1e-14)
    o.close()


def test_scalar_jacobi_eigenfunction_along_y():
    """p = cos(k y) is an eigenfunction of the D3Q15 neighbour sum: the 6 axis
    and 8 diagonal neighbours give 4 p(y) + 5 (p(y+1) + p(y-1)), so with no
    source one sweep multiplies p by (4 + 10 cos k) / 14."""
    lx, ly, lz = 6, 16, 5
    o = po.Oracle(po.Params(lx=lx, ly=ly, lz=lz, DD=2.0))
    k = 2 * np.pi * 3 / ly
    p0 = np.broadcast_to(np.cos(k * np.arange(ly))[None, :, None], (lz, ly, lx))
    o.set("p", p0)
    for a in ("u_x", "v_y", "w_z"):
        o.set(a, 0.0)
    o.set_poisson_accel(variant="scalar")
    o.poisson_accel_iterations(1)
    np.testing.assert_allclose(o.get("p"), (4 + 10 * np.cos(k)) / 14 * p0, rtol=0, atol=1e-15)
    o.close()


def test_scalar_jacobi_x_mirror_and_source():
    """Along x the neighbours beyond a wall are mirrored
    (pressure_bounce_direction): p_next(0) = (4 p(0) + 10 p(1) - 5 rhs) / 14
    and likewise at lx-1, interior cells 4 p + 5 (p(x-1) + p(x+1)); the
    source is rhs = scale * rho * div.  A periodic wrap in x would read the
    far wall."""
    lx, ly, lz = 7, 3, 4
    o = po.Oracle(po.Params(lx=lx, ly=ly, lz=lz, DD=2.0))
    rng = np.random.default_rng(11)
    px = rng.standard_normal(lx)
    o.set("p", np.broadcast_to(px[None, None, :], (lz, ly, lx)))
    rho, div, scale = 3.0, 0.02, 2.5
    o.set("rho", rho)
    o.set("u_x", div)
    o.set("v_y", 0.0)
    o.set("w_z", 0.0)
    o.set_poisson_accel(variant="scalar", scalar_source_scale=scale)
    o.poisson_accel_iterations(1)
    left = np.concatenate([[px[1]], px[:-1]])
    right = np.concatenate([px[1:], [px[-2]]])
    expect = (4 * px + 5 * (left + right) - 5 * scale * rho * div) / 14
    got = o.get("p")
    np.testing.assert_allclose(got, np.broadcast_to(expect[None, None, :], got.shape), rtol=1e-13, atol=1e-15)
    o.close()


# ---------------------------------------------------------------------------
# rebuilds from pinned primitives
# ---------------------------------------------------------------------------

def _onepass_image(st, omega=1.0):
    """One plain onepass iteration from state st (pinned above)."""
    o = _load(SMALL, st, variant="onepass", fixed_point_relax=omega)
    o.poisson_accel_iterations(1)
    h, p = o.get("h"), o.get("p")
    o.close()
    return h, p


def _q_sum(h):
    acc = np.zeros(h.shape[1:])
    for q in range(15):
        acc = acc + h[q]
    return acc


def _anderson_rebuild(st, omega, beta):
    i1, p1 = _onepass_image(st, omega)
    st1 = dict(st, h=i1, p=p1)
    i2, _ = _onepass_image(st1, omega)
    prev_res = i1 - st["h"]
    r = i2 - i1
    delta = r - prev_res
    num = float(np.sum(prev_res * delta))
    den = float(np.sum(delta * delta))
    alpha = -num / den if den > 1e-300 else 1.0
    alpha_clamped = min(max(alpha, -beta), 1.0 + beta)
    acc = (1.0 - alpha_clamped) * i1 + alpha_clamped * i2
    return alpha, alpha_clamped, acc, i1, i2


@pytest.mark.parametrize("omega", [1.0, 0.8])
def test_anderson_m1_rebuilt_from_two_onepass_images(omega):
    """Iteration 1 stores the history (residual I1 - h0, image I1) and steps
    to I1; iteration 2 mixes the images with alpha = -<R1, R2 - R1> /
    |R2 - R1|^2: h2 = (1 - alpha) I1 + alpha I2, p2 = sum_q h2."""
    st = _mid_solve_state()
    alpha, alpha_c, acc, i1, i2 = _anderson_rebuild(st, omega, beta=1.0)
    assert alpha == alpha_c  # inside the clamp for this state
    o = _load(SMALL, st, variant="onepass", fixed_point_relax=omega, anderson_m1=1)
    o.poisson_accel_iterations(2)
    h, p = o.get("h"), o.get("p")
    got_alpha = o.diag().anderson_alpha
    o.close()
    assert got_alpha == pytest.approx(alpha, rel=1e-9)
    scale = np.abs(acc).max()
    assert np.abs(h - acc).max() / scale < 1e-12
    np.testing.assert_allclose(p, _q_sum(acc), rtol=1e-12, atol=1e-14 * np.abs(p).max())
    # the weights the other way round, or the unaccelerated image, differ
    swapped = alpha * i1 + (1.0 - alpha) * i2
    assert np.abs(h - swapped).max() / scale > 1e-9
    assert np.abs(h - i2).max() / scale > 1e-9


def test_anderson_clamps_alpha_and_falls_back_on_zero_denominator():
    """alpha is clamped to [-beta, 1 + beta]: under-relaxed onepass
    (omega = 0.3) makes Anderson extrapolate, alpha = 1.6 on the seeded
    state, inside the default bound 2 and cut to 1.25 by beta = 0.25.  A zero
    state has zero residuals, hence a zero denominator, and alpha = 1."""
    st = _mid_solve_state()
    alpha, _, _, i1, i2 = _anderson_rebuild(st, 0.3, beta=1.0)
    assert 1.25 < alpha < 2.0
    for beta, expect in ((1.0, alpha), (0.25, 1.25)):
        o = _load(SMALL, st, variant="onepass", fixed_point_relax=0.3, anderson_m1=1,
                  anderson_beta_max=beta)
        o.poisson_accel_iterations(2)
        got = o.diag().anderson_alpha
        h = o.get("h")
        o.close()
        assert got == pytest.approx(expect, rel=1e-9)
        acc = (1.0 - got) * i1 + got * i2
        assert np.abs(h - acc).max() / np.abs(acc).max() < 1e-12

    o = _uniform(4.0, 0.0, 0.0, 0.0)
    o.set_poisson_accel(variant="onepass", anderson_m1=1)
    o.poisson_accel_iterations(2)
    assert o.diag().anderson_alpha == 1.0
    o.close()


def _one_check(st, ci, **accel):
    """One check interval of ci iterations with the watchdog stopping after
    the first check, from state st (h, p, p_prev as given)."""
    o = _load(SMALL, st, **accel)
    o.set_poisson(check_interval=ci, tol=1e-12, iter_cap=ci)
    o.stage("poisson")
    out = {k: o.get(k) for k in ("p", "h", "p_prev")}
    d = o.diag()
    o.close()
    return out, d


def _blocks(delta, b=4):
    lz, ly, lx = delta.shape
    nz, ny, nx = -(-lz // b), -(-ly // b), -(-lx // b)
    sums = np.zeros((nz, ny, nx))
    cells = np.zeros((nz, ny, nx))
    for z in range(lz):
        for y in range(ly):
            for x in range(lx):
                sums[z // b, y // b, x // b] += delta[z, y, x]
                cells[z // b, y // b, x // b] += 1
    return sums, cells


def test_two_grid_correction_rebuilt_from_the_strict_check():
    """At a check that has not converged, the 4^3 block sums B_b of the
    gauged p - p_prev give correction = strength B_b / cells_b (edge blocks
    truncated: the grid is 12x10x14), added to p and to h as E correction,
    and p_prev takes the corrected p.  The strict loop's state at that check
    is the uncorrected one."""
    st = _mid_solve_state()
    st["p_prev"] = st["p"] - 0.01 * np.sin(np.arange(st["p"].size).reshape(st["p"].shape))
    ci = 3
    plain, d0 = _one_check(st, ci)
    corr_out, d1 = _one_check(st, ci, two_grid=1, two_grid_strength=0.7)
    assert d0.iteration_count == d1.iteration_count == ci and not d1.converged
    sums, cells = _blocks(plain["p"] - st["p_prev"])
    corr_block = 0.7 * sums / cells
    lz, ly, lx = st["p"].shape
    corr = corr_block[np.arange(lz)[:, None, None] // 4, np.arange(ly)[None, :, None] // 4,
                      np.arange(lx)[None, None, :] // 4]
    scale = np.abs(plain["p"]).max()
    assert np.abs(corr_out["p"] - (plain["p"] + corr)).max() / scale < 1e-13
    assert np.array_equal(corr_out["p_prev"], corr_out["p"])
    hexp = plain["h"] + EW[:, None, None, None] * corr[None]
    assert np.abs(corr_out["h"] - hexp).max() / np.abs(hexp).max() < 1e-13
    assert d1.block_size == 4 and d1.block_count == corr_block.size
    assert np.isnan(d1.block_low_frequency_fraction)
    # the correction is not a plain mean over the grid, nor unscaled by strength
    assert np.abs(corr_out["p"] - (plain["p"] + 0.7 * np.mean(plain["p"] - st["p_prev"]))).max() / scale > 1e-6
    assert np.abs(corr_out["p"] - (plain["p"] + sums.ravel().mean() * 0)).max() / scale > 1e-6


def test_spatial_fraction_rebuilt_from_the_strict_check():
    """block_low_frequency_fraction = sum_b |B_b| / sum |p - p_prev|."""
    st = _mid_solve_state()
    st["p_prev"] = st["p"] * 0.9
    ci = 2
    plain, d0 = _one_check(st, ci)
    _, d1 = _one_check(st, ci, spatial_diagnostics=1)
    delta = plain["p"] - st["p_prev"]
    sums, _ = _blocks(delta)
    frac = np.abs(sums).sum() / np.abs(delta).sum()
    assert d1.block_low_frequency_fraction == pytest.approx(frac, rel=1e-12)
    assert d1.relative_error == d0.relative_error


def test_source_aware_start_equals_strict_solve_from_its_formula():
    """The source-aware start h = E p - scale tau_h (E / 3) S, with the
    projected source S, followed by the strict loop, equals the strict loop
    started from that h.  The source is dyadic with zero mean, so the
    projection is exact and numpy forms the same h."""
    st = _mid_solve_state()
    rng = np.random.default_rng(5)
    shape = st["p"].shape
    def zero_sum_dyadic():
        k = rng.integers(-8, 9, size=shape)
        return (k - np.roll(k, 1)) / 1024.0  # telescoping integers: the sum is exactly 0

    st["u_x"] = zero_sum_dyadic()
    st["v_y"] = zero_sum_dyadic()
    st["w_z"] = np.zeros(shape)
    assert st["u_x"].sum() == 0.0 and st["v_y"].sum() == 0.0
    scale = 1.5
    tau = 1.0 / st["rho"] + 0.5
    S = (st["u_x"] + st["v_y"]) + st["w_z"]
    h0 = np.stack([EW[q] * st["p"] - scale * tau * ((EW[q] / 3.0) * S) for q in range(15)])
    a = _load(SMALL, st, source_aware_hh_init=1, source_aware_hh_scale=scale)
    a.stage("poisson")
    b = _load(SMALL, dict(st, h=h0))
    b.stage("poisson")
    assert a.diag().iteration_count == b.diag().iteration_count
    for k in ("p", "h", "p_prev"):
        assert np.array_equal(a.get(k), b.get(k)), k
    # and it is not the plain equilibrium start
    c = _load(SMALL, dict(st, h=np.stack([EW[q] * st["p"] for q in range(15)])))
    c.stage("poisson")
    assert not np.array_equal(a.get("p"), c.get("p"))
    for o in (a, b, c):
        o.close()


# ---------------------------------------------------------------------------
# settings and refusals
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("kw,msg", [
    ({"pressure_relax_scale": 0.0}, "pressure relaxation scale must be positive"),
    ({"fixed_point_relax": -1.0}, "Poisson fixed-point relaxation must be positive"),
    ({"anderson_beta_max": -0.1}, "Poisson Anderson beta max must be non-negative"),
    ({"two_grid_strength": float("nan")}, "Poisson two-grid strength must be non-negative"),
])
def test_setter_validation_messages(kw, msg):
    o = po.Oracle(SMALL)
    with pytest.raises(po.OracleError, match=msg):
        o.set_poisson_accel(**kw)
    o.close()


def test_dual_residual_refuses_accelerators():
    """GPU/InamuroCUDA.cu:1846-1865."""
    o = _smoke()
    o.set_poisson(dual=True)
    o.set_poisson_accel(variant="onepass")
    with pytest.raises(po.OracleError, match="fixed-point residual requires frozen split"):
        o.step(1)
    o.close()
    o = _smoke()
    o.set_poisson(dual=True)
    o.set_poisson_accel(spatial_diagnostics=1)
    with pytest.raises(po.OracleError, match="no spatial diagnostics"):
        o.step(1)
    o.close()


def test_adjoint_refuses_accelerators():
    o = po.Oracle(SMALL)
    o.set_poisson_accel(variant="scalar")
    phi = np.full(o.shape, SMALL.fei_min)
    with pytest.raises(po.OracleError, match="strict split Poisson map"):
        o.objective(phi, phi, [100])
    o.close()
