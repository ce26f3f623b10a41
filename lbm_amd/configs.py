"""Initial fields of the benchmark workloads, built on the host in float64.

The reference initialises droplets in its constructors (sharp two-droplet
state, CPU/Inamuro.cpp:550-587; tanh or sharp droplets with a velocity,
GPU/Inamuro.cpp:491-566).  The workloads below use the same tanh profile and
the same rules -- phi = fei_min + lf (fei_max - fei_min), rho the linear blend
rho_G + lf (rho_L - rho_G), velocity v lf, and a cell is written only where
the new phi exceeds the current one -- with the droplet radius made a function
of direction:

* ``head_on_collision``: two droplets moving towards each other along x,
  each radius perturbed by real spherical harmonics 2 <= l <= 8 with seeded
  Gaussian coefficients, scaled to a given RMS over the sphere;
* ``ellipsoid``: a droplet whose interface is an ellipsoid.

Fields are float64 arrays [lz][ly][lx] (x fastest), ready for
``InamuroHIP.initFromCPU`` and the oracle's ``set``.
"""
from __future__ import annotations

import math

import numpy as np

SH_SEED = 2026
SH_LMIN, SH_LMAX = 2, 8


def legendre_table(lmax: int, x):
    """Associated Legendre functions P_l^m(x), 0 <= m <= l <= lmax, without the
    Condon-Shortley phase, by the standard three-term recurrence in l:
    P_m^m = (2m-1)!! (1-x^2)^(m/2), P_{m+1}^m = (2m+1) x P_m^m,
    (l-m) P_l^m = (2l-1) x P_{l-1}^m - (l+m-1) P_{l-2}^m."""
    x = np.asarray(x, dtype=np.float64)
    s = np.sqrt(np.maximum(0.0, 1.0 - x * x))
    P = {}
    pmm = np.ones_like(x)
    for m in range(lmax + 1):
        if m > 0:
            pmm = pmm * (2 * m - 1) * s
        P[(m, m)] = pmm
        if m + 1 <= lmax:
            P[(m + 1, m)] = (2 * m + 1) * x * pmm
        for l in range(m + 2, lmax + 1):
            P[(l, m)] = ((2 * l - 1) * x * P[(l - 1, m)] - (l + m - 1) * P[(l - 2, m)]) / (l - m)
    return P


def real_sph_harm_table(lmax: int, theta, phi):
    """Orthonormal real spherical harmonics Y_lm for all l <= lmax:
    m > 0: sqrt(2) N P_l^m(cos t) cos(m p); m < 0: sqrt(2) N P_l^|m| sin(|m| p);
    m = 0: N P_l(cos t), with N = sqrt((2l+1)/(4 pi) (l-|m|)!/(l+|m|)!)."""
    P = legendre_table(lmax, np.cos(theta))
    Y = {}
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            am = abs(m)
            norm = math.sqrt((2 * l + 1) / (4.0 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
            if m == 0:
                Y[(l, m)] = norm * P[(l, 0)]
            elif m > 0:
                Y[(l, m)] = math.sqrt(2.0) * norm * P[(l, am)] * np.cos(am * phi)
            else:
                Y[(l, m)] = math.sqrt(2.0) * norm * P[(l, am)] * np.sin(am * phi)
    return Y


def sh_modes(lmin: int = SH_LMIN, lmax: int = SH_LMAX):
    return [(l, m) for l in range(lmin, lmax + 1) for m in range(-l, l + 1)]


def sh_coefficients(n_droplets: int = 2, seed: int = SH_SEED, lmin: int = SH_LMIN, lmax: int = SH_LMAX):
    """Standard-normal coefficients a_lm, one set per droplet, in sh_modes order."""
    rng = np.random.default_rng(seed)
    modes = sh_modes(lmin, lmax)
    return [rng.standard_normal(len(modes)) for _ in range(n_droplets)]


def radius_perturbation(coeffs, sigma: float, theta, phi, lmin: int = SH_LMIN, lmax: int = SH_LMAX):
    """dR(theta, phi) = s sum a_lm Y_lm with s chosen so that the RMS of dR over
    the sphere is sigma (by orthonormality, s = sigma sqrt(4 pi / sum a^2))."""
    coeffs = np.asarray(coeffs, dtype=np.float64)
    out = np.zeros(np.broadcast(theta, phi).shape)
    norm2 = float((coeffs * coeffs).sum())
    if norm2 == 0.0:
        return out
    scale = sigma * math.sqrt(4.0 * math.pi / norm2)
    Y = real_sph_harm_table(lmax, theta, phi)
    for a, lm in zip(coeffs, sh_modes(lmin, lmax)):
        out += a * Y[lm]
    return scale * out


def background(params, shape):
    """initializeArrays, CPU/Inamuro.cpp:519-524."""
    return dict(fei=np.full(shape, params.fei_min), rho=np.full(shape, params.rho_G),
                u=np.zeros(shape), v=np.zeros(shape), w=np.zeros(shape), p=np.zeros(shape))


def add_droplet(f, params, center, radius_fn, velocity=(0.0, 0.0, 0.0), width: float = 2.0,
                reach: float = None):
    """Writes one tanh droplet into fields f (place_droplet's rules,
    GPU/Inamuro.cpp:499-533).  radius_fn(theta, phi) -> local radius.  Cells
    farther than `reach` from the centre are skipped (lf rounds to 0 there);
    None visits every cell."""
    lz, ly, lx = f["fei"].shape
    box = []
    for c, n in zip((center[2], center[1], center[0]), (lz, ly, lx)):
        if reach is None:
            box.append(slice(0, n))
        else:
            box.append(slice(max(0, int(math.floor(c - reach))), min(n, int(math.ceil(c + reach)) + 1)))
    z, y, x = np.meshgrid(np.arange(box[0].start, box[0].stop, dtype=np.float64),
                          np.arange(box[1].start, box[1].stop, dtype=np.float64),
                          np.arange(box[2].start, box[2].stop, dtype=np.float64), indexing="ij")
    dx, dy, dz = x - center[0], y - center[1], z - center[2]
    r = np.sqrt(dx * dx + dy * dy + dz * dz)
    theta = np.arccos(np.clip(np.where(r > 0, dz / np.where(r > 0, r, 1.0), 1.0), -1.0, 1.0))
    phi = np.arctan2(dy, dx)
    R = radius_fn(theta, phi)
    lf = 0.5 * (1.0 - np.tanh((r - R) / width))
    local = params.fei_min + lf * (params.fei_max - params.fei_min)
    sub = tuple(box)
    sel = (lf > 0.0) & (local > f["fei"][sub])
    f["fei"][sub][sel] = local[sel]
    f["rho"][sub][sel] = (params.rho_G + lf * (params.rho_L - params.rho_G))[sel]
    for k, vel in zip(("u", "v", "w"), velocity):
        if vel != 0.0 or k == "w":
            f[k][sub][sel] = (vel * lf)[sel]
    return f


def head_on_collision(params, shape=(128, 128, 256), radius: float = 24.0, centers_x=(80.0, 176.0),
                      speed: float = 0.02, sigma: float = 0.5, coeffs=None, width: float = 2.0):
    """Two droplets of mean radius `radius` on the x axis through the box
    centre, moving at +speed and -speed along x, radii perturbed by
    radius_perturbation (RMS sigma).  shape is (lz, ly, lx)."""
    lz, ly, lx = shape
    if coeffs is None:
        coeffs = sh_coefficients(2)
    f = background(params, shape)
    # tanh((r - R)/w) rounds to 1 beyond r - R = 19 w; |Y_lm| <= sqrt((2l+1)/(2 pi))
    # bounds the perturbation
    for cx, vx, a in zip(centers_x, (speed, -speed), coeffs):
        a = np.asarray(a, dtype=np.float64)
        n2 = float(np.dot(a, a))
        bound = 0.0 if n2 == 0.0 else (sigma * math.sqrt(4.0 * math.pi / n2) * float(np.abs(a).sum())
                                       * math.sqrt((2 * SH_LMAX + 1) / (2.0 * math.pi)))
        add_droplet(f, params, (cx, ly / 2.0, lz / 2.0),
                    lambda t, p, a=a: radius + radius_perturbation(a, sigma, t, p),
                    velocity=(vx, 0.0, 0.0), width=width, reach=radius + bound + 20.0 * width)
    return f


def ellipsoid_phi(L: int, axes, params, width: float = 2.0):
    """phi of a droplet at the centre of an L^3 box whose interface is the
    ellipsoid with semi-axes `axes`: mean + (phi_L - phi_G)/2 tanh(d / width),
    d the first-order signed distance (1 - s) s / |grad s|,
    s = |(x/a, y/b, z/c)|."""
    c = L / 2.0
    z, y, x = np.meshgrid(*(np.arange(L, dtype=np.float64) - c for _ in range(3)), indexing="ij")
    a, b, cc = axes
    s = np.sqrt((x / a) ** 2 + (y / b) ** 2 + (z / cc) ** 2)
    gs = np.sqrt((x / a ** 2) ** 2 + (y / b ** 2) ** 2 + (z / cc ** 2) ** 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        d = np.where(s > 0, (1.0 - s) * s / np.where(gs > 0, gs, 1.0), min(axes))
    return (0.5 * (params.fei_L + params.fei_G) +
            0.5 * (params.fei_L - params.fei_G) * np.tanh(d / width))
