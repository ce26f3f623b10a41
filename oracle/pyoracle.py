"""ctypes wrapper of the CPU oracle (oracle/liblbm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the cpu_baseline leg of bench.py, always as the checker and never as the thing
measured or shipped.  Arrays cross the boundary in the canonical layout:
macro fields [z][y][x] (x fastest), distributions [q][z][y][x].
"""
from __future__ import annotations

import ctypes
import os
import struct
from dataclasses import dataclass, fields

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbm_oracle.so")
REF_PROBE = os.path.join(_HERE, "_ref", "ref_probe")
REF_PROBE_INIT = os.path.join(_HERE, "_ref", "ref_probe_init")

Q = 15
FIELD_NAMES = ("fei", "rho", "u", "v", "w", "p")
ARRAY_IDS = {
    "fei": 0, "rho": 1, "u": 2, "v": 3, "w": 4, "p": 5,
    "fei_x": 6, "fei_y": 7, "fei_z": 8, "rho_x": 9, "rho_y": 10, "rho_z": 11,
    "u_x": 12, "u_y": 13, "u_z": 14, "v_x": 15, "v_y": 16, "v_z": 17,
    "w_x": 18, "w_y": 19, "w_z": 20, "fei_lap": 21, "u_lap": 22, "v_lap": 23,
    "w_lap": 24, "p_prev": 25, "f": 100, "g": 101, "h": 102,
}
STAGES = {
    "gradients": 0, "collision": 1, "stream_fg": 2, "bc_fg": 3, "macro": 4,
    "poisson": 5, "correct": 6, "update_hh": 7, "collision_p": 8,
    "stream_h": 9, "bc_h": 10, "getp": 11,
}


@dataclass
class Params:
    """The 14 positional lines of params.in (CPU/Inamuro.cpp:37-53)."""
    lx: int = 48
    ly: int = 96
    lz: int = 128
    rho_L: float = 50.0
    rho_G: float = 1.0
    mu_L: float = 0.008
    mu_G: float = 0.00016
    tauf: float = 1.0
    taug: float = 1.0
    k_f: float = 0.5
    k_g: float = 0.0005
    T: float = 0.035
    a: float = 1.0
    b: float = 6.7
    fei_max: float = 0.09714
    fei_min: float = 0.01134
    fei_L: float = 0.092
    fei_G: float = 0.015
    DD: float = 32.0
    period: int = 1
    max_time_steps: int = 20
    output_frequency: int = 20

    def params_in(self, extra: str = "") -> str:
        """Text of a params.in file in the reference's positional format."""
        r = repr
        lines = [
            f"{self.lx} {self.ly} {self.lz}", f"{self.period}", f"{self.max_time_steps}",
            f"{self.output_frequency}", f"{r(self.rho_L)} {r(self.rho_G)}",
            f"{r(self.tauf)} {r(self.taug)}", f"{r(self.mu_L)} {r(self.mu_G)}",
            f"{r(self.k_f)} {r(self.k_g)}", r(self.T), r(self.a), r(self.b),
            f"{r(self.fei_max)} {r(self.fei_min)}", f"{r(self.fei_L)} {r(self.fei_G)}",
            r(self.DD),
        ]
        return "\n".join(lines) + "\n" + extra


class _CParams(ctypes.Structure):
    _fields_ = [("lx", ctypes.c_int), ("ly", ctypes.c_int), ("lz", ctypes.c_int)] + [
        (n, ctypes.c_double) for n in (
            "rho_L", "rho_G", "mu_L", "mu_G", "tauf", "taug", "k_f", "k_g", "T", "a", "b",
            "fei_max", "fei_min", "fei_L", "fei_G", "DD")
    ]


class PoissonDiag(ctypes.Structure):
    _fields_ = [
        ("pressure_solve_call_count", ctypes.c_int32), ("hit_max_iterations", ctypes.c_int32),
        ("prev_iteration_count", ctypes.c_uint64), ("iteration_count", ctypes.c_uint64),
        ("relative_error", ctypes.c_double), ("pressure_l1_delta", ctypes.c_double),
        ("pressure_l1_norm", ctypes.c_double), ("divergence_mean", ctypes.c_double),
        ("pressure_gauge_target", ctypes.c_double), ("pressure_gauge_max_shift", ctypes.c_double),
        ("fixed_point_h_l1", ctypes.c_double), ("fixed_point_h_scale", ctypes.c_double),
        ("fixed_point_h_relative", ctypes.c_double), ("fixed_point_p_l1", ctypes.c_double),
        ("fixed_point_p_scale", ctypes.c_double), ("fixed_point_p_relative", ctypes.c_double),
        ("fixed_point_relative", ctypes.c_double), ("fixed_point_evaluated", ctypes.c_int32),
        ("converged", ctypes.c_int32), ("source_mean", ctypes.c_double),
        ("source_abs_mean", ctypes.c_double), ("projected_source_mean", ctypes.c_double),
        ("fallback_count", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("block_low_frequency_fraction", ctypes.c_double), ("block_size", ctypes.c_int32),
        ("block_count", ctypes.c_int32), ("anderson_alpha", ctypes.c_double),
    ]


POISSON_VARIANTS = {"split": 0, "fused": 1, "onepass": 2, "scalar": 3}


class PoissonAccel(ctypes.Structure):
    """orc_poisson_accel: the reference GPU solver's Poisson accelerator
    settings (GPU/InamuroCUDA.hpp:60-76)."""
    _fields_ = [
        ("variant", ctypes.c_int32), ("fused_boundary", ctypes.c_int32),
        ("pressure_relax_scale", ctypes.c_double), ("fixed_point_relax", ctypes.c_double),
        ("scalar_source_scale", ctypes.c_double), ("source_aware_hh_init", ctypes.c_int32),
        ("anderson_m1", ctypes.c_int32), ("source_aware_hh_scale", ctypes.c_double),
        ("anderson_beta_max", ctypes.c_double), ("two_grid_strength", ctypes.c_double),
        ("two_grid", ctypes.c_int32), ("spatial_diagnostics", ctypes.c_int32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.POINTER(_CParams)]
        L.orc_destroy.argtypes = [vp]
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_last_error.argtypes = [vp]
        L.orc_set_threads.argtypes = [vp, ctypes.c_int]
        L.orc_set_poisson.argtypes = [vp, ctypes.c_uint64, ctypes.c_double, ctypes.c_uint64, ctypes.c_int]
        L.orc_init_two_droplets_sharp.argtypes = [vp]
        L.orc_place_droplet.argtypes = [vp] + [ctypes.c_double] * 6 + [ctypes.c_int, ctypes.c_double]
        L.orc_get.argtypes = [vp, ctypes.c_int, dp]
        L.orc_set.argtypes = [vp, ctypes.c_int, dp]
        L.orc_step.argtypes = [vp]
        L.orc_stage.argtypes = [vp, ctypes.c_int]
        L.orc_is_first_step.argtypes = [vp]
        L.orc_set_first_step.argtypes = [vp, ctypes.c_int]
        L.orc_poisson_diag_get.argtypes = [vp, ctypes.POINTER(PoissonDiag)]
        L.orc_eos.restype = ctypes.c_double
        L.orc_eos.argtypes = [ctypes.POINTER(_CParams), ctypes.c_double]
        L.orc_rho_of_phi.restype = ctypes.c_double
        L.orc_rho_of_phi.argtypes = [ctypes.POINTER(_CParams), ctypes.c_double]
        L.orc_poisson_iterations.argtypes = [vp, ctypes.c_int]
        L.orc_fixed_point_residual.argtypes = [vp, ctypes.c_double]
        u64p = ctypes.POINTER(ctypes.c_uint64)
        L.orc_set_schedule.argtypes = [vp, u64p, ctypes.c_int]
        L.orc_set_z_walls.argtypes = [vp, ctypes.c_int]
        L.orc_set_gravity.argtypes = [vp, ctypes.c_double]
        L.orc_objective.argtypes = [vp, dp, dp, ctypes.c_int, u64p, dp]
        L.orc_adjoint.argtypes = [vp, dp, dp, ctypes.c_int, u64p, dp, dp, dp]
        L.orc_set_objective.argtypes = [vp, ctypes.c_int]
        L.orc_set_poisson_accel.argtypes = [vp, ctypes.POINTER(PoissonAccel)]
        L.orc_default_poisson_accel.argtypes = [ctypes.POINTER(PoissonAccel)]
        L.orc_poisson_accel_iterations.argtypes = [vp, ctypes.c_int]
        _lib = L
    return _lib


def _cparams(p: Params) -> _CParams:
    c = _CParams()
    for f in _CParams._fields_:
        setattr(c, f[0], getattr(p, f[0]))
    return c


def eos(p: Params, fei: float) -> float:
    return lib().orc_eos(ctypes.byref(_cparams(p)), fei)


def rho_of_phi(p: Params, fei: float) -> float:
    return lib().orc_rho_of_phi(ctypes.byref(_cparams(p)), fei)


class OracleError(RuntimeError):
    pass


class Oracle:
    """One oracle simulation (restated reference Inamuro state)."""

    def __init__(self, params: Params, threads: int = 1):
        self.params = params
        self._cp = _cparams(params)
        self.h = lib().orc_create(ctypes.byref(self._cp))
        if not self.h:
            raise ValueError("oracle rejected the grid size")
        self.shape = (params.lz, params.ly, params.lx)
        self.n = params.lx * params.ly * params.lz
        lib().orc_set_threads(self.h, threads)

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise OracleError(lib().orc_last_error(self.h).decode())

    def set_poisson(self, check_interval=100, tol=1e-3, iter_cap=0, dual=False):
        self._check(lib().orc_set_poisson(self.h, check_interval, tol, iter_cap, int(dual)))

    def set_poisson_accel(self, variant="split", **kw) -> None:
        """The reference GPU solver's Poisson accelerators: variant split,
        fused, onepass or scalar, plus any PoissonAccel field by name
        (fused_boundary, pressure_relax_scale, fixed_point_relax,
        scalar_source_scale, source_aware_hh_init, source_aware_hh_scale,
        anderson_m1, anderson_beta_max, two_grid, two_grid_strength,
        spatial_diagnostics); unnamed fields take the reference defaults."""
        a = PoissonAccel()
        lib().orc_default_poisson_accel(ctypes.byref(a))
        a.variant = POISSON_VARIANTS[variant]
        for k, v in kw.items():
            if k not in dict(PoissonAccel._fields_):
                raise KeyError(k)
            setattr(a, k, v)
        self._check(lib().orc_set_poisson_accel(self.h, ctypes.byref(a)))

    def poisson_accel_iterations(self, k: int) -> None:
        """k bare iterations of the configured accelerator variant (no
        projection, gauge or check; fresh Anderson history)."""
        self._check(lib().orc_poisson_accel_iterations(self.h, k))

    def init_two_droplets_sharp(self):
        lib().orc_init_two_droplets_sharp(self.h)

    def place_droplet(self, cx, cy, cz, vx=0.0, vy=0.0, vz=0.0, profile="tanh", width=2.0):
        lib().orc_place_droplet(self.h, cx, cy, cz, vx, vy, vz, 1 if profile == "tanh" else 0, width)

    def get(self, name: str) -> np.ndarray:
        aid = ARRAY_IDS[name]
        shape = (Q,) + self.shape if aid >= 100 else self.shape
        out = np.empty(shape, dtype=np.float64)
        self._check(lib().orc_get(self.h, aid, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def set(self, name: str, arr) -> None:
        aid = ARRAY_IDS[name]
        shape = (Q,) + self.shape if aid >= 100 else self.shape
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(arr, dtype=np.float64), shape))
        self._check(lib().orc_set(self.h, aid, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def fields(self) -> dict:
        return {k: self.get(k) for k in FIELD_NAMES}

    def step(self, n: int = 1) -> list:
        iters = []
        for _ in range(n):
            self._check(lib().orc_step(self.h))
            iters.append(self.diag().iteration_count)
        return iters

    def stage(self, name: str) -> None:
        self._check(lib().orc_stage(self.h, STAGES[name]))

    @property
    def first_step(self) -> bool:
        return bool(lib().orc_is_first_step(self.h))

    @first_step.setter
    def first_step(self, v: bool) -> None:
        lib().orc_set_first_step(self.h, int(v))

    def diag(self) -> PoissonDiag:
        d = PoissonDiag()
        lib().orc_poisson_diag_get(self.h, ctypes.byref(d))
        return d

    def fixed_point_residual(self, gauge_target: float) -> PoissonDiag:
        """The dual-mode check's kinetic fixed-point residual of the current
        state against gauge_target (fills the fixed_point_* diagnostics)."""
        self._check(lib().orc_fixed_point_residual(self.h, float(gauge_target)))
        return self.diag()

    def poisson_iterations(self, k: int) -> None:
        lib().orc_poisson_iterations(self.h, k)

    def set_z_walls(self, enabled: bool = True) -> None:
        """No-slip walls at z = 0 and lz-1 (extension; off by default)."""
        self._check(lib().orc_set_z_walls(self.h, int(bool(enabled))))

    def set_gravity(self, gz: float) -> None:
        """Gravitational acceleration along z (extension; 0 by default)."""
        lib().orc_set_gravity(self.h, float(gz))

    def set_schedule(self, iters) -> None:
        """Freeze the iteration counts of the next len(iters) Poisson solves."""
        a = np.ascontiguousarray(np.asarray(iters, dtype=np.uint64))
        self._check(lib().orc_set_schedule(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(a)))

    OBJECTIVES = {"misfit": 0, "centroid_z": 1}

    def set_objective(self, kind: str) -> None:
        """'misfit' (sum (phi_T - target)^2, the default) or 'centroid_z'
        (sum z w / sum w, w = phi_T - fei_G: the droplet's centroid height;
        the target is then ignored)."""
        self._check(lib().orc_set_objective(self.h, self.OBJECTIVES[kind]))

    def _adj_args(self, phi0, target, iters):
        phi0 = np.ascontiguousarray(np.asarray(phi0, dtype=np.float64).reshape(self.shape))
        if target is None:
            target = np.zeros(self.shape)
        target = np.ascontiguousarray(np.asarray(target, dtype=np.float64).reshape(self.shape))
        it = np.ascontiguousarray(np.asarray(iters, dtype=np.uint64))
        dptr = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        return phi0, target, it, dptr

    def objective(self, phi0, target, iters) -> float:
        """J = sum (phi_T - target)^2 of the frozen-schedule map from phi0."""
        phi0, target, it, dptr = self._adj_args(phi0, target, iters)
        J = ctypes.c_double()
        self._check(lib().orc_objective(self.h, dptr(phi0), dptr(target), len(it),
                                        it.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(J)))
        return J.value

    def adjoint(self, phi0, target, iters):
        """(J, dJ/dphi0, dJ/dk_g) of the frozen-schedule map."""
        phi0, target, it, dptr = self._adj_args(phi0, target, iters)
        J, gk = ctypes.c_double(), ctypes.c_double()
        g = np.empty(self.shape, dtype=np.float64)
        self._check(lib().orc_adjoint(self.h, dptr(phi0), dptr(target), len(it),
                                      it.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                      ctypes.byref(J), dptr(g), ctypes.byref(gk)))
        return J.value, g, gk.value


# ---------------------------------------------------------------------------
# Reference probe I/O (oracle/_ref, built only where /root/reference exists)
# ---------------------------------------------------------------------------

def ref_available() -> bool:
    return os.path.exists(REF_PROBE) and os.path.exists(REF_PROBE_INIT)


def write_init(path: str, flds: dict) -> None:
    with open(path, "wb") as fh:
        for k in FIELD_NAMES:
            fh.write(np.ascontiguousarray(flds[k], dtype=np.float64).tobytes())


def read_probe_output(path: str):
    with open(path, "rb") as fh:
        raw = fh.read()
    magic, lx, ly, lz, steps, dists = struct.unpack_from("<6i", raw, 0)
    assert magic == 0x4C424D31
    off = 24
    iters = np.frombuffer(raw, dtype=np.uint64, count=steps, offset=off).copy()
    off += 8 * steps
    step_ms = np.frombuffer(raw, dtype=np.float64, count=steps, offset=off).copy()
    off += 8 * steps
    n = lx * ly * lz
    out = {}
    for k in FIELD_NAMES:
        out[k] = np.frombuffer(raw, dtype=np.float64, count=n, offset=off).reshape(lz, ly, lx).copy()
        off += 8 * n
    if dists:
        for k in ("f", "g", "h"):
            out[k] = np.frombuffer(raw, dtype=np.float64, count=Q * n, offset=off).reshape(Q, lz, ly, lx).copy()
            off += 8 * Q * n
    return iters, step_ms, out


def read_init_probe_output(path: str):
    with open(path, "rb") as fh:
        raw = fh.read()
    magic, lx, ly, lz = struct.unpack_from("<4i", raw, 0)
    assert magic == 0x4C424D32
    n = lx * ly * lz
    off = 16
    out = {}
    for k in FIELD_NAMES:
        out[k] = np.frombuffer(raw, dtype=np.float64, count=n, offset=off).reshape(lz, ly, lx).copy()
        off += 8 * n
    return out


def param_names():
    return [f.name for f in fields(Params)]
