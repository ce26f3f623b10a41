"""ctypes binding of lbm_amd/liblbm_hip.so (the C ABI in include/lbm_hip.h).

`InamuroHIP` mirrors the reference's GPU backend class InamuroCUDA
(GPU/InamuroCUDA.hpp:16-91): performTimeStepGPU, downloadFieldsToCPU,
downloadHH, getLastPoissonDiagnostics, setPoissonConvergence,
setPoissonIterationLimit, setUsePoissonDualResidual, synchronize.  Errors
raise the exception type the reference throws, with the reference's message.

Arrays cross the boundary in the canonical layout: macro fields
float64[lz, ly, lx], distributions float64[15, lz, ly, lx].

There is no CPU fallback: if the HIP library is missing or no GPU is
present, construction fails loudly.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbm_hip.so")

Q = 15
FP64 = 0
MIXED = 1

LBM_OK = 0
LBM_ERR_ARG = -1
LBM_ERR_HIP = -2
LBM_ERR_NUMERIC = -3
LBM_ERR_OVERFLOW = -4
LBM_ERR_STATE = -5
LBM_ERR_NOMEM = -6

# Every symbol include/lbm_hip.h declares.
EXPORTED = (
    "lbm_create", "lbm_destroy", "lbm_last_error", "lbm_upload", "lbm_upload_dists",
    "lbm_set_poisson", "lbm_step", "lbm_download", "lbm_download_dists",
    "lbm_poisson_diag", "lbm_synchronize", "lbm_set_timing", "lbm_perf_get",
    "lbm_perf_reset", "lbm_mass", "lbm_init_reference", "lbm_place_droplet",
    "lbm_reset_fields", "lbm_set_iteration_schedule", "lbm_iteration_history",
    "lbm_adjoint", "lbm_set_adjoint_memory", "lbm_adjoint_stats", "lbm_set_poisson_graph",
    "lbm_set_z_walls", "lbm_set_gravity", "lbm_set_adjoint_objective",
    "lbm_set_adjoint_state_precision", "lbm_set_progress", "lbm_set_check_observer",
    "lbm_mass_noalloc", "lbm_set_poisson_launch", "lbm_get_poisson_launch",
    "lbm_poisson_accel_defaults", "lbm_set_poisson_accel", "lbm_get_poisson_accel",
)

# lbm_progress_fn: (user, phase, step, iterations)
PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64)
PROGRESS_PHASES = {0: "forward step", 1: "reverse block", 2: "reverse step"}

# Objectives of lbm_set_adjoint_objective.
OBJECTIVES = {"misfit": 0, "centroid_z": 1}


class LbmParams(ctypes.Structure):
    """lbm_params_t: the 14 positional parameters of params.in."""
    _fields_ = [("lx", ctypes.c_int32), ("ly", ctypes.c_int32), ("lz", ctypes.c_int32)] + [
        (n, ctypes.c_double) for n in (
            "rho_L", "rho_G", "mu_L", "mu_G", "tauf", "taug", "k_f", "k_g", "T", "a", "b",
            "fei_max", "fei_min", "fei_L", "fei_G", "DD")
    ]


class PoissonStepDiagnostics(ctypes.Structure):
    """lbm_poisson_diag_t (GPU/InamuroCUDA.hpp:19-43)."""
    _fields_ = [
        ("iterations", ctypes.c_uint64), ("prev_iterations", ctypes.c_uint64),
        ("pressure_solve_call_count", ctypes.c_int32), ("hit_max_iterations", ctypes.c_int32),
        ("relative_error", ctypes.c_double), ("pressure_l1_delta", ctypes.c_double),
        ("pressure_l1_norm", ctypes.c_double), ("fixed_point_h_l1", ctypes.c_double),
        ("fixed_point_h_scale", ctypes.c_double), ("fixed_point_h_relative", ctypes.c_double),
        ("fixed_point_p_l1", ctypes.c_double), ("fixed_point_p_scale", ctypes.c_double),
        ("fixed_point_p_relative", ctypes.c_double), ("fixed_point_relative", ctypes.c_double),
        ("source_mean", ctypes.c_double), ("source_abs_mean", ctypes.c_double),
        ("projected_source_mean", ctypes.c_double), ("divergence_mean", ctypes.c_double),
        ("pressure_gauge_target", ctypes.c_double), ("pressure_gauge_max_shift", ctypes.c_double),
        ("fixed_point_evaluated", ctypes.c_int32), ("pressure_converged", ctypes.c_int32),
        ("fixed_point_converged", ctypes.c_int32), ("dual_residual_enabled", ctypes.c_int32),
        ("converged", ctypes.c_int32), ("fallback_count", ctypes.c_int32),
    ]


class PoissonCheck(ctypes.Structure):
    """lbm_poisson_check_t: one convergence check, the values of one row of
    the reference's Poisson diagnostics CSV (GPU/InamuroCUDA.cu:2634-2682)."""
    _fields_ = [
        ("step", ctypes.c_int64), ("iteration", ctypes.c_uint64),
        ("pressure_l1_delta", ctypes.c_double), ("pressure_l1_norm", ctypes.c_double),
        ("relative_error", ctypes.c_double), ("fixed_point_h_l1", ctypes.c_double),
        ("fixed_point_h_scale", ctypes.c_double), ("fixed_point_h_relative", ctypes.c_double),
        ("fixed_point_p_l1", ctypes.c_double), ("fixed_point_p_scale", ctypes.c_double),
        ("fixed_point_p_relative", ctypes.c_double), ("fixed_point_relative", ctypes.c_double),
        ("fixed_point_evaluated", ctypes.c_int32), ("pressure_converged", ctypes.c_int32),
        ("fixed_point_converged", ctypes.c_int32), ("dual_residual_enabled", ctypes.c_int32),
        ("converged", ctypes.c_int32), ("block_size", ctypes.c_int32),
        ("block_count", ctypes.c_int32), ("pad", ctypes.c_int32),
        ("block_low_frequency_fraction", ctypes.c_double), ("anderson_alpha", ctypes.c_double),
    ]


POISSON_VARIANTS = {"split": 0, "fused": 1, "onepass": 2, "scalar": 3}


class PoissonAccel(ctypes.Structure):
    """lbm_poisson_accel_t: the reference GPU solver's Poisson accelerator
    settings (GPU/InamuroCUDA.hpp:60-76, 88)."""
    _fields_ = [
        ("variant", ctypes.c_int32), ("fused_boundary", ctypes.c_int32),
        ("pressure_relax_scale", ctypes.c_double), ("fixed_point_relax", ctypes.c_double),
        ("scalar_source_scale", ctypes.c_double), ("source_aware_hh_init", ctypes.c_int32),
        ("anderson_m1", ctypes.c_int32), ("source_aware_hh_scale", ctypes.c_double),
        ("anderson_beta_max", ctypes.c_double), ("two_grid_strength", ctypes.c_double),
        ("two_grid", ctypes.c_int32), ("spatial_diagnostics", ctypes.c_int32),
    ]


# lbm_check_fn: (user, const lbm_poisson_check_t*)
CHECK_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(PoissonCheck))


class PerformanceMetrics(ctypes.Structure):
    """lbm_perf_t (GPU/InamuroCUDA.hpp:115-132)."""
    _fields_ = [
        ("steps", ctypes.c_int64), ("poisson_iterations", ctypes.c_uint64),
        ("poisson_checks", ctypes.c_uint64), ("collide_ms", ctypes.c_double),
        ("macro_ms", ctypes.c_double), ("poisson_ms", ctypes.c_double),
        ("check_ms", ctypes.c_double), ("correct_ms", ctypes.c_double),
        ("poisson_iter_launches", ctypes.c_uint64),
    ]


class AdjointStats(ctypes.Structure):
    """lbm_adjoint_stats_t: cost record of the last adjoint call."""
    _fields_ = [
        ("steps", ctypes.c_int32), ("checkpoint_interval", ctypes.c_int32),
        ("store_slots", ctypes.c_int64), ("forward_iterations", ctypes.c_uint64),
        ("recomputed_iterations", ctypes.c_uint64), ("adjoint_iterations", ctypes.c_uint64),
        ("forward_ms", ctypes.c_double), ("reverse_ms", ctypes.c_double),
        ("final_relative_change", ctypes.c_double), ("iterations_to_tolerance", ctypes.c_int64),
    ]


_lib = None


def load_library():
    """Loads liblbm_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH} (run python -c 'import __graft_entry__ as g; g.build()')")
    L = ctypes.CDLL(LIB_PATH)
    vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
    L.lbm_create.argtypes = [ctypes.POINTER(LbmParams), ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    L.lbm_destroy.argtypes = [vp]
    L.lbm_last_error.restype = ctypes.c_char_p
    L.lbm_last_error.argtypes = [vp]
    L.lbm_upload.argtypes = [vp] + [dp] * 6
    L.lbm_upload_dists.argtypes = [vp, dp, dp, dp, ctypes.c_int]
    L.lbm_set_poisson.argtypes = [vp, ctypes.c_int32, ctypes.c_double, ctypes.c_uint64, ctypes.c_int]
    L.lbm_step.argtypes = [vp, ctypes.c_int32]
    L.lbm_download.argtypes = [vp] + [dp] * 6
    L.lbm_download_dists.argtypes = [vp, dp, dp, dp]
    L.lbm_poisson_diag.argtypes = [vp, ctypes.POINTER(PoissonStepDiagnostics)]
    L.lbm_synchronize.argtypes = [vp]
    L.lbm_set_timing.argtypes = [vp, ctypes.c_int]
    L.lbm_perf_get.argtypes = [vp, ctypes.POINTER(PerformanceMetrics)]
    L.lbm_perf_reset.argtypes = [vp]
    L.lbm_mass.argtypes = [vp, dp]
    L.lbm_init_reference.argtypes = [vp]
    L.lbm_reset_fields.argtypes = [vp]
    L.lbm_place_droplet.argtypes = [vp] + [ctypes.c_double] * 6 + [ctypes.c_int, ctypes.c_double]
    L.lbm_set_iteration_schedule.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.c_int64]
    L.lbm_iteration_history.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.c_int64,
                                        ctypes.POINTER(ctypes.c_int64)]
    L.lbm_adjoint.argtypes = [vp, dp, dp, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64),
                              dp, dp, dp]
    L.lbm_set_adjoint_memory.argtypes = [vp, ctypes.c_int64]
    L.lbm_set_poisson_graph.argtypes = [vp, ctypes.c_int]
    L.lbm_set_z_walls.argtypes = [vp, ctypes.c_int]
    L.lbm_set_gravity.argtypes = [vp, ctypes.c_double]
    L.lbm_set_adjoint_objective.argtypes = [vp, ctypes.c_int32]
    L.lbm_set_adjoint_state_precision.argtypes = [vp, ctypes.c_int32]
    L.lbm_set_progress.argtypes = [vp, PROGRESS_FN, vp]
    L.lbm_adjoint_stats.argtypes = [vp, ctypes.POINTER(AdjointStats)]
    L.lbm_set_check_observer.argtypes = [vp, CHECK_FN, vp]
    L.lbm_mass_noalloc.argtypes = [vp, ctypes.c_int64, vp, vp]
    L.lbm_set_poisson_launch.argtypes = [vp, ctypes.c_int]
    L.lbm_get_poisson_launch.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.lbm_poisson_accel_defaults.restype = None
    L.lbm_poisson_accel_defaults.argtypes = [ctypes.POINTER(PoissonAccel)]
    L.lbm_set_poisson_accel.argtypes = [vp, ctypes.POINTER(PoissonAccel)]
    L.lbm_get_poisson_accel.argtypes = [vp, ctypes.POINTER(PoissonAccel)]
    _lib = L
    return L


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class SolverParams:
    """The physical parameters of params.in (CPU/Inamuro.hpp:39-68 defaults)."""
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

    def to_c(self) -> LbmParams:
        c = LbmParams()
        for name, _ in LbmParams._fields_:
            setattr(c, name, getattr(self, name))
        return c


def _raise(code: int, msg: str):
    if code == LBM_ERR_ARG:
        raise ValueError(msg)
    if code == LBM_ERR_OVERFLOW:
        raise OverflowError(msg)
    raise RuntimeError(msg)


class InamuroHIP:
    """GPU time step of the Inamuro two-phase D3Q15 solver (one context)."""

    def __init__(self, params, precision: int = FP64):
        L = load_library()
        self.params = params
        self.precision = precision
        self._cp = params.to_c() if hasattr(params, "to_c") else SolverParams(
            **{k: getattr(params, k) for k, _ in LbmParams._fields_}).to_c()
        h = ctypes.c_void_p()
        rc = L.lbm_create(ctypes.byref(self._cp), precision, ctypes.byref(h))
        if rc != LBM_OK:
            _raise(rc, L.lbm_last_error(None).decode())
        self._h = h
        self.shape = (self._cp.lz, self._cp.ly, self._cp.lx)
        self._objective = "misfit"

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            load_library().lbm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != LBM_OK:
            _raise(rc, load_library().lbm_last_error(self._h).decode())

    # -- data movement ------------------------------------------------------
    def _field(self, a):
        if a is None:
            return None
        return np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), self.shape))

    def initFromCPU(self, fei=None, rho=None, u=None, v=None, w=None, p=None):
        arrs = [self._field(x) for x in (fei, rho, u, v, w, p)]
        self._check(load_library().lbm_upload(self._h, *[_ptr(a) for a in arrs]))

    def upload_fields(self, fields: dict):
        self.initFromCPU(**{k: fields.get(k) for k in ("fei", "rho", "u", "v", "w", "p")})

    def upload_dists(self, f=None, g=None, h=None, first_step=False):
        shp = (Q,) + self.shape
        arrs = [None if a is None else np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(shp))
                for a in (f, g, h)]
        self._check(load_library().lbm_upload_dists(self._h, *[_ptr(a) for a in arrs], int(first_step)))

    def downloadFieldsToCPU(self) -> dict:
        out = {k: np.empty(self.shape) for k in ("fei", "rho", "u", "v", "w", "p")}
        self._check(load_library().lbm_download(self._h, *[_ptr(out[k]) for k in ("fei", "rho", "u", "v", "w", "p")]))
        return out

    def download_dists(self, which=("f", "g", "h")) -> dict:
        shp = (Q,) + self.shape
        out = {k: np.empty(shp) for k in which}
        self._check(load_library().lbm_download_dists(
            self._h, _ptr(out.get("f")), _ptr(out.get("g")), _ptr(out.get("h"))))
        return out

    def downloadHH(self) -> np.ndarray:
        return self.download_dists(("h",))["h"]

    def initializeDroplets(self):
        """CPU-tree initialisation: two sharp droplets (CPU/Inamuro.cpp:550-587)."""
        self._check(load_library().lbm_init_reference(self._h))

    def reset_fields(self):
        """Background state of Inamuro::initializeArrays (CPU/Inamuro.cpp:511-548)."""
        self._check(load_library().lbm_reset_fields(self._h))

    def place_droplet(self, cx, cy, cz, vx=0.0, vy=0.0, vz=0.0, profile="tanh", width=2.0):
        """GPU-tree place_droplet (GPU/Inamuro.cpp:499-533)."""
        prof = {"sharp": 0, "tanh": 1}[profile]
        self._check(load_library().lbm_place_droplet(self._h, cx, cy, cz, vx, vy, vz, prof, width))

    # -- configuration ------------------------------------------------------
    def setPoisson(self, check_interval=100, tolerance=1e-3, iteration_limit=0, dual_residual=False):
        """dual_residual: False/0 off, True/1 the dual exit, 2 the kinetic
        residual evaluated without changing the exit (the fixed-point shadow)."""
        dual = int(dual_residual)
        self._check(load_library().lbm_set_poisson(self._h, int(check_interval), float(tolerance),
                                                   int(iteration_limit), dual))
        self._poisson = (int(check_interval), float(tolerance), int(iteration_limit), dual)

    def _cfg(self):
        return getattr(self, "_poisson", (100, 1e-3, 0, False))

    def setPoissonConvergence(self, check_interval: int, tolerance: float):
        _, _, lim, dual = self._cfg()
        self.setPoisson(check_interval, tolerance, lim, dual)

    def setPoissonIterationLimit(self, max_iterations: int):
        ci, tol, _, dual = self._cfg()
        self.setPoisson(ci, tol, max_iterations, dual)

    def setUsePoissonDualResidual(self, enabled: bool):
        ci, tol, lim, dual = self._cfg()
        self.setPoisson(ci, tol, lim, 1 if enabled else (0 if dual == 1 else dual))

    def setEnablePoissonFixedPointShadow(self, enabled: bool):
        """GPU/InamuroCUDA.cu:2568-2571: the kinetic residual at every check,
        reported in the diagnostics, without changing the exit."""
        ci, tol, lim, dual = self._cfg()
        if dual != 1:
            self.setPoisson(ci, tol, lim, 2 if enabled else 0)

    def setUsePoissonDeterministicReductions(self, enabled: bool):
        """GPU/InamuroCUDA.cu:2563-2566.  Every reduction of this build is a
        fixed-shape tree and deterministic whatever the flag; accepted for the
        reference's interface."""

    def set_poisson_launch(self, config: int = -1):
        """Launch shape of the Poisson kernel: -1 chooses by timing at the
        first solve (default), 0..2 fixes one; results are identical."""
        self._check(load_library().lbm_set_poisson_launch(self._h, int(config)))

    def poisson_launch(self) -> int:
        out = ctypes.c_int()
        self._check(load_library().lbm_get_poisson_launch(self._h, ctypes.byref(out)))
        return out.value

    def set_z_walls(self, enabled: bool = True):
        """No-slip walls at z = 0 and lz-1 (extension; off by default)."""
        self._check(load_library().lbm_set_z_walls(self._h, int(bool(enabled))))

    def set_gravity(self, gz: float):
        """Gravitational acceleration along z (extension; 0 by default)."""
        self._check(load_library().lbm_set_gravity(self._h, float(gz)))

    # -- the reference's Poisson accelerators (FP64) --------------------------
    def poisson_accel(self) -> dict:
        a = PoissonAccel()
        self._check(load_library().lbm_get_poisson_accel(self._h, ctypes.byref(a)))
        return {k: getattr(a, k) for k, _ in PoissonAccel._fields_}

    def set_poisson_accel(self, variant=None, **kw):
        """Changes the named accelerator settings (lbm_poisson_accel_t fields;
        variant 'split', 'fused', 'onepass' or 'scalar'), keeping the others."""
        a = PoissonAccel()
        self._check(load_library().lbm_get_poisson_accel(self._h, ctypes.byref(a)))
        if variant is not None:
            a.variant = POISSON_VARIANTS[variant] if isinstance(variant, str) else int(variant)
        names = dict(PoissonAccel._fields_)
        for k, v in kw.items():
            if k not in names:
                raise KeyError(k)
            setattr(a, k, v)
        self._check(load_library().lbm_set_poisson_accel(self._h, ctypes.byref(a)))

    def _variant_flag(self, name: str, enabled: bool):
        # setUseFusedPoisson / setUseOnePassPoisson / setUseScalarPoisson keep
        # three flags; doPressurePoisson gives scalar, then onepass, then fused
        # precedence (GPU/InamuroCUDA.cu:2216-2296).
        flags = getattr(self, "_variant_flags", {"fused": False, "onepass": False, "scalar": False})
        flags = dict(flags, **{name: bool(enabled)})
        variant = next((v for v in ("scalar", "onepass", "fused") if flags[v]), "split")
        self.set_poisson_accel(variant=variant)
        self._variant_flags = flags

    def setUseFusedPoisson(self, enabled: bool):
        self._variant_flag("fused", enabled)

    def setUseOnePassPoisson(self, enabled: bool):
        self._variant_flag("onepass", enabled)

    def setUseScalarPoisson(self, enabled: bool):
        self._variant_flag("scalar", enabled)

    def setScalarPoissonSourceScale(self, scale: float):
        self.set_poisson_accel(scalar_source_scale=float(scale))

    def setUseSourceAwareHHInit(self, enabled: bool):
        self.set_poisson_accel(source_aware_hh_init=int(bool(enabled)))

    def setSourceAwareHHScale(self, scale: float):
        self.set_poisson_accel(source_aware_hh_scale=float(scale))

    def setPressureRelaxScale(self, scale: float):
        self.set_poisson_accel(pressure_relax_scale=float(scale))

    def setPoissonFixedPointRelax(self, omega: float):
        self.set_poisson_accel(fixed_point_relax=float(omega))

    def setUsePoissonAndersonM1(self, enabled: bool):
        self.set_poisson_accel(anderson_m1=int(bool(enabled)))

    def setPoissonAndersonBetaMax(self, value: float):
        self.set_poisson_accel(anderson_beta_max=float(value))

    def setUsePoissonTwoGridCorrection(self, enabled: bool):
        self.set_poisson_accel(two_grid=int(bool(enabled)))

    def setPoissonTwoGridStrength(self, value: float):
        self.set_poisson_accel(two_grid_strength=float(value))

    def setUseFusedBoundaryPressure(self, enabled: bool):
        self.set_poisson_accel(fused_boundary=int(bool(enabled)))

    def setUsePoissonSpatialDiagnostics(self, enabled: bool):
        self.set_poisson_accel(spatial_diagnostics=int(bool(enabled)))

    def setUsePoissonGraph(self, enabled: bool):
        """Replay each check-interval segment as a captured HIP graph (default on)."""
        self._check(load_library().lbm_set_poisson_graph(self._h, int(bool(enabled))))

    def set_iteration_schedule(self, iterations):
        """Freeze the Poisson iteration count of the next len(iterations) steps."""
        arr = np.ascontiguousarray(np.asarray(list(iterations), dtype=np.uint64))
        self._check(load_library().lbm_set_iteration_schedule(
            self._h, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(arr)))

    def iteration_history(self) -> list:
        n = ctypes.c_int64()
        self._check(load_library().lbm_iteration_history(self._h, None, 0, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.uint64)
        self._check(load_library().lbm_iteration_history(
            self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), n.value, ctypes.byref(n)))
        return [int(x) for x in out[:n.value]]

    # -- stepping -----------------------------------------------------------
    def performTimeStepGPU(self, n: int = 1):
        self._check(load_library().lbm_step(self._h, int(n)))

    def step(self, n: int = 1) -> list:
        """Runs n steps; returns the Poisson iteration count of each."""
        its = []
        for _ in range(n):
            self.performTimeStepGPU(1)
            its.append(self.getLastPoissonDiagnostics().iterations)
        return its

    def synchronize(self):
        self._check(load_library().lbm_synchronize(self._h))

    def getLastPoissonDiagnostics(self) -> PoissonStepDiagnostics:
        d = PoissonStepDiagnostics()
        self._check(load_library().lbm_poisson_diag(self._h, ctypes.byref(d)))
        return d

    def set_check_observer(self, fn):
        """fn(record) after every completed Poisson convergence check, with
        a dict of the lbm_poisson_check_t fields; None switches it off."""
        if fn is None:
            self._check_cb = None
            self._check(load_library().lbm_set_check_observer(self._h, CHECK_FN(), None))
            return

        def cb(_u, rec):
            r = rec.contents
            fn({k: getattr(r, k) for k, _ in PoissonCheck._fields_ if k != "pad"})

        self._check_cb = CHECK_FN(cb)
        self._check(load_library().lbm_set_check_observer(self._h, self._check_cb, None))

    def getGridSize(self):
        return (self._cp.lx, self._cp.ly, self._cp.lz)

    def getAlgorithmName(self) -> str:
        return "Inamuro"

    # -- timing ---------------------------------------------------------------
    def set_timing(self, enabled: bool):
        self._check(load_library().lbm_set_timing(self._h, int(bool(enabled))))

    def perf(self) -> PerformanceMetrics:
        p = PerformanceMetrics()
        self._check(load_library().lbm_perf_get(self._h, ctypes.byref(p)))
        return p

    def reset_perf(self):
        self._check(load_library().lbm_perf_reset(self._h))

    def setEnablePoissonDetailTiming(self, enabled: bool):
        """GPU/InamuroCUDA.cu:2558-2561: one fused kernel per iteration here,
        so the detail is the iteration and check totals of perf(); turns
        event timing on."""
        if enabled:
            self.set_timing(True)

    def resetPerformanceMetrics(self):
        self.reset_perf()

    # -- discrete adjoint (no counterpart in the reference) ----------------------
    def initial_rho(self, phi0):
        """rho0 of the adjoint's initial state: linear in phi0 (place_droplet's rule)."""
        c = self._cp
        return c.rho_G + (np.asarray(phi0, dtype=np.float64) - c.fei_min) * (c.rho_L - c.rho_G) / (c.fei_max - c.fei_min)

    def set_adjoint_objective(self, kind: str):
        """'misfit' (J = sum (phi_T - target)^2, the default) or 'centroid_z'
        (J = sum z w / sum w, w = phi_T - fei_G: the droplet's centroid
        height; no target)."""
        self._check(load_library().lbm_set_adjoint_objective(self._h, OBJECTIVES[kind]))
        self._objective = kind

    def set_progress(self, fn):
        """fn(phase, step, iterations) with phase 'forward step' / 'reverse
        block' / 'reverse step', called by lbm_adjoint as its work finishes;
        None switches reports off."""
        if fn is None:
            self._progress = None
            self._check(load_library().lbm_set_progress(self._h, PROGRESS_FN(), None))
            return
        self._progress = PROGRESS_FN(lambda _u, ph, st, it: fn(PROGRESS_PHASES.get(ph, str(ph)), st, it))
        self._check(load_library().lbm_set_progress(self._h, self._progress, None))

    def set_adjoint_state_precision(self, precision: str):
        """'fp64' (default) or 'fp32': storage of the adjoint Poisson state
        between iterations (arithmetic and reductions stay FP64)."""
        self._check(load_library().lbm_set_adjoint_state_precision(self._h, {"fp64": 0, "fp32": 1}[precision]))

    def adjoint(self, phi0, target, steps: int, checkpoint_interval: int = 100, iterations=None):
        """(J, dJ/dphi0, dJ/dk_g) of the objective (set_adjoint_objective)
        after `steps` steps from phi0; `iterations` freezes the Poisson counts,
        otherwise the forward pass's own counts are frozen for the reverse
        sweep.  target may be None for the centroid objective."""
        phi0 = self._field(phi0)
        target = None if target is None else self._field(target)
        grad = np.empty(self.shape)
        J, gk = ctypes.c_double(), ctypes.c_double()
        it_ptr = None
        if iterations is not None:
            its = np.ascontiguousarray(np.asarray(list(iterations), dtype=np.uint64))
            if len(its) != steps:
                raise ValueError("one iteration count per step")
            it_ptr = its.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
        self._check(load_library().lbm_adjoint(self._h, _ptr(phi0), None if target is None else _ptr(target),
                                               int(steps),
                                               int(checkpoint_interval), it_ptr, ctypes.byref(J),
                                               _ptr(grad), ctypes.byref(gk)))
        return J.value, grad, gk.value

    def objective(self, phi0, target, iterations) -> float:
        """Forward pass of the adjoint's map with frozen iteration counts."""
        phi0 = self._field(phi0)
        self.initFromCPU(fei=phi0, rho=self.initial_rho(phi0), u=0.0, v=0.0, w=0.0, p=0.0)
        self.set_iteration_schedule(iterations)
        self.performTimeStepGPU(len(iterations))
        fei = self.downloadFieldsToCPU()["fei"]
        if self._objective == "centroid_z":
            w = fei - self._cp.fei_G
            return float((np.arange(self.shape[0])[:, None, None] * w).sum() / w.sum())
        return float(((fei - self._field(target)) ** 2).sum())

    def set_adjoint_memory(self, iterate_bytes: int):
        self._check(load_library().lbm_set_adjoint_memory(self._h, int(iterate_bytes)))

    def adjoint_stats(self) -> AdjointStats:
        st = AdjointStats()
        self._check(load_library().lbm_adjoint_stats(self._h, ctypes.byref(st)))
        return st

    def mass(self) -> float:
        out = ctypes.c_double()
        self._check(load_library().lbm_mass(self._h, ctypes.byref(out)))
        return out.value
