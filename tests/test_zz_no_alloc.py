"""lbm_mass_noalloc (the allocation-free variant of lbm_mass) through the C
ABI: on two calls in a row, the second with a smaller input, its sum of a
device array equals lbm_mass's sum of the same phi field bit for bit (the
same kernels in the same order).  Device buffers of the test come from the
HIP runtime the library itself links (hipMalloc in the caller, not in the
entry point)."""
import ctypes
import os

import numpy as np
import pytest

from lbm_amd import hip
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip_runtime():
    rt = ctypes.CDLL(os.path.join("/opt/rocm/lib", "libamdhip64.so"))
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


def device_sum(rt, lib, values):
    """Uploads `values`, sums them with lbm_mass_noalloc on the null stream,
    returns the host copy of the device result."""
    a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
    d_in, d_out = ctypes.c_void_p(), ctypes.c_void_p()
    assert rt.hipMalloc(ctypes.byref(d_in), a.nbytes) == 0
    assert rt.hipMalloc(ctypes.byref(d_out), 8) == 0
    try:
        assert rt.hipMemcpy(d_in, a.ctypes.data, a.nbytes, H2D) == 0
        rc = lib.lbm_mass_noalloc(d_in, a.size, d_out, None)
        assert rc == hip.LBM_OK, lib.lbm_last_error(None)
        out = np.zeros(1)
        assert rt.hipMemcpy(out.ctypes.data, d_out, 8, D2H) == 0
        assert rt.hipDeviceSynchronize() == 0
        return float(out[0])
    finally:
        rt.hipFree(d_in)
        rt.hipFree(d_out)


def phi_field(P, cx, cy, cz):
    o = po.Oracle(P)
    o.place_droplet(cx, cy, cz, profile="tanh", width=2.0)
    return o.fields()


def test_noalloc_sum_equals_lbm_mass_on_two_calls():
    lib = hip.load_library()
    rt = hip_runtime()
    cases = [(po.Params(lx=48, ly=40, lz=36, DD=20.0), (20.5, 19.0, 17.0)),   # 69,120 cells
             (po.Params(lx=10, ly=9, lz=7, DD=4.0), (4.0, 4.5, 3.0))]          # 630 cells
    for i, (P, c) in enumerate(cases):
        print(f"no-alloc call {i + 1}: {P.lx}x{P.ly}x{P.lz}", flush=True)
        f = phi_field(P, *c)
        got = device_sum(rt, lib, f["fei"])  # first call on the device resolves the scratch
        sim = hip.InamuroHIP(P)
        sim.upload_fields(f)
        want = sim.mass()
        sim.close()
        assert np.float64(got).view(np.uint64) == np.float64(want).view(np.uint64), (i, got, want)
        assert abs(got - float(np.sum(f["fei"]))) <= 1e-12 * abs(got)


def test_noalloc_rejects_bad_arguments():
    lib = hip.load_library()
    d = ctypes.c_void_p(16)  # never dereferenced: every call below is rejected first
    assert lib.lbm_mass_noalloc(None, 10, d, None) == hip.LBM_ERR_ARG
    assert b"null" in lib.lbm_last_error(None)
    assert lib.lbm_mass_noalloc(d, 0, d, None) == hip.LBM_ERR_ARG
    assert lib.lbm_mass_noalloc(d, 1 << 31, d, None) == hip.LBM_ERR_ARG
    assert b"2^31-1" in lib.lbm_last_error(None)
