"""fdtd_solve_noalloc (no device allocation, scratch in a static device
array) against the allocating path fdtd_create + fdtd_step +
fdtd_read_listener_series on the same scenes: listener series and final
pressure field identical bit for bit, on two calls in a row with the second
scene smaller than the first (the scratch array keeps the first call's
contents), plus the size query and the scratch limit."""
import ctypes as C

import numpy as np
import pytest

from acoustics_amd import capi, scenes
from acoustics_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the engine library links (the caller allocates its
    state buffer with it; torch's bundled runtime is a different instance)."""
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _noalloc(lib, sb, steps):
    hip = _hip()
    need = C.c_int64(0)
    capi.check(lib.fdtd_solve_noalloc(sb.ref(), steps, None, 0, None, None, C.byref(need)), "size query")
    assert need.value > 0
    buf = C.c_void_p()
    assert hip.hipMalloc(C.byref(buf), need.value) == 0
    try:
        assert buf.value % 256 == 0
        out = np.zeros((steps, sb.n_listeners), dtype=np.float64)
        st = lib.fdtd_solve_noalloc(sb.ref(), steps, buf, need.value, None, out.ctypes.data_as(C.POINTER(C.c_double)),
                                    None)
        capi.check(st, "fdtd_solve_noalloc")
        # the newest pressure slot (p^steps) in the device layout: two slots of
        # padded rows (row stride: n2 rounded up to four; slot: rounded up to 64 cells)
        n0, n1, n2 = sb.divisions
        stride = (n2 + 3) // 4 * 4
        npad = (n0 * n1 * stride + 63) // 64 * 64
        slot = (steps + 1) & 1
        raw = np.zeros(npad, dtype=np.float32)
        assert hip.hipMemcpy(raw.ctypes.data_as(C.c_void_p), C.c_void_p(buf.value + 4 * slot * npad), 4 * npad, 2) == 0
        field = raw[: n0 * n1 * stride].reshape(n0, n1, stride)[:, :, :n2].astype(np.float64)
    finally:
        hip.hipFree(buf)
    return out, field


def _allocating(sb, steps):
    with Solver(sb) as s:
        s.step(steps)
        return s.series(), s.pressure()


def test_noalloc_matches_allocating_path(hip_lib):
    cases = [
        # first call: a sphere with rasterized ghost cells, 48^3, 200 steps
        scenes.config2(seed=41, n=48, radius=0.03, shell=0.07, pml=8, n_listeners=16)[0],
        # second call, smaller: free field 32^3 with a pressure source
        scenes.config1(seed=42, n=32)[0],
    ]
    steps = [200, 150]
    for sb, n in zip(cases, steps):
        print(f"no-alloc case: divisions={sb.divisions} steps={n}", flush=True)
        got, field = _noalloc(hip_lib, sb, n)
        ref, pref = _allocating(sb, n)
        assert np.any(ref != 0)
        assert np.array_equal(got, ref)
        assert np.array_equal(field, pref)
