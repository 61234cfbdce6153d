"""Python plumbing over the C-ABI (tests and bench drive the engine through it).

The engine itself is the HIP library; these classes only own handles and
copy results into numpy arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi


class Solver:
    """One scene on the device (fdtd_create ... fdtd_destroy)."""

    def __init__(self, scene_buffers, steps_per_graph=0, tile_config=-1, series_capacity=0, stream=None):
        self.lib = capi.load_library()
        self.sb = scene_buffers
        opt = capi.Options()
        self.lib.fdtd_options_default(C.byref(opt))
        opt.steps_per_graph = steps_per_graph
        opt.tile_config = tile_config
        opt.series_capacity = series_capacity
        if stream is not None:
            opt.stream = stream
        h = C.c_void_p()
        capi.check(self.lib.fdtd_create(scene_buffers.ref(), C.byref(opt), C.byref(h)), "fdtd_create")
        self.h = h

    @classmethod
    def from_xml(cls, path, steps_per_graph=0):
        self = cls.__new__(cls)
        self.lib = capi.load_library()
        self.sb = None
        opt = capi.Options()
        self.lib.fdtd_options_default(C.byref(opt))
        opt.steps_per_graph = steps_per_graph
        h = C.c_void_p()
        capi.check(self.lib.fdtd_create_from_xml(path.encode(), C.byref(opt), C.byref(h)), "fdtd_create_from_xml")
        self.h = h
        return self

    def info(self):
        inf = capi.Info()
        capi.check(self.lib.fdtd_get_info(self.h, C.byref(inf)), "fdtd_get_info")
        return {k: getattr(inf, k) for k, _ in capi.Info._fields_}

    def step(self, n):
        capi.check(self.lib.fdtd_step(self.h, int(n)), "fdtd_step")

    def synchronize(self):
        capi.check(self.lib.fdtd_synchronize(self.h), "fdtd_synchronize")

    def listen(self):
        L = self.info()["n_listeners"]
        out = np.zeros(L)
        capi.check(self.lib.fdtd_listen(self.h, out.ctypes.data_as(C.POINTER(C.c_double))), "fdtd_listen")
        return out

    def series(self, first=0, n=None):
        inf = self.info()
        L = inf["n_listeners"]
        if n is None:
            n = inf["step_index"] - first
        out = np.zeros((n, L))
        capi.check(self.lib.fdtd_read_listener_series(self.h, int(first), int(n),
                                                      out.ctypes.data_as(C.POINTER(C.c_double))),
                   "fdtd_read_listener_series")
        return out

    def pressure(self):
        inf = self.info()
        out = np.zeros(inf["n_cells"])
        capi.check(self.lib.fdtd_read_pressure(self.h, out.ctypes.data_as(C.POINTER(C.c_double))),
                   "fdtd_read_pressure")
        if self.sb is not None:
            return out.reshape(self.sb.divisions)
        return out

    def set_timing(self, on=True):
        capi.check(self.lib.fdtd_set_timing(self.h, 1 if on else 0), "fdtd_set_timing")

    def last_kernel_ms(self):
        a, b = C.c_double(), C.c_double()
        capi.check(self.lib.fdtd_last_kernel_ms(self.h, C.byref(a), C.byref(b)), "fdtd_last_kernel_ms")
        return a.value, b.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.fdtd_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
