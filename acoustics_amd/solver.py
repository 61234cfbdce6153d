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

    def __init__(self, scene_buffers, steps_per_graph=0, tile_config=-1, series_capacity=0, stream=None,
                 pml_state=capi.FDTD_PML_STATE_MERGED):
        self.lib = capi.load_library()
        self.sb = scene_buffers
        opt = capi.Options()
        self.lib.fdtd_options_default(C.byref(opt))
        opt.steps_per_graph = steps_per_graph
        opt.tile_config = tile_config
        opt.series_capacity = series_capacity
        opt.pml_state = pml_state
        if stream is not None:
            opt.stream = stream
        h = C.c_void_p()
        capi.check(self.lib.fdtd_create(scene_buffers.ref(), C.byref(opt), C.byref(h)), "fdtd_create")
        self.h = h

    @classmethod
    def from_xml(cls, path, steps_per_graph=0, pml_state=capi.FDTD_PML_STATE_MERGED):
        self = cls.__new__(cls)
        self.lib = capi.load_library()
        self.sb = None
        opt = capi.Options()
        self.lib.fdtd_options_default(C.byref(opt))
        opt.steps_per_graph = steps_per_graph
        opt.pml_state = pml_state
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
        if getattr(self, "_n_listeners", None) is None:
            self._n_listeners = self.info()["n_listeners"]
        out = np.zeros(self._n_listeners)
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

    def save_snapshot(self, pattern, simulator_id=None, numbering=0):
        capi.check(self.lib.fdtd_save_snapshot(self.h, pattern.encode(), simulator_id.encode() if simulator_id else None,
                                               int(numbering)), "fdtd_save_snapshot")

    def load_snapshot(self, pattern, simulator_id=None, numbering=0):
        capi.check(self.lib.fdtd_load_snapshot(self.h, pattern.encode(), simulator_id.encode() if simulator_id else None,
                                               int(numbering)), "fdtd_load_snapshot")

    def set_timing(self, on=True):
        """on: True / FDTD_TIMING_PER_STEP (events around each step's kernels),
        FDTD_TIMING_SPAN (events around each fdtd_step call only), or False."""
        mode = capi.FDTD_TIMING_PER_STEP if on is True else (capi.FDTD_TIMING_OFF if on is False else int(on))
        capi.check(self.lib.fdtd_set_timing(self.h, mode), "fdtd_set_timing")

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


def _scene_array(scene_buffers):
    arr = (capi.Scene * len(scene_buffers))()
    for i, sb in enumerate(scene_buffers):
        arr[i] = sb.scene
    return arr


class Batch:
    """Independent scenes advanced together (fdtd_batch_*)."""

    def __init__(self, scene_buffers, num_steps, steps_per_graph=32, tile_config=-1):
        self.lib = capi.load_library()
        self.sbs = list(scene_buffers)
        self.steps = [int(s) for s in num_steps]
        self._arr = _scene_array(self.sbs)
        ns = (C.c_int * len(self.steps))(*self.steps)
        opt = capi.BatchOptions()
        self.lib.fdtd_batch_options_default(C.byref(opt))
        opt.steps_per_graph = steps_per_graph
        opt.tile_config = tile_config
        h = C.c_void_p()
        capi.check(self.lib.fdtd_batch_create(self._arr, ns, len(self.sbs), C.byref(opt), C.byref(h)),
                   "fdtd_batch_create")
        self.h = h

    def run(self):
        capi.check(self.lib.fdtd_batch_run(self.h), "fdtd_batch_run")

    def step(self, n):
        capi.check(self.lib.fdtd_batch_step(self.h, int(n)), "fdtd_batch_step")

    def series(self, i):
        done = C.c_int64()
        capi.check(self.lib.fdtd_batch_steps_done(self.h, i, C.byref(done)), "fdtd_batch_steps_done")
        out = np.zeros((done.value, self.sbs[i].n_listeners))
        capi.check(self.lib.fdtd_batch_read_listener_series(self.h, i, out.ctypes.data_as(C.POINTER(C.c_double))),
                   "fdtd_batch_read_listener_series")
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.fdtd_batch_destroy(self.h)
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


def time_chunk(settings, num_steps, chunk):
    """Settings of one time-parallel chunk (fdtd_time_chunk): returns
    (settings dict, first step in the whole run, steps the chunk runs)."""
    lib = capi.load_library()
    src = capi.SceneBuffers(settings).scene.settings
    out = capi.Settings()
    first, n = C.c_int64(), C.c_int64()
    capi.check(lib.fdtd_time_chunk(C.byref(src), int(num_steps), int(chunk), C.byref(out), C.byref(first),
                                   C.byref(n)), "fdtd_time_chunk")
    d = dict(settings)
    d["start_time"] = out.start_time
    d["stop_boundary_acc_time"] = out.stop_boundary_acc_time
    return d, first.value, n.value


def time_parallel_run(scene_buffers, num_steps, steps_per_graph=32, tile_config=-1):
    """Runs every chunk of a time-parallel scene as one device batch and
    sums the chunks' listener series (fdtd_run_time_parallel).  Returns
    (series [num_steps][L], steps each chunk ran)."""
    lib = capi.load_library()
    sb = scene_buffers
    k = sb.scene.settings.num_time_chunks
    opt = capi.BatchOptions()
    lib.fdtd_batch_options_default(C.byref(opt))
    opt.steps_per_graph = steps_per_graph
    opt.tile_config = tile_config
    out = np.zeros((int(num_steps), sb.n_listeners))
    run = (C.c_int64 * max(1, k))()
    capi.check(lib.fdtd_run_time_parallel(C.byref(sb.scene), int(num_steps), C.byref(opt),
                                          out.ctypes.data_as(C.POINTER(C.c_double)), run),
               "fdtd_run_time_parallel")
    return out, [run[i] for i in range(k)]


class Server:
    """Continuous-batching request loop (fdtd_server_*)."""

    def __init__(self, hbm_budget_bytes=64 << 30, max_active=0, steps_per_block=64):
        self.lib = capi.load_library()
        opt = capi.ServerOptions()
        self.lib.fdtd_server_options_default(C.byref(opt))
        opt.hbm_budget_bytes = int(hbm_budget_bytes)
        opt.max_active = int(max_active)
        opt.steps_per_block = int(steps_per_block)
        h = C.c_void_p()
        capi.check(self.lib.fdtd_server_create(C.byref(opt), C.byref(h)), "fdtd_server_create")
        self.h = h
        self.meta = {}

    def submit(self, scene_buffers, num_steps):
        rid = C.c_int64()
        capi.check(self.lib.fdtd_server_submit(self.h, scene_buffers.ref(), int(num_steps), C.byref(rid)),
                   "fdtd_server_submit")
        self.meta[rid.value] = (int(num_steps), scene_buffers.n_listeners)
        return rid.value

    def poll(self, rid):
        return self.lib.fdtd_server_poll(self.h, rid)

    def wait(self, rid, timeout_ms=-1):
        return self.lib.fdtd_server_wait(self.h, rid, int(timeout_ms))

    def cancel(self, rid):
        capi.check(self.lib.fdtd_server_cancel(self.h, rid), "fdtd_server_cancel")

    def take(self, rid):
        """Returns (status, series or None, elapsed_ms)."""
        steps, L = self.meta[rid]
        out = np.zeros((steps, L))
        el = C.c_double()
        st = self.lib.fdtd_server_take_result(self.h, rid, out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(el))
        return st, (out if st == 0 else None), el.value

    def stats(self):
        s = capi.ServerStats()
        capi.check(self.lib.fdtd_server_get_stats(self.h, C.byref(s)), "fdtd_server_get_stats")
        return {k: getattr(s, k) for k, _ in capi.ServerStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            self.lib.fdtd_server_destroy(self.h)
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


def scene_device_bytes(scene_buffers):
    lib = capi.load_library()
    return lib.fdtd_scene_device_bytes(scene_buffers.ref())
