"""ctypes mirror of include/fdtd_acoustics.h and the loader of the HIP library.

The product path is the C-ABI library ``acoustics_amd/lib/libfdtd_acoustics.so``
built from ``acoustics_amd/csrc``.  Loading fails loudly when it is missing:
there is no CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libfdtd_acoustics.so")

FDTD_BOUNDARY_RASTERIZE = 0
FDTD_BOUNDARY_FULLY_COUPLED = 1
FDTD_SCHEME_COLLOCATED = 0
FDTD_SCHEME_STAGGERED = 1
FDTD_SAMPLE_CENTER = 0
FDTD_SAMPLE_SEVEN = 1
FDTD_PML_STATE_MERGED = 0
FDTD_PML_STATE_SPLIT = 1
FDTD_TIMING_OFF = 0
FDTD_TIMING_PER_STEP = 1
FDTD_TIMING_SPAN = 2
FDTD_OBJ_SPHERE = 0
FDTD_OBJ_ELLIPSOID = 1

STATUS_NAMES = {
    0: "FDTD_OK", 1: "FDTD_ERR_INVALID", 2: "FDTD_ERR_NO_DEVICE", 3: "FDTD_ERR_DEVICE",
    4: "FDTD_ERR_OOM", 5: "FDTD_ERR_PARSE", 6: "FDTD_ERR_UNSUPPORTED", 7: "FDTD_ERR_NONFINITE",
    8: "FDTD_ERR_BUSY", 9: "FDTD_ERR_CANCELLED", 10: "FDTD_ERR_NOT_FOUND",
}


class Settings(C.Structure):
    _fields_ = [
        ("sound_speed", C.c_double), ("density", C.c_double), ("cell_size", C.c_double),
        ("time_step", C.c_double), ("stop_time", C.c_double), ("pml_width", C.c_double),
        ("pml_strength", C.c_double), ("air_viscosity_alpha", C.c_double), ("start_time", C.c_double),
        ("boundary_condition_preset", C.c_int), ("boundary_handling", C.c_int), ("use_mesh", C.c_int),
        ("divisions", C.c_int * 3), ("min_bound", C.c_double * 3),
        ("scheme", C.c_int), ("time_parallel", C.c_int), ("num_time_chunks", C.c_int), ("sampling", C.c_int),
        ("overlap_time", C.c_double), ("early_termination_ratio", C.c_double),
        ("stop_boundary_acc_time", C.c_double),
    ]


class Object(C.Structure):
    _fields_ = [
        ("type", C.c_int), ("reserved", C.c_int), ("center", C.c_double * 3),
        ("semi_axes", C.c_double * 3), ("rotation", C.c_double * 9),
        ("velocity", C.c_double * 3), ("angular_velocity", C.c_double * 3),
    ]


FDTD_VIB_GAUSSIAN, FDTD_VIB_HARMONIC, FDTD_VIB_CHIRP = 0, 1, 2


class ImpulseSource(C.Structure):
    _fields_ = [("object", C.c_int), ("kind", C.c_int), ("amplitude", C.c_double),
                ("width", C.c_double), ("onset", C.c_double), ("omega", C.c_double), ("phase", C.c_double),
                ("decay_rate", C.c_double), ("decay_center", C.c_double), ("t0", C.c_double), ("t1", C.c_double),
                ("omega1", C.c_double)]


class Impact(C.Structure):
    _fields_ = [("timestamp", C.c_double), ("support", C.c_double), ("impulse", C.c_double * 3),
                ("position", C.c_double * 3)]


class AccNoiseSource(C.Structure):
    _fields_ = [("object", C.c_int), ("n_impacts", C.c_int), ("mass", C.c_double),
                ("center_of_mass", C.c_double * 3), ("inv_inertia", C.c_double * 9),
                ("impacts", C.POINTER(Impact))]


class PressureSource(C.Structure):
    _fields_ = [("position", C.c_double * 3), ("width_time", C.c_double), ("width_space", C.c_double),
                ("offset_time", C.c_double), ("normalize_constant", C.c_double)]


class SpeakerSource(C.Structure):
    _fields_ = [("position", C.c_double * 3), ("width_space", C.c_double), ("start_time", C.c_double),
                ("sample_rate", C.c_double), ("n_samples", C.c_int), ("reserved", C.c_int),
                ("samples", C.POINTER(C.c_double))]


class PlaneConstraint(C.Structure):
    _fields_ = [("direction", C.c_int), ("sign", C.c_int), ("height", C.c_double)]


class Kinematics(C.Structure):
    _fields_ = [("object", C.c_int), ("n_frames", C.c_int), ("step_size", C.c_double),
                ("frames", C.POINTER(C.c_double))]


class Scene(C.Structure):
    _fields_ = [
        ("settings", Settings),
        ("n_objects", C.c_int), ("objects", C.POINTER(Object)),
        ("n_impulse_sources", C.c_int), ("impulse_sources", C.POINTER(ImpulseSource)),
        ("n_pressure_sources", C.c_int), ("pressure_sources", C.POINTER(PressureSource)),
        ("n_listeners", C.c_int), ("listeners", C.POINTER(C.c_double)),
        ("n_speaker_sources", C.c_int), ("speaker_sources", C.POINTER(SpeakerSource)),
        ("n_constraints", C.c_int), ("constraints", C.POINTER(PlaneConstraint)),
        ("n_acc_noise_sources", C.c_int), ("acc_noise_sources", C.POINTER(AccNoiseSource)),
        ("n_kinematics", C.c_int), ("kinematics", C.POINTER(Kinematics)),
    ]


class Options(C.Structure):
    _fields_ = [("stream", C.c_void_p), ("device", C.c_int), ("steps_per_graph", C.c_int),
                ("series_capacity", C.c_int), ("tile_config", C.c_int), ("pml_state", C.c_int)]


class Info(C.Structure):
    _fields_ = [("n_cells", C.c_int64), ("n_pml_cells", C.c_int64), ("n_ghost_cells", C.c_int64),
                ("n_solid_cells", C.c_int64), ("n_listeners", C.c_int64), ("step_index", C.c_int64),
                ("time", C.c_double), ("bytes_per_step", C.c_double), ("device_bytes", C.c_int64),
                ("n_pml_merged_cells", C.c_int64), ("state_bytes_per_step", C.c_double)]


class BatchOptions(C.Structure):
    _fields_ = [("stream", C.c_void_p), ("steps_per_graph", C.c_int), ("tile_config", C.c_int)]


class ServerOptions(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_active", C.c_int), ("steps_per_block", C.c_int),
                ("device", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("cancelled", C.c_int64), ("active", C.c_int64), ("queued", C.c_int64),
                ("blocks_run", C.c_int64), ("peak_resident_bytes", C.c_int64)]


# Every symbol include/fdtd_acoustics.h declares, with (restype, argtypes).
_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_D = C.POINTER(C.c_double)
SIGNATURES = {
    "fdtd_status_string": (C.c_char_p, [C.c_int]),
    "fdtd_settings_default": (None, [C.POINTER(Settings)]),
    "fdtd_parse_xml": (C.c_int, [C.c_char_p, _PP]),
    "fdtd_parsed_scene": (C.POINTER(Scene), [_P]),
    "fdtd_parsed_num_steps": (C.c_int, [_P]),
    "fdtd_parsed_free": (None, [_P]),
    "fdtd_last_error": (C.c_char_p, []),
    "fdtd_options_default": (None, [C.POINTER(Options)]),
    "fdtd_create": (C.c_int, [C.POINTER(Scene), C.POINTER(Options), _PP]),
    "fdtd_create_from_xml": (C.c_int, [C.c_char_p, C.POINTER(Options), _PP]),
    "fdtd_destroy": (C.c_int, [_P]),
    "fdtd_step": (C.c_int, [_P, C.c_int]),
    "fdtd_synchronize": (C.c_int, [_P]),
    "fdtd_listen": (C.c_int, [_P, _D]),
    "fdtd_read_listener_series": (C.c_int, [_P, C.c_int64, C.c_int64, _D]),
    "fdtd_read_pressure": (C.c_int, [_P, _D]),
    "fdtd_get_info": (C.c_int, [_P, C.POINTER(Info)]),
    "fdtd_set_timing": (C.c_int, [_P, C.c_int]),
    "fdtd_last_kernel_ms": (C.c_int, [_P, _D, _D]),
    "fdtd_batch_options_default": (None, [C.POINTER(BatchOptions)]),
    "fdtd_batch_create": (C.c_int, [C.POINTER(Scene), C.POINTER(C.c_int), C.c_int,
                                    C.POINTER(BatchOptions), _PP]),
    "fdtd_batch_run": (C.c_int, [_P]),
    "fdtd_batch_step": (C.c_int, [_P, C.c_int]),
    "fdtd_batch_steps_done": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int64)]),
    "fdtd_batch_read_listener_series": (C.c_int, [_P, C.c_int, _D]),
    "fdtd_batch_destroy": (C.c_int, [_P]),
    "fdtd_time_chunk": (C.c_int, [C.POINTER(Settings), C.c_int64, C.c_int, C.POINTER(Settings),
                                  C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "fdtd_run_time_parallel": (C.c_int, [C.POINTER(Scene), C.c_int64, C.POINTER(BatchOptions), _D,
                                         C.POINTER(C.c_int64)]),
    "fdtd_server_options_default": (None, [C.POINTER(ServerOptions)]),
    "fdtd_server_create": (C.c_int, [C.POINTER(ServerOptions), _PP]),
    "fdtd_server_submit": (C.c_int, [_P, C.POINTER(Scene), C.c_int, C.POINTER(C.c_int64)]),
    "fdtd_server_poll": (C.c_int, [_P, C.c_int64]),
    "fdtd_server_wait": (C.c_int, [_P, C.c_int64, C.c_int]),
    "fdtd_server_take_result": (C.c_int, [_P, C.c_int64, _D, _D]),
    "fdtd_server_cancel": (C.c_int, [_P, C.c_int64]),
    "fdtd_server_get_stats": (C.c_int, [_P, C.POINTER(ServerStats)]),
    "fdtd_scene_device_bytes": (C.c_int64, [C.POINTER(Scene)]),
    "fdtd_plan_check": (C.c_int, [C.POINTER(Scene), C.c_int, C.POINTER(C.c_int64)]),
    "fdtd_parsed_output_pattern": (C.c_char_p, [C.c_void_p]),
    "fdtd_composite_filename": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]),
    "fdtd_write_matrix": (C.c_int, [C.c_char_p, C.POINTER(C.c_double), C.c_int, C.c_int]),
    "fdtd_run_xml": (C.c_int, [C.c_char_p, C.c_char_p, C.c_void_p, C.POINTER(C.c_int64)]),
    "fdtd_save_snapshot": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int]),
    "fdtd_load_snapshot": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int]),
    "fdtd_server_destroy": (C.c_int, [_P]),
    "fdtd_solve_noalloc": (C.c_int, [C.POINTER(Scene), C.c_int, _P, C.c_int64, _P, _D, C.POINTER(C.c_int64)]),
}

_lib = None


class FDTDError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        name = STATUS_NAMES.get(status, str(status))
        detail = ""
        if _lib is not None:
            msg = _lib.fdtd_last_error()
            if msg:
                detail = ": " + msg.decode(errors="replace")
        super().__init__(f"{what} failed with {name}{detail}")


def load_library(path: str | None = None):
    """Loads the HIP C-ABI library; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"HIP library {p} is missing: run __graft_entry__.build() first")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def last_error():
    """Error text of the library's last failed call on this thread."""
    msg = load_library().fdtd_last_error()
    return msg.decode(errors="replace") if msg else ""


def check(status, what=""):
    if status != 0:
        raise FDTDError(status, what)


class SceneBuffers:
    """Owns the ctypes arrays a Scene points at (keeps them alive)."""

    def __init__(self, settings: dict, objects=(), impulse_sources=(), pressure_sources=(), listeners=None,
                 speaker_sources=(), constraints=(), acc_noise_sources=(), kinematics=()):
        self.settings_dict = dict(settings)
        s = Settings()
        for k, v in settings.items():
            if k in ("divisions", "min_bound"):
                arr = getattr(s, k)
                for i in range(3):
                    arr[i] = v[i]
            else:
                setattr(s, k, v)
        self.objects_list = list(objects)
        self.impulse_list = list(impulse_sources)
        self.pressure_list = list(pressure_sources)
        self._objs = (Object * max(1, len(self.objects_list)))()
        for i, o in enumerate(self.objects_list):
            self._objs[i].type = o["type"]
            for d in range(3):
                self._objs[i].center[d] = o["center"][d]
                self._objs[i].semi_axes[d] = o["semi_axes"][d]
            rot = np.asarray(o.get("rotation", np.eye(3)), dtype=np.float64).reshape(9)
            for q in range(9):
                self._objs[i].rotation[q] = rot[q]
            for d in range(3):
                self._objs[i].velocity[d] = o.get("velocity", (0.0, 0.0, 0.0))[d]
                self._objs[i].angular_velocity[d] = o.get("angular_velocity", (0.0, 0.0, 0.0))[d]
        self._isrc = (ImpulseSource * max(1, len(self.impulse_list)))()
        # Gaussian: dict(object, amplitude, width, onset); harmonic: dict(object,
        # kind=FDTD_VIB_HARMONIC, omega, [amplitude=1, phase=0, decay_rate=-1,
        # decay_center=0, t0=0, t1=0])
        for i, s_ in enumerate(self.impulse_list):
            kind = int(s_.get("kind", FDTD_VIB_GAUSSIAN))
            self._isrc[i].object = s_["object"]
            self._isrc[i].kind = kind
            if kind == FDTD_VIB_CHIRP:
                # dict(object, kind=FDTD_VIB_CHIRP, omega, omega1, t0, t1)
                self._isrc[i].omega = s_["omega"]
                self._isrc[i].omega1 = s_["omega1"]
                self._isrc[i].t0 = s_["t0"]
                self._isrc[i].t1 = s_["t1"]
            elif kind == FDTD_VIB_HARMONIC:
                self._isrc[i].amplitude = s_.get("amplitude", 1.0)
                self._isrc[i].omega = s_["omega"]
                self._isrc[i].phase = s_.get("phase", 0.0)
                self._isrc[i].decay_rate = s_.get("decay_rate", -1.0)
                self._isrc[i].decay_center = s_.get("decay_center", 0.0)
                self._isrc[i].t0 = s_.get("t0", 0.0)
                self._isrc[i].t1 = s_.get("t1", 0.0)
            else:
                self._isrc[i].amplitude = s_["amplitude"]
                self._isrc[i].width = s_["width"]
                self._isrc[i].onset = s_["onset"]
        self._psrc = (PressureSource * max(1, len(self.pressure_list)))()
        for i, s_ in enumerate(self.pressure_list):
            for d in range(3):
                self._psrc[i].position[d] = s_["position"][d]
            self._psrc[i].width_time = s_["width_time"]
            self._psrc[i].width_space = s_["width_space"]
            self._psrc[i].offset_time = s_["offset_time"]
            self._psrc[i].normalize_constant = s_["normalize_constant"]
        lst = np.zeros((0, 3)) if listeners is None else np.asarray(listeners, dtype=np.float64).reshape(-1, 3)
        self.listeners_np = np.ascontiguousarray(lst)
        self.scene = Scene()
        self.scene.settings = s
        self.scene.n_objects = len(self.objects_list)
        self.scene.objects = C.cast(self._objs, C.POINTER(Object))
        self.scene.n_impulse_sources = len(self.impulse_list)
        self.scene.impulse_sources = C.cast(self._isrc, C.POINTER(ImpulseSource))
        self.scene.n_pressure_sources = len(self.pressure_list)
        self.scene.pressure_sources = C.cast(self._psrc, C.POINTER(PressureSource))
        self.scene.n_listeners = self.listeners_np.shape[0]
        self.scene.listeners = self.listeners_np.ctypes.data_as(C.POINTER(C.c_double))
        # speaker sources: dict(position, width_space, start_time, sample_rate, samples)
        self.speaker_list = list(speaker_sources)
        self._spk_samples = [np.ascontiguousarray(sp["samples"], dtype=np.float64) for sp in self.speaker_list]
        self._spk = (SpeakerSource * max(1, len(self.speaker_list)))()
        for i, sp in enumerate(self.speaker_list):
            for d in range(3):
                self._spk[i].position[d] = sp["position"][d]
            self._spk[i].width_space = sp["width_space"]
            self._spk[i].start_time = sp["start_time"]
            self._spk[i].sample_rate = sp["sample_rate"]
            self._spk[i].n_samples = self._spk_samples[i].shape[0]
            self._spk[i].samples = self._spk_samples[i].ctypes.data_as(C.POINTER(C.c_double))
        self.scene.n_speaker_sources = len(self.speaker_list)
        self.scene.speaker_sources = C.cast(self._spk, C.POINTER(SpeakerSource))
        # plane constraints: dict(direction, sign, height)
        self.constraint_list = list(constraints)
        self._cons = (PlaneConstraint * max(1, len(self.constraint_list)))()
        for i, c in enumerate(self.constraint_list):
            self._cons[i].direction = int(c.get("direction", 1))
            self._cons[i].sign = int(c.get("sign", 1))
            self._cons[i].height = float(c.get("height", 0.0))
        self.scene.n_constraints = len(self.constraint_list)
        self.scene.constraints = C.cast(self._cons, C.POINTER(PlaneConstraint))
        # acceleration noise: dict(object, mass, center_of_mass, inv_inertia
        # (3x3), impacts=[dict(timestamp, support, impulse, position)])
        self.acc_noise_list = list(acc_noise_sources)
        self._impacts = []
        self._ans = (AccNoiseSource * max(1, len(self.acc_noise_list)))()
        for i, a in enumerate(self.acc_noise_list):
            imps = list(a.get("impacts", ()))
            arr = (Impact * max(1, len(imps)))()
            for q, im in enumerate(imps):
                arr[q].timestamp = im["timestamp"]
                arr[q].support = im["support"]
                for d in range(3):
                    arr[q].impulse[d] = im["impulse"][d]
                    arr[q].position[d] = im.get("position", (0.0, 0.0, 0.0))[d]
            self._impacts.append(arr)
            self._ans[i].object = a["object"]
            self._ans[i].n_impacts = len(imps)
            self._ans[i].mass = a["mass"]
            com = a.get("center_of_mass", (0.0, 0.0, 0.0))
            inv = np.asarray(a["inv_inertia"], dtype=np.float64).reshape(9)
            for d in range(3):
                self._ans[i].center_of_mass[d] = com[d]
            for q in range(9):
                self._ans[i].inv_inertia[q] = inv[q]
            self._ans[i].impacts = C.cast(arr, C.POINTER(Impact))
        self.scene.n_acc_noise_sources = len(self.acc_noise_list)
        self.scene.acc_noise_sources = C.cast(self._ans, C.POINTER(AccNoiseSource))
        # keyframe animations: dict(object, step_size, frames=(n, 7) tx ty tz qw qx qy qz)
        self.kinematics_list = list(kinematics)
        self._kin_frames = [np.ascontiguousarray(np.asarray(k["frames"], dtype=np.float64).reshape(-1, 7))
                            for k in self.kinematics_list]
        self._kin = (Kinematics * max(1, len(self.kinematics_list)))()
        for i, k in enumerate(self.kinematics_list):
            self._kin[i].object = k["object"]
            self._kin[i].n_frames = self._kin_frames[i].shape[0]
            self._kin[i].step_size = k["step_size"]
            self._kin[i].frames = self._kin_frames[i].ctypes.data_as(C.POINTER(C.c_double))
        self.scene.n_kinematics = len(self.kinematics_list)
        self.scene.kinematics = C.cast(self._kin, C.POINTER(Kinematics))

    @property
    def n_listeners(self):
        return self.listeners_np.shape[0]

    @property
    def divisions(self):
        return tuple(self.settings_dict["divisions"])

    @property
    def n_cells(self):
        d = self.divisions
        return d[0] * d[1] * d[2]

    def ref(self):
        return C.byref(self.scene)
