"""Python binding of liblyap_amd.so, the MI355X Lyapunov library.

The library's C ABI (include/lyap.h) is the product; this module mirrors its
types with ctypes and wraps its entries for tests, the bench and Python
callers. Device buffers are passed as integer device addresses (for example
``tensor.data_ptr()`` of a torch tensor on ``cuda:0``): torch is only used as
an allocator and for streams, never for the computation.

The reference surface this mirrors: structs.hpp (types), params.cu
(params_init), scene.cu (scene_convert_sequence, scene_cam_recalculate,
scene_lights_recalculate), kernel.hpp (kernel_calc_render,
kernel_calc_volume).

There is no CPU fallback: if the shared library is missing, ``load()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional, Sequence

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblyap_amd.so")

MAX_SEQ = 32
MAX_LIGHTS = 16

OK = 0
ERR_INVALID_ARGUMENT = -1
ERR_BAD_SEQUENCE = -2
ERR_UNSUPPORTED = -3
ERR_NO_DEVICE = -4
ERR_DEVICE = -5
ERR_BUDGET = -6
ERR_QUEUE_FULL = -7
ERR_NOT_FOUND = -8
ERR_CANCELLED = -9
ERR_PENDING = -10

PREC_PLAIN = 0
PREC_PRODUCT = 1
PREC_FP64 = 2

REFINE_WALK = 0
REFINE_BISECT = 1

REQ_VIEW = 0
REQ_VOLUME = 1

SCHEDULE_BATCHED = 0
SCHEDULE_DIRECT = 1
SCHEDULE_LEVELS = 2

(EDIT_LOCK, EDIT_ROTPAN, EDIT_MAG, EDIT_NTHRESH, EDIT_OTHRESH, EDIT_CTHRESH, EDIT_D, EDIT_JITTER,
 EDIT_L_PAN, EDIT_L_RANGE, EDIT_L_COLOR) = range(11)
UPDATE_CAM, UPDATE_LIGHTS, UPDATE_CALC, UPDATE_RENDER, UPDATE_QUIT = 1, 2, 4, 8, 16


# ---------------------------------------------------------------- types ----

class Vec(C.Structure):
    """Vec (vec3.hpp:12): 12 bytes."""
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Quat(C.Structure):
    """Quat (quat.hpp:17): 16 bytes; 16-byte aligned in C (padding is explicit
    in the structs below, since ctypes here has no alignment attribute)."""
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class Color(C.Structure):
    """Color (color.hpp:12): 16 bytes; 16-byte aligned in C."""
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class LyapCam(C.Structure):
    """LyapCam / LyapLight (structs.hpp:26-51): 224 bytes."""
    _fields_ = [
        ("C", Vec), ("_pad0", C.c_uint32), ("Q", Quat), ("M", C.c_float), ("V", Vec),
        ("S0", Vec), ("SDX", Vec), ("SDY", Vec),
        ("textureWidth", C.c_uint32), ("textureHeight", C.c_uint32),
        ("renderWidth", C.c_uint32), ("renderHeight", C.c_uint32),
        ("renderDenominator", C.c_uint32),
        ("lightInnerCone", C.c_float), ("lightOuterCone", C.c_float),
        ("lightRange", C.c_float), ("_pad1", C.c_uint32 * 3),
        ("ambient", Color), ("diffuseColor", Color), ("diffusePower", C.c_float),
        ("_pad2", C.c_uint32 * 3),
        ("specularColor", Color), ("specularPower", C.c_float),
        ("specularHardness", C.c_float), ("_pad3", C.c_uint32 * 2), ("chaosColor", Color),
    ]


LyapLight = LyapCam


class LyapParams(C.Structure):
    """LyapParams (structs.hpp:53-68): 56 bytes."""
    _fields_ = [
        ("d", C.c_float), ("settle", C.c_uint32), ("accum", C.c_uint32),
        ("stepMethod", C.c_uint32), ("nearThreshold", C.c_float),
        ("nearMultiplier", C.c_float), ("opaqueThreshold", C.c_float),
        ("chaosThreshold", C.c_float), ("depth", C.c_float), ("jitter", C.c_float),
        ("refine", C.c_float), ("gradient", C.c_float), ("lMin", C.c_float),
        ("lMax", C.c_float),
    ]


class LyapPoint(C.Structure):
    """LyapPoint (structs.hpp:70-76): 36 bytes."""
    _fields_ = [("P", Vec), ("N", Vec), ("a", C.c_float), ("c", C.c_float), ("l", C.c_float)]


class RGBA(C.Structure):
    _fields_ = [("r", C.c_ubyte), ("g", C.c_ubyte), ("b", C.c_ubyte), ("a", C.c_ubyte)]


class VolumeMode(C.Structure):
    _fields_ = [
        ("precision", C.c_int32), ("renorm", C.c_uint32), ("fp64_margin", C.c_float),
        ("converge_tau", C.c_float), ("converge_band", C.c_float),
        ("reserved", C.c_uint32 * 3),
    ]


class RenderMode(C.Structure):
    _fields_ = [
        ("refine", C.c_int32), ("bisect_steps", C.c_int32), ("normal_h", C.c_float),
        ("precision", C.c_int32), ("fp64_margin", C.c_float),
        ("lanes_per_ray", C.c_uint32), ("schedule", C.c_int32), ("reserved", C.c_uint32),
    ]


class View(C.Structure):
    _fields_ = [
        ("cam", LyapCam), ("prm", LyapParams), ("seq", C.c_int32 * (MAX_SEQ + 1)),
        ("width", C.c_uint32), ("height", C.c_uint32),
        ("rgba", C.c_void_p), ("points", C.c_void_p), ("depth", C.c_void_p),
        ("lambda_", C.c_void_p), ("evals", C.c_void_p), ("lights", C.c_void_p),
        ("num_lights", C.c_uint32), ("reserved", C.c_uint32),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("tile_rays", C.c_uint32), ("waves_per_cu", C.c_uint32),
        ("lanes_per_ray", C.c_uint32), ("slots", C.c_uint32),
        ("level_budget", C.c_uint32 * 3), ("voxel_tile", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {"tile_rays": self.tile_rays, "waves_per_cu": self.waves_per_cu,
                "lanes_per_ray": self.lanes_per_ray, "slots": self.slots,
                "level_budget": list(self.level_budget), "voxel_tile": self.voxel_tile}


class VolumeBlock(C.Structure):
    """LyapVolumeBlock: z slices [z_begin, z_end) of a volume grid."""
    _fields_ = [
        ("prm", LyapParams), ("seq", C.c_int32 * (MAX_SEQ + 1)), ("lo", C.c_float * 3),
        ("hi", C.c_float * 3), ("n", C.c_uint32 * 3), ("z_begin", C.c_uint32), ("z_end", C.c_uint32),
        ("out", C.c_void_p),
    ]


class Request(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("priority", C.c_uint32), ("_pad0", C.c_uint32 * 2), ("view", View),
        ("vol_prm", LyapParams), ("vol_seq", C.c_int32 * (MAX_SEQ + 1)),
        ("vol_lo", C.c_float * 3), ("vol_hi", C.c_float * 3), ("vol_n", C.c_uint32 * 3),
        ("vol_out", C.c_void_p), ("deadline_ms", C.c_double),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_size_t), ("render_slots", C.c_uint32),
        ("volume_slots", C.c_uint32), ("slice_ms", C.c_double), ("chunk_rays", C.c_uint32),
        ("flags", C.c_uint32), ("mode", RenderMode), ("vol_mode", VolumeMode),
        ("max_queued", C.c_uint32),
    ]


class Result(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("kind", C.c_uint32), ("submit_ms", C.c_double),
        ("start_ms", C.c_double), ("finish_ms", C.c_double), ("rays", C.c_uint64),
        ("voxels", C.c_uint64), ("evals", C.c_uint64), ("rgba", C.c_void_p),
        ("lambda_", C.c_void_p),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64), ("completed", C.c_uint64), ("cancelled", C.c_uint64),
        ("failed", C.c_uint64), ("rays", C.c_uint64), ("voxels", C.c_uint64),
        ("evals", C.c_uint64), ("launches", C.c_uint64), ("resident_bytes", C.c_size_t),
        ("peak_resident_bytes", C.c_size_t), ("queued", C.c_uint32), ("active", C.c_uint32),
        ("device_ms", C.c_double),
    ]


class Editor(C.Structure):
    """LyapEditor: the viewer's editable state (lyap_interactive.cu:36-71)."""
    _fields_ = [
        ("prm", LyapParams), ("_pad0", C.c_uint32 * 2), ("cam", LyapCam),
        ("lights", LyapLight * MAX_LIGHTS), ("home_light", LyapLight),
        ("num_lights", C.c_uint32), ("mode", C.c_int32), ("cur_light", C.c_int32),
        ("light_index", C.c_int32), ("dominant", C.c_int32), ("window_w", C.c_uint32),
        ("window_h", C.c_uint32), ("render_denominator", C.c_uint32), ("pending", C.c_uint32),
        ("_pad1", C.c_uint32 * 3),
    ]


assert C.sizeof(LyapParams) == 56
assert C.sizeof(LyapPoint) == 36
assert C.sizeof(LyapCam) == 224
assert C.sizeof(RGBA) == 4
assert LyapCam.ambient.offset == 128 and LyapCam.chaosColor.offset == 208
assert C.sizeof(View) == 480 and C.sizeof(Request) == 736
assert C.sizeof(Tuning) == 32 and C.sizeof(VolumeBlock) == 240


# ------------------------------------------------------------- library ----

class LyapError(RuntimeError):
    def __init__(self, status: int, message: str = ""):
        self.status = status
        super().__init__(f"lyap status {status}: {message}")


_lib: Optional[C.CDLL] = None

_P = C.c_void_p
_SIGS = {
    "lyap_strerror": (C.c_char_p, [C.c_int]),
    "lyap_params_init": (None, [C.POINTER(LyapParams), C.POINTER(LyapCam), C.POINTER(LyapLight), C.POINTER(C.c_uint)]),
    "lyap_scene_convert_sequence": (C.c_int, [C.c_char_p, C.POINTER(C.c_int32), C.c_size_t, C.POINTER(C.c_size_t)]),
    "lyap_scene_cam_recalculate": (None, [C.POINTER(LyapCam), C.c_uint, C.c_uint, C.c_uint]),
    "lyap_scene_lights_recalculate": (None, [C.POINTER(LyapLight), C.c_size_t]),
    "lyap_quat_axis_angle": (None, [C.POINTER(Vec), C.c_float, C.c_int, C.POINTER(Quat)]),
    "lyap_quat_between": (None, [C.POINTER(Vec), C.POINTER(Vec), C.c_float, C.POINTER(Quat)]),
    "lyap_quat_transform": (None, [C.POINTER(Quat), C.POINTER(Vec), C.POINTER(Vec)]),
    "lyap_cam_look_at": (None, [C.POINTER(LyapCam), C.POINTER(Vec), C.POINTER(Vec), C.c_float, C.c_uint, C.c_uint]),
    "lyap_default_sequence": (C.c_char_p, []),
    "lyap_create": (C.c_int, [C.POINTER(_P), C.c_int, C.c_size_t]),
    "lyap_destroy": (None, [_P]),
    "lyap_sync": (C.c_int, [_P]),
    "lyap_get_stream": (_P, [_P]),
    "lyap_last_error": (C.c_char_p, [_P]),
    "lyap_bytes_in_use": (C.c_size_t, [_P]),
    "lyap_device_alloc": (C.c_int, [_P, C.POINTER(_P), C.c_size_t]),
    "lyap_device_free": (C.c_int, [_P, _P]),
    "lyap_copy_to_device": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "lyap_copy_to_host": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "lyap_calc_volume": (C.c_int, [_P, _P, C.POINTER(LyapParams), C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint, C.c_uint, C.c_uint, C.POINTER(VolumeMode)]),
    "lyap_calc_volume_noalloc": (C.c_int, [_P, _P, C.POINTER(LyapParams), C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint, C.c_uint, C.c_uint, C.POINTER(VolumeMode)]),
    "lyap_calc_points": (C.c_int, [_P, _P, _P, C.c_size_t, C.POINTER(LyapParams), C.POINTER(C.c_int32), C.POINTER(VolumeMode)]),
    "lyap_last_reevaluated": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "lyap_calc_points_halves": (C.c_int, [_P, _P, _P, _P, _P, C.c_size_t, C.POINTER(LyapParams),
                                          C.POINTER(C.c_int32), C.c_uint]),
    "lyap_calc_render": (C.c_int, [_P, _P, _P, C.POINTER(LyapCam), C.POINTER(LyapParams), C.POINTER(C.c_int32), _P, C.c_uint, C.c_uint, C.c_uint, _P, _P, _P, C.POINTER(RenderMode)]),
    "lyap_shade_points": (C.c_int, [_P, _P, _P, C.POINTER(LyapCam), _P, C.c_uint, C.c_uint, C.c_uint]),
    "lyap_last_evals": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "lyap_debug_wave_log": (C.c_int, [_P, _P]),
    "lyap_timing_enable": (C.c_int, [_P, C.c_int]),
    "lyap_timing_collect": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "lyap_render_batch": (C.c_int, [_P, C.POINTER(View), C.c_uint, C.POINTER(RenderMode), C.POINTER(Tuning)]),
    "lyap_autotune": (C.c_int, [_P, C.POINTER(View), C.c_uint, C.POINTER(RenderMode), C.POINTER(Tuning)]),
    "lyap_run_batch": (C.c_int, [_P, C.POINTER(View), C.c_uint, C.POINTER(VolumeBlock), C.c_uint,
                                 C.POINTER(RenderMode), C.POINTER(VolumeMode), C.POINTER(Tuning)]),
    "lyap_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "lyap_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "lyap_autotune_batch": (C.c_int, [_P, C.POINTER(View), C.c_uint, C.POINTER(VolumeBlock), C.c_uint,
                                      C.POINTER(RenderMode), C.POINTER(VolumeMode), C.POINTER(Tuning)]),
    "lyap_frame_capture": (C.c_int, [_P, C.POINTER(View), C.c_uint, C.POINTER(RenderMode), C.POINTER(Tuning), C.POINTER(_P)]),
    "lyap_frame_replay": (C.c_int, [_P, _P]),
    "lyap_frame_destroy": (None, [_P]),
    "lyap_server_create": (C.c_int, [_P, C.POINTER(ServerConfig), C.POINTER(_P)]),
    "lyap_server_destroy": (None, [_P]),
    "lyap_server_submit": (C.c_int, [_P, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "lyap_server_progress": (C.c_int, [_P, C.c_double]),
    "lyap_server_poll": (C.c_int, [_P, C.c_uint64, C.POINTER(Result)]),
    "lyap_server_release": (C.c_int, [_P, C.c_uint64]),
    "lyap_server_cancel": (C.c_int, [_P, C.c_uint64]),
    "lyap_server_clock_ms": (C.c_double, [_P]),
    "lyap_server_stats": (C.c_int, [_P, C.POINTER(ServerStats)]),
    "lyap_server_drain": (C.c_int, [_P, C.c_double]),
    "lyap_server_finished": (C.c_int, [_P, C.POINTER(C.c_uint64), C.c_uint, C.POINTER(C.c_uint)]),
    "lyap_editor_init": (C.c_int, [C.POINTER(Editor), C.c_uint, C.c_uint]),
    "lyap_editor_key": (C.c_uint, [C.POINTER(Editor), C.c_int]),
    "lyap_editor_spaceball_rot": (C.c_uint, [C.POINTER(Editor), C.c_int, C.c_int, C.c_int]),
    "lyap_editor_spaceball_pan": (C.c_uint, [C.POINTER(Editor), C.c_int, C.c_int, C.c_int]),
    "lyap_editor_spaceball_button": (C.c_uint, [C.POINTER(Editor), C.c_int]),
    "lyap_editor_commit": (C.c_uint, [C.POINTER(Editor), C.c_uint, C.c_uint]),
    "lyap_ease_in_out_quart": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_double]),
    "lyap_flythrough_frames": (C.c_uint, [C.c_double, C.POINTER(C.c_double), C.c_uint]),
    "lyap_flythrough_camera": (None, [C.c_double, C.POINTER(LyapCam)]),
    "lyap_write_volume_raw": (C.c_int, [C.c_char_p, _P, C.c_size_t]),
    "lyap_write_points_raw": (C.c_int, [C.c_char_p, _P, C.c_size_t]),
    "lyap_write_ppm": (C.c_int, [C.c_char_p, _P, C.c_uint, C.c_uint]),
    "lyap_output_leaf": (C.c_int, [C.c_char_p, C.c_size_t, C.c_char_p, C.c_long, C.c_uint, C.c_uint,
                                   C.c_char_p, C.POINTER(LyapCam), C.POINTER(LyapParams)]),
    "lyap_output_name": (C.c_int, [C.c_char_p, C.c_size_t, C.c_char_p, C.c_uint, C.c_uint, C.c_uint,
                                   C.c_char_p]),
}


def load(path: str = LIB_PATH) -> C.CDLL:
    """Loads liblyap_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> Sequence[str]:
    return list(_SIGS)


def strerror(status: int) -> str:
    return load().lyap_strerror(status).decode()


# -------------------------------------------------------- host helpers ----

def params_init():
    """params_init() defaults (params.cu:21-114): (prm, cam, lights, n_lights)."""
    prm = LyapParams()
    cam = LyapCam()
    lights = (LyapLight * MAX_LIGHTS)()
    n = C.c_uint()
    load().lyap_params_init(C.byref(prm), C.byref(cam), lights, C.byref(n))
    return prm, cam, lights, n.value


def default_sequence() -> str:
    return load().lyap_default_sequence().decode()


def convert_sequence(text: str) -> list:
    """scene_convert_sequence (scene.cu:69-108): letter indices plus -1."""
    lib = load()
    n = C.c_size_t()
    rc = lib.lyap_scene_convert_sequence(text.encode(), None, 0, C.byref(n))
    if rc == ERR_BAD_SEQUENCE:
        raise LyapError(rc, f"bad sequence {text!r}")
    buf = (C.c_int32 * n.value)()
    rc = lib.lyap_scene_convert_sequence(text.encode(), buf, n.value, C.byref(n))
    if rc != OK:
        raise LyapError(rc, f"sequence {text!r}")
    return list(buf)


def seq_array(seq) -> C.Array:
    if isinstance(seq, str):
        seq = convert_sequence(seq)
    seq = list(seq)
    if not seq or seq[-1] != -1:
        seq = seq + [-1]
    return (C.c_int32 * len(seq))(*seq)


def cam_recalculate(cam: LyapCam, w: int, h: int, denom: int = 1) -> LyapCam:
    load().lyap_scene_cam_recalculate(C.byref(cam), w, h, denom)
    return cam


def lights_recalculate(lights, n: int):
    load().lyap_scene_lights_recalculate(lights, n)
    return lights


def quat_transform(q: Quat, v) -> tuple:
    out = Vec()
    load().lyap_quat_transform(C.byref(q), C.byref(Vec(*v)), C.byref(out))
    return (out.x, out.y, out.z)


def quat_axis_angle(axis, angle: float, degrees: bool = True) -> Quat:
    q = Quat()
    load().lyap_quat_axis_angle(C.byref(Vec(*axis)), angle, int(degrees), C.byref(q))
    return q


def quat_between(p, q, scale: float = 1.0) -> Quat:
    out = Quat()
    load().lyap_quat_between(C.byref(Vec(*p)), C.byref(Vec(*q)), scale, C.byref(out))
    return out


def cam_look_at(eye, target, M: float, w: int, h: int) -> LyapCam:
    cam = LyapCam()
    load().lyap_cam_look_at(C.byref(cam), C.byref(Vec(*eye)), C.byref(Vec(*target)), M, w, h)
    return cam


def make_params(**kw) -> LyapParams:
    """params_init() defaults with fields overridden by keyword."""
    prm, _, _, _ = params_init()
    for k, v in kw.items():
        setattr(prm, k, v)
    return prm


def volume_mode(precision=PREC_PLAIN, renorm=0, fp64_margin=0.0, converge_tau=0.0,
                converge_band=0.0) -> VolumeMode:
    m = VolumeMode()
    m.precision = precision
    m.renorm = renorm
    m.fp64_margin = fp64_margin
    m.converge_tau = converge_tau
    m.converge_band = converge_band
    return m


def render_mode(refine=REFINE_WALK, bisect_steps=8, normal_h=0.0, precision=PREC_PLAIN,
                fp64_margin=0.0, lanes_per_ray=0, schedule=SCHEDULE_BATCHED) -> RenderMode:
    m = RenderMode()
    m.schedule = schedule
    m.refine = refine
    m.bisect_steps = bisect_steps
    m.normal_h = normal_h
    m.precision = precision
    m.fp64_margin = fp64_margin
    m.lanes_per_ray = lanes_per_ray
    return m


# ------------------------------------------------------------- context ----

def _check(rc: int, ctx=None, what: str = ""):
    if rc != OK:
        msg = ""
        if ctx is not None:
            try:
                msg = load().lyap_last_error(ctx).decode()
            except Exception:
                msg = ""
        raise LyapError(rc, f"{what}: {strerror(rc)} {msg}".strip())


class Context:
    """A lyap_ctx: one device, one stream, the library's scratch."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self.lib = load()
        h = C.c_void_p()
        _check(self.lib.lyap_create(C.byref(h), device, hbm_budget), None, "lyap_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.lyap_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def stream(self) -> int:
        return int(self.lib.lyap_get_stream(self.h) or 0)

    def sync(self):
        _check(self.lib.lyap_sync(self.h), self.h, "lyap_sync")

    def bytes_in_use(self) -> int:
        return int(self.lib.lyap_bytes_in_use(self.h))

    def calc_volume(self, out_ptr: int, prm: LyapParams, seq, lo, hi, n,
                    mode: Optional[VolumeMode] = None):
        nx, ny, nz = (n, n, n) if isinstance(n, int) else n
        lo3 = (C.c_float * 3)(*lo)
        hi3 = (C.c_float * 3)(*hi)
        rc = self.lib.lyap_calc_volume(self.h, out_ptr, C.byref(prm), seq_array(seq), lo3, hi3,
                                       nx, ny, nz, C.byref(mode) if mode is not None else None)
        _check(rc, self.h, "lyap_calc_volume")

    def calc_volume_noalloc(self, out_ptr: int, prm: LyapParams, seq, lo, hi, n,
                            mode: Optional[VolumeMode] = None):
        """lyap_calc_volume_noalloc: calc_volume with no device allocation
        (a fixed per-device scratch array; one z slice at most 4,194,304 voxels)."""
        nx, ny, nz = (n, n, n) if isinstance(n, int) else n
        lo3 = (C.c_float * 3)(*lo)
        hi3 = (C.c_float * 3)(*hi)
        rc = self.lib.lyap_calc_volume_noalloc(self.h, out_ptr, C.byref(prm), seq_array(seq), lo3, hi3,
                                               nx, ny, nz, C.byref(mode) if mode is not None else None)
        _check(rc, self.h, "lyap_calc_volume_noalloc")

    def calc_points(self, out_ptr: int, xyz_ptr: int, n: int, prm: LyapParams, seq,
                    mode: Optional[VolumeMode] = None):
        rc = self.lib.lyap_calc_points(self.h, out_ptr, xyz_ptr, n, C.byref(prm), seq_array(seq),
                                       C.byref(mode) if mode is not None else None)
        _check(rc, self.h, "lyap_calc_points")

    def calc_points_halves(self, lam_ptr: int, h1_ptr: int, h2_ptr: int, xyz_ptr: int, n: int,
                           prm: LyapParams, seq, renorm: int = 16):
        rc = self.lib.lyap_calc_points_halves(self.h, lam_ptr, h1_ptr, h2_ptr, xyz_ptr, n, C.byref(prm),
                                              seq_array(seq), renorm)
        _check(rc, self.h, "lyap_calc_points_halves")

    def last_reevaluated(self) -> int:
        c = C.c_uint64()
        _check(self.lib.lyap_last_reevaluated(self.h, C.byref(c)), self.h, "lyap_last_reevaluated")
        return c.value

    def calc_render(self, rgba_ptr: int, points_ptr: int, cam: LyapCam, prm: LyapParams, seq,
                    lights_ptr: int, num_lights: int, w: int, h: int, depth_ptr: int = 0,
                    lambda_ptr: int = 0, evals_ptr: int = 0, mode: Optional[RenderMode] = None):
        rc = self.lib.lyap_calc_render(self.h, rgba_ptr, points_ptr or None, C.byref(cam), C.byref(prm),
                                       seq_array(seq), lights_ptr or None, num_lights, w, h,
                                       depth_ptr or None, lambda_ptr or None, evals_ptr or None,
                                       C.byref(mode) if mode is not None else None)
        _check(rc, self.h, "lyap_calc_render")

    def shade_points(self, rgba_ptr: int, points_ptr: int, cam: LyapCam, lights_ptr: int,
                     num_lights: int, w: int, h: int):
        """RGBA of a rendered frame's LyapPoint buffer under (new) lights."""
        rc = self.lib.lyap_shade_points(self.h, rgba_ptr or None, points_ptr or None, C.byref(cam),
                                        lights_ptr or None, num_lights, w, h)
        _check(rc, self.h, "lyap_shade_points")

    def timing(self, on: bool):
        """Brackets every render kernel launch with HIP events while on; off pauses."""
        _check(self.lib.lyap_timing_enable(self.h, 1 if on else 0), self.h, "lyap_timing_enable")

    def timing_collect(self):
        """(summed kernel ms, launches) recorded since the last collection."""
        ms = C.c_double()
        n = C.c_uint64()
        _check(self.lib.lyap_timing_collect(self.h, C.byref(ms), C.byref(n)), self.h, "lyap_timing_collect")
        return ms.value, n.value

    def last_evals(self) -> int:
        c = C.c_uint64()
        _check(self.lib.lyap_last_evals(self.h, C.byref(c)), self.h, "lyap_last_evals")
        return c.value

    # ---- batched views -------------------------------------------------
    def render_batch(self, views, mode: Optional[RenderMode] = None, tuning: Optional[Tuning] = None):
        arr = (View * len(views))(*views)
        rc = self.lib.lyap_render_batch(self.h, arr, len(views), C.byref(mode) if mode is not None else None,
                                        C.byref(tuning) if tuning is not None else None)
        _check(rc, self.h, "lyap_render_batch")

    def autotune(self, views, mode: Optional[RenderMode] = None) -> Tuning:
        arr = (View * len(views))(*views)
        best = Tuning()
        rc = self.lib.lyap_autotune(self.h, arr, len(views), C.byref(mode) if mode is not None else None,
                                    C.byref(best))
        _check(rc, self.h, "lyap_autotune")
        return best

    def run_batch(self, views, blocks, mode: Optional[RenderMode] = None,
                  vol_mode: Optional[VolumeMode] = None, tuning: Optional[Tuning] = None):
        """Views and volume blocks in one persistent launch sequence (lyap_run_batch)."""
        va = (View * max(len(views), 1))(*views)
        ba = (VolumeBlock * max(len(blocks), 1))(*blocks)
        rc = self.lib.lyap_run_batch(self.h, va, len(views), ba, len(blocks),
                                     C.byref(mode) if mode is not None else None,
                                     C.byref(vol_mode) if vol_mode is not None else None,
                                     C.byref(tuning) if tuning is not None else None)
        _check(rc, self.h, "lyap_run_batch")

    def autotune_batch(self, views, blocks, mode: Optional[RenderMode] = None,
                       vol_mode: Optional[VolumeMode] = None) -> Tuning:
        va = (View * max(len(views), 1))(*views)
        ba = (VolumeBlock * max(len(blocks), 1))(*blocks)
        best = Tuning()
        rc = self.lib.lyap_autotune_batch(self.h, va, len(views), ba, len(blocks),
                                          C.byref(mode) if mode is not None else None,
                                          C.byref(vol_mode) if vol_mode is not None else None, C.byref(best))
        _check(rc, self.h, "lyap_autotune_batch")
        return best

    def set_tuning(self, tuning: Optional[Tuning]):
        """Sets (or with None clears) the context's tuned launch shape."""
        _check(self.lib.lyap_set_tuning(self.h, C.byref(tuning) if tuning is not None else None), self.h,
               "lyap_set_tuning")

    def get_tuning(self) -> Tuning:
        t = Tuning()
        _check(self.lib.lyap_get_tuning(self.h, C.byref(t)), self.h, "lyap_get_tuning")
        return t

    def frame_capture(self, views, mode: Optional[RenderMode] = None, tuning: Optional[Tuning] = None):
        arr = (View * len(views))(*views)
        f = C.c_void_p()
        rc = self.lib.lyap_frame_capture(self.h, arr, len(views), C.byref(mode) if mode is not None else None,
                                         C.byref(tuning) if tuning is not None else None, C.byref(f))
        _check(rc, self.h, "lyap_frame_capture")
        return Frame(self, f)


class Frame:
    """A captured frame (lyap_frame): replay() relaunches its graph."""

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self.h = handle

    def replay(self):
        _check(self.ctx.lib.lyap_frame_replay(self.ctx.h, self.h), self.ctx.h, "lyap_frame_replay")

    def close(self):
        if self.h:
            self.ctx.lib.lyap_frame_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_view(cam: LyapCam, prm: LyapParams, seq, w: int, h: int, rgba: int, points: int = 0,
              depth: int = 0, lam: int = 0, evals: int = 0, lights: int = 0, num_lights: int = 0) -> View:
    """A View for the batched entries; device outputs as integer addresses."""
    v = View()
    v.cam = cam
    v.prm = prm
    s = list(seq_array(seq))
    if len(s) > MAX_SEQ + 1:
        raise LyapError(ERR_BAD_SEQUENCE, "sequence longer than %d" % MAX_SEQ)
    for i, x in enumerate(s):
        v.seq[i] = x
    v.width, v.height = w, h
    v.rgba, v.points, v.depth, v.lambda_, v.evals = rgba, points or None, depth or None, lam or None, evals or None
    v.lights, v.num_lights = lights or None, num_lights
    return v


def make_block(prm: LyapParams, seq, lo, hi, n, out: int, z_begin: int = 0, z_end: int = 0) -> VolumeBlock:
    """A VolumeBlock: slices [z_begin, z_end) of the n (or (nx, ny, nz)) grid over
    [lo, hi); `out` is the device address of slice z_begin's first voxel."""
    b = VolumeBlock()
    b.prm = prm
    s = list(seq_array(seq))
    if len(s) > MAX_SEQ + 1:
        raise LyapError(ERR_BAD_SEQUENCE, "sequence longer than %d" % MAX_SEQ)
    for i, x in enumerate(s):
        b.seq[i] = x
    nx, ny, nz = (n, n, n) if isinstance(n, int) else n
    for k in range(3):
        b.lo[k], b.hi[k] = lo[k], hi[k]
    b.n[0], b.n[1], b.n[2] = nx, ny, nz
    b.z_begin, b.z_end = z_begin, z_end
    b.out = out or None
    return b


# ------------------------------------------------------------- serving ----

def view_request(view: View, deadline_ms: float = 0.0) -> Request:
    """A VIEW request; leave view.rgba 0 to have the server keep RGBA and lambda resident."""
    r = Request()
    r.kind = REQ_VIEW
    r.view = view
    r.deadline_ms = deadline_ms
    return r


def volume_request(prm: LyapParams, seq, lo, hi, n, out: int = 0, deadline_ms: float = 0.0) -> Request:
    """A VOLUME request over the box [lo, hi) with n (or (nx, ny, nz)) samples per axis."""
    r = Request()
    r.kind = REQ_VOLUME
    r.vol_prm = prm
    s = list(seq_array(seq))
    for i, x in enumerate(s):
        r.vol_seq[i] = x
    nx, ny, nz = (n, n, n) if isinstance(n, int) else n
    for k in range(3):
        r.vol_lo[k] = lo[k]
        r.vol_hi[k] = hi[k]
    r.vol_n[0], r.vol_n[1], r.vol_n[2] = nx, ny, nz
    r.vol_out = out or None
    r.deadline_ms = deadline_ms
    return r


SERVER_SEPARATE_VOLUMES = 1


def server_config(hbm_budget_bytes: int = 0, render_slots: int = 0, volume_slots: int = 0,
                  slice_ms: float = 0.0, chunk_rays: int = 0, mode: Optional[RenderMode] = None,
                  vol_mode: Optional[VolumeMode] = None, flags: int = 0,
                  max_queued: int = 0) -> ServerConfig:
    c = ServerConfig()
    c.flags = flags
    c.max_queued = max_queued
    c.hbm_budget_bytes = hbm_budget_bytes
    c.render_slots, c.volume_slots = render_slots, volume_slots
    c.slice_ms, c.chunk_rays = slice_ms, chunk_rays
    c.mode = mode if mode is not None else render_mode(REFINE_BISECT, 8, 1e-3, PREC_PRODUCT)
    c.vol_mode = vol_mode if vol_mode is not None else volume_mode(PREC_PRODUCT)
    return c


class Server:
    """lyap_server: continuous batching of view and volume requests on the context's device."""

    def __init__(self, ctx: Context, cfg: Optional[ServerConfig] = None):
        self.ctx = ctx
        self.lib = ctx.lib
        h = C.c_void_p()
        _check(self.lib.lyap_server_create(ctx.h, C.byref(cfg) if cfg is not None else None, C.byref(h)),
               ctx.h, "lyap_server_create")
        self.h = h

    def submit(self, req: Request) -> int:
        t = C.c_uint64()
        _check(self.lib.lyap_server_submit(self.h, C.byref(req), C.byref(t)), self.ctx.h, "lyap_server_submit")
        return t.value

    def progress(self, wait_ms: float = 0.0):
        _check(self.lib.lyap_server_progress(self.h, wait_ms), self.ctx.h, "lyap_server_progress")

    def poll(self, ticket: int) -> Result:
        r = Result()
        _check(self.lib.lyap_server_poll(self.h, ticket, C.byref(r)), self.ctx.h, "lyap_server_poll")
        return r

    def release(self, ticket: int):
        _check(self.lib.lyap_server_release(self.h, ticket), self.ctx.h, "lyap_server_release")

    def cancel(self, ticket: int):
        _check(self.lib.lyap_server_cancel(self.h, ticket), self.ctx.h, "lyap_server_cancel")

    def clock_ms(self) -> float:
        return float(self.lib.lyap_server_clock_ms(self.h))

    def stats(self) -> ServerStats:
        s = ServerStats()
        _check(self.lib.lyap_server_stats(self.h, C.byref(s)), self.ctx.h, "lyap_server_stats")
        return s

    def finished(self, cap: int = 256) -> list:
        """Tickets finished since the last call, in finishing order."""
        out = []
        buf = (C.c_uint64 * cap)()
        n = C.c_uint()
        while True:
            _check(self.lib.lyap_server_finished(self.h, buf, cap, C.byref(n)), self.ctx.h, "lyap_server_finished")
            out.extend(buf[i] for i in range(n.value))
            if n.value < cap:
                return out

    def drain(self, timeout_ms: float = 0.0):
        _check(self.lib.lyap_server_drain(self.h, timeout_ms), self.ctx.h, "lyap_server_drain")

    def close(self):
        if self.h:
            self.lib.lyap_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------- editing, files ----

def editor(window_w: int = 512, window_h: int = 512) -> Editor:
    """A LyapEditor initialised like the viewer (params_init + its controls)."""
    ed = Editor()
    _check(load().lyap_editor_init(C.byref(ed), window_w, window_h), None, "lyap_editor_init")
    return ed


def ease_in_out_quart(t: float, b: float = 0.0, c: float = 1.0, d: float = 1.0) -> float:
    return float(load().lyap_ease_in_out_quart(t, b, c, d))


def flythrough_times(step: float = 0.01) -> list:
    """The frame times t of scale.pl's loop (accumulated in double)."""
    n = load().lyap_flythrough_frames(step, None, 0)
    buf = (C.c_double * max(n, 1))()
    load().lyap_flythrough_frames(step, buf, n)
    return [buf[i] for i in range(n)]


def flythrough_camera(eased: float, w: int, h: int, M: Optional[float] = None) -> LyapCam:
    """The scripted camera for an eased value, recalculated for a w x h frame
    (M from params_init unless given)."""
    _, cam, _, _ = params_init()
    if M is not None:
        cam.M = M
    load().lyap_flythrough_camera(eased, C.byref(cam))
    load().lyap_scene_cam_recalculate(C.byref(cam), w, h, 1)
    return cam


def output_leaf(kind: str, stamp: int, w: int, h: int, sequence: str, cam: LyapCam, prm: LyapParams) -> str:
    buf = C.create_string_buffer(512)
    _check(load().lyap_output_leaf(buf, 512, kind.encode(), stamp, w, h, sequence.encode(), C.byref(cam),
                                   C.byref(prm)), None, "lyap_output_leaf")
    return buf.value.decode()


def output_name(leaf: str, hours: int, minutes: int, seconds: int, ext: str) -> str:
    buf = C.create_string_buffer(600)
    _check(load().lyap_output_name(buf, 600, leaf.encode(), hours, minutes, seconds, ext.encode()), None,
           "lyap_output_name")
    return buf.value.decode()


def write_ppm(path: str, rgba, w: int, h: int):
    """rgba: a host buffer of w*h RGBA bytes (numpy uint8 array or bytes)."""
    import numpy as np
    a = np.ascontiguousarray(np.asarray(rgba, dtype=np.uint8).reshape(-1))
    if a.size < 4 * w * h:
        raise LyapError(ERR_INVALID_ARGUMENT, "rgba buffer smaller than w*h*4")
    _check(load().lyap_write_ppm(path.encode(), a.ctypes.data_as(C.c_void_p), w, h), None, "lyap_write_ppm")


def write_volume_raw(path: str, exps):
    import numpy as np
    a = np.ascontiguousarray(np.asarray(exps, dtype=np.float32).reshape(-1))
    _check(load().lyap_write_volume_raw(path.encode(), a.ctypes.data_as(C.c_void_p), a.size), None,
           "lyap_write_volume_raw")


def write_points_raw(path: str, points, n: int):
    """points: a host buffer of n LyapPoint records (36 bytes each)."""
    import numpy as np
    a = np.ascontiguousarray(np.asarray(points).view(np.uint8).reshape(-1))
    if a.size < 36 * n:
        raise LyapError(ERR_INVALID_ARGUMENT, "points buffer smaller than n records")
    _check(load().lyap_write_points_raw(path.encode(), a.ctypes.data_as(C.c_void_p), n), None,
           "lyap_write_points_raw")
