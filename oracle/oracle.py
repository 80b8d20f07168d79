"""ctypes wrapper of liblyap_oracle.so, the CPU restatement of the reference.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker. The product never imports it.
See lyap_oracle.cpp for what it restates (kernel.cu:25-154, 156-532).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

import lyapunov3d_amd as lyap

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblyap_oracle.so")

F64 = 0    # Real = double (DOUBLE_REALS)
F32W = 1   # Real = float, as written (double temporaries from the literals)

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    lib.oracle_lyap_points.argtypes = [C.c_int, P, C.c_size_t, C.c_float, C.c_uint, C.c_uint, P, P, C.c_int]
    lib.oracle_lyap_halves.argtypes = [P, C.c_size_t, C.c_float, C.c_uint, C.c_uint, C.c_uint, P, P, P, C.c_int]
    lib.oracle_volume.argtypes = [C.c_int, P, C.POINTER(lyap.LyapParams), P, P, P, C.c_uint, C.c_uint,
                                  C.c_uint, C.c_uint, C.c_uint, C.c_int]
    lib.oracle_coord.argtypes = [C.c_float, C.c_float, C.c_uint, C.c_uint]
    lib.oracle_coord.restype = C.c_float
    lib.oracle_render.argtypes = [C.c_int, C.POINTER(lyap.RenderMode), P, P, P, P, P, P, P,
                                  C.POINTER(lyap.LyapCam), C.POINTER(lyap.LyapParams), P, P,
                                  C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_int]
    lib.oracle_shade.argtypes = [P, C.c_size_t, C.POINTER(lyap.LyapCam), P, C.c_uint, P, P]
    lib.oracle_ray_setup.argtypes = [C.POINTER(lyap.LyapCam), C.POINTER(lyap.LyapParams), C.c_uint, C.c_uint,
                                     P, P, P]
    lib.oracle_ray_setup.restype = C.c_int
    lib.oracle_max_threads.restype = C.c_int
    _lib = lib
    return lib


def threads() -> int:
    return int(load().oracle_max_threads())


def _seq(seq) -> np.ndarray:
    if isinstance(seq, str):
        seq = lyap.convert_sequence(seq)
    s = list(seq)
    if s[-1] != -1:
        s.append(-1)
    return np.ascontiguousarray(s, dtype=np.int32)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def lyap_points(xyz, d, settle, accum, seq, mode=F64, nthreads=0) -> np.ndarray:
    xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
    out = np.empty(xyz.shape[0], dtype=np.float64)
    s = _seq(seq)
    load().oracle_lyap_points(mode, _ptr(xyz), xyz.shape[0], d, settle, accum, _ptr(s), _ptr(out),
                              nthreads or threads())
    return out


def lyap_halves(xyz, d, settle, accum, split, seq, nthreads=0):
    xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
    a = np.empty(xyz.shape[0]); b = np.empty(xyz.shape[0])
    s = _seq(seq)
    load().oracle_lyap_halves(_ptr(xyz), xyz.shape[0], d, settle, accum, split, _ptr(s), _ptr(a), _ptr(b),
                              nthreads or threads())
    return a, b


def coord(lo, hi, i, n) -> float:
    return float(load().oracle_coord(lo, hi, i, n))


def grid_coords(lo, hi, n) -> np.ndarray:
    """Per-axis coordinates lo + ((hi - lo) * i) / n in float32."""
    lo = np.float32(lo); hi = np.float32(hi)
    i = np.arange(n, dtype=np.float32)
    return (lo + ((hi - lo) * i) / np.float32(n)).astype(np.float32)


def volume(prm, seq, lo, hi, n, mode=F64, z_range=None, nthreads=0) -> np.ndarray:
    nx, ny, nz = (n, n, n) if isinstance(n, int) else n
    z0, z1 = z_range if z_range else (0, nz)
    out = np.empty((z1 - z0, ny, nx), dtype=np.float32)
    s = _seq(seq)
    lo3 = np.asarray(lo, dtype=np.float32); hi3 = np.asarray(hi, dtype=np.float32)
    load().oracle_volume(mode, _ptr(out), C.byref(prm), _ptr(s), _ptr(lo3), _ptr(hi3), nx, ny, nz, z0, z1,
                         nthreads or threads())
    return out


def render(cam, prm, seq, lights, num_lights, w, h, mode=None, lam=F64, rows=None, nthreads=0):
    """Returns a dict of per-pixel arrays; rows = (begin, step) to subsample."""
    r0, rs = rows if rows else (0, 1)
    out = {
        "rgba": np.zeros((h, w, 4), dtype=np.uint8),
        "points": np.zeros((h, w, 9), dtype=np.float32),
        "depth": np.full((h, w), np.nan, dtype=np.float32),
        "lambda": np.full((h, w), np.nan, dtype=np.float32),
        "evals": np.zeros((h, w), dtype=np.uint32),
        "hit_index": np.full((h, w), -2, dtype=np.int32),
        "margin": np.full((h, w), np.nan, dtype=np.float32),
    }
    s = _seq(seq)
    load().oracle_render(lam, C.byref(mode) if mode is not None else None,
                         _ptr(out["rgba"]), _ptr(out["points"]), _ptr(out["depth"]), _ptr(out["lambda"]),
                         _ptr(out["evals"]), _ptr(out["hit_index"]), _ptr(out["margin"]),
                         C.byref(cam), C.byref(prm), _ptr(s), C.cast(lights, C.c_void_p), num_lights,
                         w, h, r0, rs, nthreads or threads())
    return out


def shade(points: np.ndarray, cam, lights, num_lights):
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 9)
    col = np.empty((pts.shape[0], 4), dtype=np.float32)
    rgba = np.empty((pts.shape[0], 4), dtype=np.uint8)
    load().oracle_shade(_ptr(pts), pts.shape[0], C.byref(cam), C.cast(lights, C.c_void_p), num_lights,
                        _ptr(col), _ptr(rgba))
    return col, rgba


def ray_setup(cam, prm, sx, sy):
    V = np.empty(3, dtype=np.float32)
    t0 = C.c_float(); t1 = C.c_float()
    hit = load().oracle_ray_setup(C.byref(cam), C.byref(prm), sx, sy, _ptr(V), C.byref(t0), C.byref(t1))
    return bool(hit), V, t0.value, t1.value
