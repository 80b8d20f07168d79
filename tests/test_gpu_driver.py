"""The compiled C++ driver (tests/driver/lyap_driver.cpp) against the oracle.

The driver calls liblyap_amd.so through lyap.h only, the way the reference's
own drivers would after the drop-in:
 - `calculate`: lyap_calculate.cu's flow with params_init defaults (BCABA,
   d = 2.1, settle 18 + accum 1008) over the reference's fixed box [0, 4)^3,
   written as exps.raw. This box holds the r < 1 region and the r = 0 faces
   (lambda NaN: the derivative there is 0). FP64 matches the FP64 oracle to
   1e-6 with the same NaNs; the FP32 forms sit at the reference's own FP32
   level (the oracle's FP32-as-written form) less 0.002;
 - `interactive`: one frame of lyap_interactive.cu with the reference's
   method (refinement walk, jitter 0.5, near step, normal step * gradient,
   the scripted camera and spotlight), written as Points raw and P3 PPM.
   FP64 reproduces the oracle's hit records; the FP32 (PLAIN) walk agrees on
   hits and on the hit's ray parameter within one march step at least as
   often as the oracle's FP32-as-written walk, less 0.002.
"""
import os
import subprocess

import numpy as np
import pytest

import lyapunov3d_amd as lyap
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER_DIR = os.path.join(ROOT, "tests", "driver")
DRIVER = os.path.join(DRIVER_DIR, "lyap_driver")
SLACK = 0.002


def driver(*args):
    if not os.path.exists(DRIVER):
        subprocess.run(["make", "-C", DRIVER_DIR], check=True, capture_output=True)
    print("driver:", " ".join(str(a) for a in args), flush=True)
    r = subprocess.run([DRIVER, *[str(a) for a in args]], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def vol_stats(got, ref, h1, h2):
    stable = (ref < -1e-3) & (np.abs(h1 - h2) < 1e-4)
    d = np.abs(got[stable].astype(np.float64) - ref[stable])
    within = float(np.mean(np.nan_to_num(d, nan=9.0) <= 1e-3)) if d.size else 1.0
    sign = float(np.mean((np.sign(got) == np.sign(ref)) | (np.isnan(got) & np.isnan(ref))))
    return within, sign


@pytest.fixture(scope="module")
def calc_reference():
    prm, _, _, _ = lyap.params_init()
    seq = lyap.default_sequence()
    n = 64
    lo, hi = (0.0, 0.0, 0.0), (4.0, 4.0, 4.0)
    f64 = O.volume(prm, seq, lo, hi, n, mode=O.F64)
    f32 = O.volume(prm, seq, lo, hi, n, mode=O.F32W)
    c = O.grid_coords(0.0, 4.0, n)
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    xyz = np.stack([xx, yy, zz], -1).reshape(-1, 3)
    h1, h2 = O.lyap_halves(xyz, prm.d, prm.settle, prm.accum, prm.accum // 2, seq)
    return n, f64, f32, h1.reshape(f64.shape), h2.reshape(f64.shape)


@pytest.mark.parametrize("prec", ["fp64", "plain", "product"])
def test_calculate_flow(tmp_path, prec, calc_reference):
    n, f64, f32, h1, h2 = calc_reference
    out = tmp_path / "exps.raw"
    driver("calculate", n, prec, out)
    got = np.fromfile(out, dtype=np.float32)
    assert got.size == n ** 3
    got = got.reshape(n, n, n)
    # index 0 on an axis is coordinate 0: r = 0 there, derivative 0, lambda NaN
    assert np.all(np.isnan(got[0, :, :])) and np.all(np.isnan(f64[0, :, :]))
    assert np.mean(np.isnan(f64)) < 0.2
    # the r < 1 corner: the orbit dies towards 0 and lambda is the mean log of r there
    assert np.all(f64[1:4, 1:4, 1:4] < 0)
    if prec == "fp64":
        assert np.array_equal(np.isnan(got), np.isnan(f64))
        m = ~np.isnan(f64)
        assert np.max(np.abs(got[m].astype(np.float64) - f64[m])) <= 1e-6
    else:
        ref = f64.astype(np.float64)
        w, s = vol_stats(got, ref, h1, h2)
        rw, rs = vol_stats(f32, ref, h1, h2)
        print(f"calculate {prec}: within {w:.5f} sign {s:.5f} (reference FP32 {rw:.5f} {rs:.5f})")
        assert w >= rw - SLACK and s >= rs - SLACK


def read_points(path, n):
    a = np.fromfile(path, dtype=np.float32)
    assert a.size == 9 * n
    return a.reshape(n, 9)


def read_ppm(path, w, h):
    text = open(path).read().split()
    assert text[:4] == ["P3", str(w), str(h), "255"]
    v = np.array([int(x) for x in text[4:]], dtype=np.int64)
    assert v.size == 3 * w * h
    return v.reshape(h, w, 3)


@pytest.fixture(scope="module")
def frame_reference():
    prm, cam, lights, n_lights = lyap.params_init()
    w, h = 64, 48
    lyap.cam_recalculate(cam, w, h, 1)
    lyap.lights_recalculate(lights, n_lights)
    seq = lyap.default_sequence()
    f64 = O.render(cam, prm, seq, lights, n_lights, w, h, mode=lyap.render_mode(), lam=O.F64)
    f32 = O.render(cam, prm, seq, lights, n_lights, w, h, mode=lyap.render_mode(), lam=O.F32W)
    return (w, h, cam, prm), f64, f32


def ray_params(cam, prm, w, h, P):
    """Ray parameter t of each hit position and the largest march step of
    the reference method there (1.5 x the step: jitter 0.5)."""
    t = np.full(w * h, np.nan)
    step = np.zeros(w * h)
    C = np.array([cam.C.x, cam.C.y, cam.C.z], dtype=np.float64)
    for i in range(w * h):
        y, x = divmod(i, w)
        _, V, t0, t1 = O.ray_setup(cam, prm, x, y)
        V = np.asarray(V, dtype=np.float64)
        t[i] = np.dot(P[i] - C, V) / np.dot(V, V)
        step[i] = 1.5 * np.linalg.norm(V) / prm.depth
    return t, step


@pytest.mark.parametrize("prec", ["fp64", "plain"])
def test_interactive_frame(tmp_path, prec, frame_reference):
    (w, h, cam, prm), f64, f32 = frame_reference
    pts_path, ppm_path = tmp_path / "Points.raw", tmp_path / "Render.ppm"
    driver("interactive", w, h, prec, pts_path, ppm_path)
    pts = read_points(pts_path, w * h)
    rgb = read_ppm(ppm_path, w, h)
    ref_pts = f64["points"].reshape(-1, 9)
    hit = ~np.isnan(pts[:, 6])
    ref_hit = ~np.isnan(ref_pts[:, 6])
    assert ref_hit.mean() > 0.5
    if prec == "fp64":
        assert np.array_equal(hit, ref_hit)
        assert np.array_equal(pts[hit, :3], ref_pts[hit, :3])
        np.testing.assert_allclose(pts[hit, 3:6], ref_pts[hit, 3:6], atol=1e-3)
        np.testing.assert_allclose(pts[hit, 6:8], ref_pts[hit, 6:8], atol=1e-5)
        assert np.max(np.abs(rgb - f64["rgba"][..., :3].astype(np.int64))) <= 2
    else:
        t_ref, step = ray_params(cam, prm, w, h, ref_pts[:, :3].astype(np.float64))

        def agreement(p):
            hp = ~np.isnan(p[:, 6])
            both = hp & ref_hit
            tp, _ = ray_params(cam, prm, w, h, p[:, :3].astype(np.float64))
            within = np.mean(np.abs(tp[both] - t_ref[both]) <= step[both])
            return float(np.mean(hp == ref_hit)), float(within)

        a, wn = agreement(pts)
        ra, rw = agreement(f32["points"].reshape(-1, 9))
        print(f"interactive plain: hit agreement {a:.4f} within one step {wn:.4f} "
              f"(reference FP32 {ra:.4f} {rw:.4f})")
        assert a >= ra - SLACK and wn >= rw - SLACK
        # misses are transparent black; hits are shaded
        assert np.all(rgb[~hit.reshape(h, w)] == 0)
