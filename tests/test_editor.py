"""Headless editing (lyap_editor_*), the fly-through camera path and the
on-disk formats, on the CPU.

Expected values are worked out by hand from the reference's handlers
(lyap_interactive.cu:144-463), scale.pl and the writers
(lyap_interactive.cu:569-655, lyap_calculate.cu:79-86).
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import lyapunov3d_amd as lyap


def lib():
    return lyap.load()


def key(ed, k):
    return lib().lyap_editor_key(C.byref(ed), ord(k))


def rot(ed, x, y, z):
    return lib().lyap_editor_spaceball_rot(C.byref(ed), x, y, z)


def pan(ed, x, y, z):
    return lib().lyap_editor_spaceball_pan(C.byref(ed), x, y, z)


def f32(x):
    return float(np.float32(x))


def test_init_matches_params_init():
    ed = lyap.editor(640, 480)
    prm, cam, lights, n = lyap.params_init()
    assert bytes(ed.prm) == bytes(prm)
    assert bytes(ed.cam) == bytes(cam)
    assert ed.num_lights == n == 1
    assert bytes(ed.home_light) == bytes(lights[0])
    assert ed.mode == lyap.EDIT_ROTPAN and ed.cur_light == -1 and ed.light_index == 0
    assert ed.dominant == 0 and ed.render_denominator == 1
    assert ed.pending == lyap.UPDATE_CAM | lyap.UPDATE_LIGHTS | lyap.UPDATE_CALC


def test_mode_nudges_and_clamps():
    ed = lyap.editor()
    assert key(ed, "d") == 0 and ed.mode == lyap.EDIT_D
    assert rot(ed, 100, 0, 0) == lyap.UPDATE_CALC
    # d (float) + x / 10000.0 in double, stored as float
    assert ed.prm.d == f32(f32(2.1) + 100 / 10000.0)
    rot(ed, 100000, 0, 0)
    assert ed.prm.d == ed.prm.lMax == 4.0
    rot(ed, -10 ** 6, 0, 0)
    assert ed.prm.d == ed.prm.lMin == 0.0
    key(ed, "j")
    rot(ed, -5, 0, 0)
    assert ed.prm.jitter == f32(0.5 - 5 / 10000.0)
    rot(ed, 10 ** 6, 0, 0)
    assert ed.prm.jitter == 1.0
    rot(ed, -10 ** 6, 0, 0)
    assert ed.prm.jitter == 0.0
    # opaque threshold may not pass the near threshold, and the other way
    key(ed, "t")
    rot(ed, 10000, 0, 0)                  # -0.75 + 1 > near (-1): pinned to near
    assert ed.prm.opaqueThreshold == ed.prm.nearThreshold == -1.0
    key(ed, "y")
    rot(ed, -10000, 0, 0)                 # near -2 < opaque: pinned to opaque
    assert ed.prm.nearThreshold == ed.prm.opaqueThreshold == -1.0
    key(ed, "c")
    rot(ed, 2500, 0, 0)
    assert ed.prm.chaosThreshold == f32(-0.5 + 0.25)
    key(ed, "m")
    assert rot(ed, 250, 0, 0) == lyap.UPDATE_CAM | lyap.UPDATE_CALC
    assert ed.cam.M == f32(f32(0.45) + 250 / 2500.0)


def test_step_keys():
    ed = lyap.editor()
    assert key(ed, "z") == lyap.UPDATE_CALC and ed.prm.depth == 4095.0
    key(ed, "Z")
    assert ed.prm.depth == 2047.5
    key(ed, "A")
    assert ed.prm.depth == 4095.0
    key(ed, "a")
    assert ed.prm.depth == 4096.0
    ed.prm.depth = 1.5
    key(ed, "z")
    assert ed.prm.depth == 1.0
    key(ed, "Z")
    assert ed.prm.depth == 1.0
    key(ed, "s")
    assert ed.prm.accum == 1009
    key(ed, "S")
    assert ed.prm.accum == 2018
    key(ed, "X")
    assert ed.prm.accum == 1009
    key(ed, "x")
    assert ed.prm.accum == 1008
    ed.prm.accum = 1
    key(ed, "x")
    assert ed.prm.accum == 1              # 0 goes back to 1
    ed.prm.accum = 1
    key(ed, "X")
    assert ed.prm.accum == 1
    key(ed, "[")
    assert ed.prm.stepMethod == 1
    key(ed, "}")
    assert ed.prm.stepMethod == 2
    assert key(ed, "-") == lyap.UPDATE_CALC and ed.render_denominator == 2
    assert key(ed, "=") == lyap.UPDATE_CALC and ed.render_denominator == 1
    assert key(ed, "+") == 0 and ed.render_denominator == 1
    assert key(ed, "q") == lyap.UPDATE_QUIT
    assert key(ed, "?") == 0


def soften(v, lim, rng, scale):
    if v <= -lim:
        return scale * (v + lim) / (rng - lim)
    if v >= lim:
        return scale * (v - lim) / (rng - lim)
    return 0.0


def test_pan_moves_along_the_view_axes():
    ed = lyap.editor()
    c0 = (ed.cam.C.x, ed.cam.C.y, ed.cam.C.z)
    assert pan(ed, 40, -40, 49) == lyap.UPDATE_CAM | lyap.UPDATE_LIGHTS | lyap.UPDATE_CALC
    assert (ed.cam.C.x, ed.cam.C.y, ed.cam.C.z) == c0      # inside the dead zone
    pan(ed, 100, 0, 0)
    d = soften(100, 50, 500, 0.05)
    ax = lyap.quat_transform(ed.cam.Q, (1.0, 0.0, 0.0))
    moved = np.array([ed.cam.C.x, ed.cam.C.y, ed.cam.C.z]) - np.array(c0)
    np.testing.assert_allclose(moved, d * np.array(ax), atol=1e-6)
    # y is flipped (screen down), z passes through
    ed2 = lyap.editor()
    pan(ed2, 0, 300, 0)
    up = lyap.quat_transform(ed2.cam.Q, (0.0, 1.0, 0.0))
    moved = np.array([ed2.cam.C.x, ed2.cam.C.y, ed2.cam.C.z]) - np.array(c0)
    np.testing.assert_allclose(moved, -soften(300, 50, 500, 0.05) * np.array(up), atol=1e-6)


def test_dominant_axis_lock():
    ed = lyap.editor()
    lib().lyap_editor_spaceball_button(C.byref(ed), 1)
    assert ed.dominant == 1
    c0 = np.array([ed.cam.C.x, ed.cam.C.y, ed.cam.C.z])
    pan(ed, 100, 300, 0)                  # (x, -y, z) -> only the y part stays
    up = np.array(lyap.quat_transform(ed.cam.Q, (0.0, 1.0, 0.0)))
    moved = np.array([ed.cam.C.x, ed.cam.C.y, ed.cam.C.z]) - c0
    np.testing.assert_allclose(moved, -soften(300, 50, 500, 0.05) * up, atol=1e-6)
    lib().lyap_editor_spaceball_button(C.byref(ed), 2)
    assert ed.dominant == 1
    lib().lyap_editor_spaceball_button(C.byref(ed), 1)
    assert ed.dominant == 0


def qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return (aw * bx + ax * bw + ay * bz - az * by,
            aw * by - ax * bz + ay * bw + az * bx,
            aw * bz + ax * by - ay * bx + az * bw,
            aw * bw - ax * bx - ay * by - az * bz)


def test_rotate_composes_a_body_rotation():
    ed = lyap.editor()
    q0 = (ed.cam.Q.x, ed.cam.Q.y, ed.cam.Q.z, ed.cam.Q.w)
    assert rot(ed, 0, 0, 0) == 0
    assert rot(ed, 500, 0, 0) == lyap.UPDATE_CAM | lyap.UPDATE_CALC
    # softened axis (1, 0, 0), length 1: 0.05 rad about the camera's x axis
    h = 0.025
    want = qmul(q0, (math.sin(h), 0.0, 0.0, math.cos(h)))
    got = (ed.cam.Q.x, ed.cam.Q.y, ed.cam.Q.z, ed.cam.Q.w)
    np.testing.assert_allclose(got, want, atol=2e-7)
    assert abs(sum(v * v for v in got) - 1.0) < 1e-6


def test_lights_select_pan_colour_and_home():
    ed = lyap.editor(512, 512)
    assert key(ed, "1") == lyap.UPDATE_CAM | lyap.UPDATE_RENDER | lyap.UPDATE_LIGHTS
    assert ed.cur_light == 0 and ed.mode == lyap.EDIT_ROTPAN
    L0 = np.array([ed.lights[0].C.x, ed.lights[0].C.y, ed.lights[0].C.z])
    c0 = np.array([ed.cam.C.x, ed.cam.C.y, ed.cam.C.z])
    pan(ed, 100, 0, 0)                    # moves the light, not the camera
    moved = np.array([ed.lights[0].C.x, ed.lights[0].C.y, ed.lights[0].C.z]) - L0
    ax = np.array(lyap.quat_transform(ed.lights[0].Q, (1.0, 0.0, 0.0)))
    np.testing.assert_allclose(moved, soften(100, 50, 500, 0.05) * ax, atol=1e-6)
    assert np.array_equal(c0, [ed.cam.C.x, ed.cam.C.y, ed.cam.C.z])
    key(ed, "o")
    assert ed.mode == lyap.EDIT_L_COLOR
    pan(ed, -500, 0, 0)                   # delta x = -0.25
    assert ed.lights[0].diffuseColor.x == f32(1.0 - 0.25)
    pan(ed, 10 ** 6, 0, 0)
    assert ed.lights[0].diffuseColor.x == 1.0
    assert bytes(ed.lights[0].specularColor) == bytes(ed.lights[0].diffuseColor)
    # a light that does not exist is not edited
    key(ed, "3")
    key(ed, "l")
    before = bytes(ed.lights[2])
    assert pan(ed, 400, 0, 0) == 0 and bytes(ed.lights[2]) == before
    # home: light 0 back to its first state, at the camera
    key(ed, "1")
    assert key(ed, "h") == lyap.UPDATE_LIGHTS | lyap.UPDATE_RENDER
    L = ed.lights[0]
    assert (L.C.x, L.C.y, L.C.z) == (ed.cam.C.x, ed.cam.C.y, ed.cam.C.z)
    assert (L.Q.x, L.Q.y, L.Q.z, L.Q.w) == (ed.cam.Q.x, ed.cam.Q.y, ed.cam.Q.z, ed.cam.Q.w)
    assert L.diffuseColor.x == ed.home_light.diffuseColor.x
    key(ed, "0")
    assert ed.cur_light == -1


def test_commit_recalculates_what_changed():
    ed = lyap.editor(512, 512)
    ed.pending = 0
    key(ed, "m")
    rot(ed, 500, 0, 0)
    flags = lib().lyap_editor_commit(C.byref(ed), 256, 128)
    assert flags == lyap.UPDATE_CAM | lyap.UPDATE_CALC and ed.pending == 0
    ref = lyap.LyapCam.from_buffer_copy(bytes(ed.cam))
    lyap.cam_recalculate(ref, 256, 128, 1)
    assert bytes(ref) == bytes(ed.cam)
    assert ed.cam.renderWidth == 256 and ed.cam.renderHeight == 128


def test_ease_in_out():
    e = lyap.ease_in_out_quart
    assert e(0.0) == 0.0 and e(1.0) == 1.0 and e(0.5) == 0.5
    assert e(0.25) == 0.125 and e(0.75) == 0.875
    assert e(0.3, 2.0, 4.0, 2.0) == 2.0 + 2.0 * 0.3 ** 2


def test_flythrough_frames():
    ts = lyap.flythrough_times(0.01)
    assert len(ts) == 101
    t = 0.0
    for x in ts:
        assert x == t
        t += 0.01
    assert ts[-1] <= 1.00001 and ts[-1] + 0.01 > 1.00001


def test_flythrough_camera_path():
    prm, cam0, _, _ = lyap.params_init()
    # the end of the path is the default scripted camera
    cam1 = lyap.flythrough_camera(1.0, 64, 64)
    assert (cam1.C.x, cam1.C.y, cam1.C.z) == (cam0.C.x, cam0.C.y, cam0.C.z)
    assert (cam1.Q.x, cam1.Q.y, cam1.Q.z, cam1.Q.w) == (cam0.Q.x, cam0.Q.y, cam0.Q.z, cam0.Q.w)
    # the middle: nlerp of -20 and +20 degrees is the identity, so the view
    # axis is -(1,1,1)/sqrt 3 and the camera one unit behind (3.55, 3.55, 3.55)
    cm = lyap.flythrough_camera(0.5, 64, 64)
    s = 1.0 / math.sqrt(3.0)
    np.testing.assert_allclose([cm.C.x, cm.C.y, cm.C.z], [3.55 + s] * 3, atol=2e-6)
    np.testing.assert_allclose([cm.V.x, cm.V.y, cm.V.z], [-s] * 3, atol=2e-6)
    # the two ends are mirror images in the plane spanned by (1,1,1) and
    # the side vector (-1,1,1): their view axes make equal angles with -(1,1,1)
    ca = lyap.flythrough_camera(0.0, 64, 64)
    cb = lyap.flythrough_camera(1.0, 64, 64)
    va = np.array([ca.V.x, ca.V.y, ca.V.z])
    vb = np.array([cb.V.x, cb.V.y, cb.V.z])
    axis = -np.ones(3) / math.sqrt(3.0)
    assert abs(math.degrees(math.acos(va @ axis)) - 20.0) < 1e-3
    assert abs(math.degrees(math.acos(vb @ axis)) - 20.0) < 1e-3
    assert abs(math.degrees(math.acos(va @ vb)) - 40.0) < 1e-3


def test_ppm_format(tmp_path):
    rgba = np.array([[[255, 0, 7, 9], [1, 22, 133, 0], [0, 0, 0, 0]],
                     [[10, 20, 30, 40], [99, 100, 101, 255], [5, 6, 7, 8]]], dtype=np.uint8)
    path = str(tmp_path / "x.ppm")
    lyap.write_ppm(path, rgba, 3, 2)
    want = "P3\n3 2\n255\n255   0   7   1  22 133   0   0   0  10  20  30  99 100 101   5   6   7 \n"
    assert open(path).read() == want


def test_raw_formats(tmp_path):
    v = np.arange(4 * 3 * 2, dtype=np.float32).reshape(2, 3, 4) * 0.5 - 3.0
    p = str(tmp_path / "exps.raw")
    lyap.write_volume_raw(p, v)
    assert os.path.getsize(p) == v.nbytes
    assert np.array_equal(np.fromfile(p, dtype=np.float32).reshape(2, 3, 4), v)
    pts = (lyap.LyapPoint * 3)()
    for i in range(3):
        pts[i].P.x, pts[i].N.z, pts[i].a, pts[i].c, pts[i].l = i, -i, 0.5 * i, 2.0, float("nan")
    raw = np.frombuffer(bytes(pts), dtype=np.uint8)
    q = str(tmp_path / "Points.raw")
    lyap.write_points_raw(q, raw, 3)
    assert open(q, "rb").read() == bytes(pts)


def test_output_names():
    prm, cam, _, _ = lyap.params_init()
    cam.C.x, cam.C.y, cam.C.z = 4.125, 4.125, -0.3
    prm.depth = 8192.0
    leaf = lyap.output_leaf("Render", 1526604923, 3840, 2160, "BCABA", cam, prm)
    # field formats as in the reference's own file name (scale.pl:66)
    assert leaf == ("Render_1526604923_3840x2160_BCABA_cx=4.125_cy=4.125_cz=-0.30000001"
                    "_step=2_D=2.1_i=18,1008_d=8192_j=0.5_r=32_ot=-0.75")
    assert lyap.output_name(leaf, 44, 4, 2, "ppm") == leaf + "_time=44h04m02s.ppm"
    pleaf = lyap.output_leaf("Points", 7, 16, 8, "AB", cam, prm)
    assert pleaf.startswith("Points_7_16x8_AB_cx=")
    with pytest.raises(lyap.LyapError):
        lyap.output_leaf("Render", 1, 1, 1, "A" * 600, cam, prm)
