"""Host scene helpers of liblyap_amd.so (params.cu, scene.cu, quat.hpp), CPU only."""
import math

import numpy as np
import pytest

import lyapunov3d_amd as lyap


@pytest.mark.parametrize("text,expect", [
    ("BCABA", [1, 2, 0, 1, 0, -1]),                  # params.cu default
    ("ABBAC", [0, 1, 1, 0, 2, -1]),
    ("abc", [0, 1, 2, -1]),                          # lower case
    ("A2B", [0, 0, 0, 1, -1]),                       # digit n: n more copies
    ("2A", [1, 1, 0, -1]),                           # B precedes the first letter
    ("D1", [3, 3, -1]),
    ("A6B6C6D6", [0] * 7 + [1] * 7 + [2] * 7 + [3] * 7 + [-1]),
])
def test_convert_sequence(text, expect):
    assert lyap.convert_sequence(text) == expect


@pytest.mark.parametrize("text", ["ABE", "A0", "", "A B", "x"])
def test_bad_sequence(text):
    with pytest.raises(lyap.LyapError) as e:
        lyap.convert_sequence(text)
    assert e.value.status == lyap.ERR_BAD_SEQUENCE


def test_params_defaults():
    """params.cu:23-36 and the default light of params.cu:96-109."""
    prm, cam, lights, n = lyap.params_init()
    f = np.float32
    assert (prm.d, prm.settle, prm.accum, prm.stepMethod) == (f(2.1), 18, 1008, 2)
    assert (prm.nearThreshold, prm.nearMultiplier, prm.opaqueThreshold, prm.chaosThreshold) == (-1.0, 2.0, -0.75, -0.5)
    assert (prm.depth, prm.jitter, prm.refine, prm.gradient, prm.lMin, prm.lMax) == (4096.0, 0.5, 32.0, f(0.01), 0.0, 4.0)
    assert n == 1 and lyap.default_sequence() == "BCABA"
    L = lights[0]
    assert (L.C.x, L.C.y, L.C.z) == (6.0, 5.0, 3.0)
    assert (L.diffuseColor.x, L.diffuseColor.y, L.diffuseColor.z) == (1.0, 0.25, 0.125)
    assert cam.M == f(0.45)
    # scripted camera: one unit behind (3.1,3.1,3.1) along a unit view axis
    C = np.array([cam.C.x, cam.C.y, cam.C.z])
    fwd = np.array(lyap.quat_transform(cam.Q, (0, 0, 1)))
    assert np.linalg.norm(C - 3.1) == pytest.approx(1.0, abs=1e-5)
    assert np.allclose(C + fwd, 3.1, atol=1e-5)        # looks at the scripted point


def test_cam_recalculate_identity():
    """scene.cu:31-63 with Q = 1: V = z, S0 = (-M,-M,1), SDX = (2M/w,0,0)."""
    cam = lyap.LyapCam()
    cam.Q.w = 1.0
    cam.M = 0.5
    lyap.cam_recalculate(cam, 4, 8, 1)
    assert (cam.V.x, cam.V.y, cam.V.z) == (0.0, 0.0, 1.0)
    assert (cam.S0.x, cam.S0.y, cam.S0.z) == (-0.5, -0.5, 1.0)
    assert (cam.SDX.x, cam.SDX.y, cam.SDX.z) == (0.25, 0.0, 0.0)
    assert (cam.SDY.x, cam.SDY.y, cam.SDY.z) == (0.0, 0.125, 0.0)
    assert (cam.renderWidth, cam.renderHeight, cam.renderDenominator) == (4, 8, 1)
    # cone cosines: angle between z and (-M,-M,1.5) / (-M,-M,1)
    assert cam.lightInnerCone == pytest.approx(1.5 / math.sqrt(0.5 + 2.25), rel=1e-6)
    assert cam.lightOuterCone == pytest.approx(1.0 / math.sqrt(1.5), rel=1e-6)


def test_cam_recalculate_rotated_and_clamped():
    cam = lyap.LyapCam()
    q = lyap.quat_axis_angle((0, 0, 1), 90.0, True)      # z-rotation by 90 degrees
    cam.Q = q
    cam.M = 0.0                                          # clamped to 1e-6
    cam.renderDenominator = 2
    lyap.cam_recalculate(cam, 8, 8, 0)
    assert cam.M == np.float32(1e-6) and cam.renderWidth == 4
    cam.M = 0.5
    lyap.cam_recalculate(cam, 8, 8, 1)
    assert np.allclose([cam.SDX.x, cam.SDX.y, cam.SDX.z], [0.0, 0.125, 0.0], atol=1e-7)
    assert np.allclose([cam.S0.x, cam.S0.y, cam.S0.z], [0.5, -0.5, 1.0], atol=1e-7)


def test_cam_normalise_quirk():
    """QUAT::normalize (quat.hpp:147-165) multiplies by |q| outside the 1e-6
    band around |q|^2 = 1: q = (0,0,0,2) becomes (0,0,0,4); the view axis,
    normalised afterwards, is still +z."""
    cam = lyap.LyapCam()
    cam.Q.x, cam.Q.y, cam.Q.z, cam.Q.w = 0.0, 0.0, 0.0, 2.0
    cam.M = 0.5
    lyap.cam_recalculate(cam, 4, 4, 1)
    assert (cam.Q.w, cam.V.z) == (4.0, 1.0)


def vec(v):
    return (v.x, v.y, v.z)


def test_transform_sums_left_to_right_in_float():
    """QUAT::transform (quat.hpp:279-284) adds its eight float products in
    the written order. q = (0, 2^-12, 0, 1), v = (1, 0, 2^-13): the x
    component is 1 + 2 y w v_z - y^2 v_x = 1 + 2^-24 - 2^-24. In float the
    first sum 1 + 2^-24 rounds (to even) to 1 and the last term then gives
    1 - 2^-24; evaluated exactly (or in double, rounded once) it is 1."""
    q = lyap.Quat(0.0, 2.0 ** -12, 0.0, 1.0)
    out = lyap.quat_transform(q, (1.0, 0.0, 2.0 ** -13))
    assert out[0] == 1.0 - 2.0 ** -24


@pytest.mark.parametrize("q,axes", [((0.0, 0.0, 0.0, 1.0), (0, 1, 2)), ((0.5, 0.5, 0.5, 0.5), (2, 0, 1))])
def test_camera_vectors_bitwise(q, axes):
    """scene_cam_recalculate (scene.cu:31-63) on rotations whose products
    are exact: the identity, and (1/2, 1/2, 1/2, 1/2), the 120 degree turn
    about (1,1,1) that maps (a, b, c) to (c, a, b). M = 1/2 and a 4 x 4 image:
    S0 = Q (-1/2, -1/2, 1), SDX = Q (1/4, 0, 0), SDY = Q (0, 1/4, 0), V = Q z,
    bit for bit."""
    cam = lyap.LyapCam()
    cam.Q.x, cam.Q.y, cam.Q.z, cam.Q.w = q
    cam.M = 0.5
    lyap.cam_recalculate(cam, 4, 4, 1)

    def rot(v):
        return tuple(v[a] for a in axes)

    assert vec(cam.S0) == rot((-0.5, -0.5, 1.0))
    assert vec(cam.SDX) == rot((0.25, 0.0, 0.0))
    assert vec(cam.SDY) == rot((0.0, 0.25, 0.0))
    assert vec(cam.V) == rot((0.0, 0.0, 1.0))


def test_quaternions():
    q = lyap.quat_axis_angle((1, 0, 0), 180.0, True)
    assert np.allclose(lyap.quat_transform(q, (0, 1, 0)), (0, -1, 0), atol=1e-6)
    q = lyap.quat_axis_angle((0, 1, 0), math.pi / 2, False)
    assert np.allclose(lyap.quat_transform(q, (0, 0, 1)), (1, 0, 0), atol=1e-6)
    v = (0.6, 0.0, 0.8)
    q = lyap.quat_between((0, 0, 1), v, 1.0)
    assert np.allclose(lyap.quat_transform(q, (0, 0, 1)), v, atol=1e-6)
    # half the angle with scale 0.5
    q = lyap.quat_between((0, 0, 1), (1, 0, 0.0001), 0.5)
    r = lyap.quat_transform(q, (0, 0, 1))
    assert np.allclose(r, (math.sqrt(0.5), 0, math.sqrt(0.5)), atol=1e-3)
    # the reference returns the identity for perpendicular vectors (quat.hpp:204-208)
    q = lyap.quat_between((0, 0, 1), (1, 0, 0), 1.0)
    assert (q.x, q.y, q.z, q.w) == (0.0, 0.0, 0.0, 1.0)


@pytest.mark.parametrize("eye", [(7, 3, 3), (3, -2, 3), (1, 1, 8), (3, 3, -3), (3, 3, 9)])
def test_look_at(eye):
    cam = lyap.cam_look_at(eye, (3, 3, 3), 0.4, 16, 16)
    d = np.array([3, 3, 3], float) - np.array(eye, float)
    assert np.allclose([cam.V.x, cam.V.y, cam.V.z], d / np.linalg.norm(d), atol=1e-6)
    # centre ray through pixel corner (8,8) points at the target
    V = np.array([cam.S0.x, cam.S0.y, cam.S0.z]) + 8 * np.array([cam.SDX.x, cam.SDX.y, cam.SDX.z]) + \
        8 * np.array([cam.SDY.x, cam.SDY.y, cam.SDY.z])
    assert np.allclose(V / np.linalg.norm(V), d / np.linalg.norm(d), atol=1e-5)


def test_lights_recalculate():
    _, _, lights, n = lyap.params_init()
    lyap.lights_recalculate(lights, n)
    L = lights[0]
    V = np.array(lyap.quat_transform(L.Q, (0, 0, 1)))
    assert np.allclose([L.V.x, L.V.y, L.V.z], V / np.linalg.norm(V), atol=1e-6)
    inner = np.array(lyap.quat_transform(L.Q, (-0.5, -0.5, 1.5)))
    assert L.lightInnerCone == pytest.approx(float(V @ inner / np.linalg.norm(inner) / np.linalg.norm(V)), abs=1e-6)
    assert L.lightInnerCone > L.lightOuterCone
