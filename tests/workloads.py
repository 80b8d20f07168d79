"""The seeded workloads the tests and the bench share.

Each function returns plain ctypes structs (camera, parameters, lights) built
with the library's host scene helpers; nothing here touches a GPU.
"""
from __future__ import annotations

import math

import numpy as np

import lyapunov3d_amd as lyap

CENTER = (3.0, 3.0, 3.0)


def volume_config(name: str):
    """(params, sequence, lo, hi, n) of the volume workloads."""
    if name == "c1":     # 64^3 over [2,4]^3, ABBAC, 100 + 400
        prm = lyap.make_params(settle=100, accum=400)
        return prm, "ABBAC", (2.0, 2.0, 2.0), (4.0, 4.0, 4.0), 64
    if name == "c3":     # 256^3 over [2.5,4]^3, ABCABCCA, 200 + 2000
        prm = lyap.make_params(settle=200, accum=2000)
        return prm, "ABCABCCA", (2.5, 2.5, 2.5), (4.0, 4.0, 4.0), 256
    raise KeyError(name)


C3_MODE = dict(precision=lyap.PREC_PRODUCT, fp64_margin=0.01)
C3_STRICT = dict(precision=lyap.PREC_PRODUCT, fp64_margin=0.01,
                 converge_tau=1e-4, converge_band=0.1)
# the same test up to min(half) < 0.2: also re-evaluates weakly chaotic
# voxels, which FP32 rounding can show as chaotic where FP64 settles
C3_STRICT_WIDE = dict(precision=lyap.PREC_PRODUCT, fp64_margin=0.01,
                      converge_tau=1e-4, converge_band=0.2)


def seeded_quat(seed: int):
    """Unit quaternion from a normal 4-vector, normalised in FP64."""
    g = np.random.default_rng(seed).normal(size=4)
    g = g / np.linalg.norm(g)
    return tuple(float(v) for v in g)


def render_config(w: int = 512, h: int = 512, seed: int = 1, seq: str = "AABAB",
                  settle: int = 100, accum: int = 300, depth: float = 500.0):
    """The plain render workload: camera 4 from the [2,4]^3 centre along a
    seeded quaternion's +z axis, 45 degree field of view, uniform steps
    (stepMethod 1, depth 500), opaque at lambda <= 0, default spotlight."""
    prm, cam, lights, n_lights = lyap.params_init()
    prm.settle = settle
    prm.accum = accum
    prm.stepMethod = 1
    prm.depth = depth
    prm.jitter = 0.0
    prm.nearMultiplier = 1.0
    prm.opaqueThreshold = 0.0
    prm.lMin = 2.0
    prm.lMax = 4.0
    qx, qy, qz, qw = seeded_quat(seed)
    cam.Q.x, cam.Q.y, cam.Q.z, cam.Q.w = qx, qy, qz, qw
    fwd = lyap.quat_transform(cam.Q, (0.0, 0.0, 1.0))
    # C = centre - 4 * forward, formed in double from the float forward axis
    cam.C.x = CENTER[0] - 4.0 * fwd[0]
    cam.C.y = CENTER[1] - 4.0 * fwd[1]
    cam.C.z = CENTER[2] - 4.0 * fwd[2]
    cam.M = math.tan(math.radians(22.5))
    lyap.cam_recalculate(cam, w, h, 1)
    lyap.lights_recalculate(lights, n_lights)
    return cam, prm, seq, lights, n_lights


def c2_mode(precision=lyap.PREC_PRODUCT, fp64_margin=0.01):
    return lyap.render_mode(refine=lyap.REFINE_BISECT, bisect_steps=8, normal_h=1e-3,
                            precision=precision, fp64_margin=fp64_margin)


def random_sequence(rng, lo=3, hi=12, letters="ABC") -> str:
    n = int(rng.integers(lo, hi + 1))
    return "".join(letters[int(i)] for i in rng.integers(0, len(letters), n))


def batch_views(n_views: int = 64, seed: int = 4, sides=(128, 256, 384, 512, 768, 1024),
                accum_range=(200, 2000), settle: int = 100):
    """The batched workload: views of seeded side, camera on a sphere of
    radius U[3,6] around the centre looking at it, sequence of 3-12 letters
    over A-C, accumulation U[accum_range]. Returns a list of dicts."""
    rng = np.random.default_rng(seed)
    prm0, _, lights, n_lights = lyap.params_init()
    lyap.lights_recalculate(lights, n_lights)
    views = []
    for _ in range(n_views):
        side = int(rng.choice(sides))
        radius = float(rng.uniform(3.0, 6.0))
        d = rng.normal(size=3)
        d = d / np.linalg.norm(d)
        eye = tuple(float(CENTER[k] + radius * d[k]) for k in range(3))
        cam = lyap.cam_look_at(eye, CENTER, math.tan(math.radians(22.5)), side, side)
        prm = lyap.make_params(settle=settle, accum=int(rng.integers(accum_range[0], accum_range[1] + 1)),
                               stepMethod=1, depth=500.0, jitter=0.0, nearMultiplier=1.0,
                               opaqueThreshold=0.0, lMin=2.0, lMax=4.0)
        views.append(dict(cam=cam, prm=prm, seq=random_sequence(rng), w=side, h=side))
    return views, lights, n_lights


def serving_stream(seconds: float = 60.0, rate: float = 200.0, seed: int = 5, settle: int = 100):
    """The serving workload: Poisson arrivals at `rate` per second for
    `seconds`. 70% are views (side U{256,512,1024}, camera and sequence as in
    batch_views), 30% volume blocks (edge U{32,64,128,256} over a random cube
    inside [2,4]^3 with side U[0.25,2]); accumulation U{200..4000}, settle
    100. Returns (requests, lights, n_lights); each request is a dict with
    its arrival time in ms."""
    rng = np.random.default_rng(seed)
    _, _, lights, n_lights = lyap.params_init()
    lyap.lights_recalculate(lights, n_lights)
    out = []
    t = 0.0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= seconds:
            break
        accum = int(rng.integers(200, 4001))
        if rng.random() < 0.7:
            side = int(rng.choice((256, 512, 1024)))
            radius = float(rng.uniform(3.0, 6.0))
            d = rng.normal(size=3)
            d = d / np.linalg.norm(d)
            eye = tuple(float(CENTER[k] + radius * d[k]) for k in range(3))
            cam = lyap.cam_look_at(eye, CENTER, math.tan(math.radians(22.5)), side, side)
            prm = lyap.make_params(settle=settle, accum=accum, stepMethod=1, depth=500.0, jitter=0.0,
                                   nearMultiplier=1.0, opaqueThreshold=0.0, lMin=2.0, lMax=4.0)
            out.append(dict(kind="view", t_ms=1000.0 * t, cam=cam, prm=prm, seq=random_sequence(rng),
                            w=side, h=side))
        else:
            edge = int(rng.choice((32, 64, 128, 256)))
            s = float(rng.uniform(0.25, 2.0))
            lo = tuple(float(2.0 + rng.uniform(0.0, 2.0 - s)) for _ in range(3))
            hi = tuple(x + s for x in lo)
            prm = lyap.make_params(settle=settle, accum=accum)
            out.append(dict(kind="volume", t_ms=1000.0 * t, prm=prm, seq=random_sequence(rng),
                            lo=lo, hi=hi, n=edge))
    return out, lights, n_lights
