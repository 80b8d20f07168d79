"""Fly-throughs and D-letter slices as batches, on the GPU.

 - the scripted camera path (scale.pl) rendered as one batch gives, frame by
   frame, the bytes of rendering each frame alone, and its FP64 frames are
   the oracle's exactly;
 - a sweep of the fourth coordinate d for a sequence with the letter D is one
   batch too; d changes lambda only through D (a sequence without D gives the
   same bytes for every d), and the FP64 slices are the oracle's.
"""
import numpy as np
import pytest

import lyapunov3d_amd as lyap
from oracle import oracle as O
from tests import workloads as W
from tests.gpu_util import RenderBuffers, device_lights, torch_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    torch_cuda()
    c = lyap.Context(0)
    yield c
    c.close()


def scene(side, accum=120, settle=30, d=2.1):
    prm, _, lights, n = lyap.params_init()
    prm.settle, prm.accum, prm.d = settle, accum, d
    prm.stepMethod, prm.depth, prm.jitter, prm.nearMultiplier = 1, 200.0, 0.0, 1.0
    prm.opaqueThreshold, prm.lMin, prm.lMax = 0.0, 2.0, 4.0
    lyap.lights_recalculate(lights, n)
    return prm, lights, n


def same(a, b):
    for k in ("rgba", "depth", "lambda", "evals", "points"):
        assert np.array_equal(a[k], b[k], equal_nan=True), k


def batch_and_single(ctx, frames, ld, n, mode):
    bufs = [RenderBuffers(f["w"], f["h"]) for f in frames]
    api = [lyap.make_view(f["cam"], f["prm"], f["seq"], f["w"], f["h"], b.rgba.data_ptr(), b.points.data_ptr(),
                          b.depth.data_ptr(), b.lam.data_ptr(), b.evals.data_ptr(), ld.data_ptr(), n)
           for f, b in zip(frames, bufs)]
    ctx.render_batch(api, mode)
    ctx.sync()
    got = [b.fetch() for b in bufs]
    for f, g in zip(frames, got):
        one = RenderBuffers(f["w"], f["h"])
        one.render(ctx, f["cam"], f["prm"], f["seq"], ld, n, mode)
        ctx.sync()
        same(g, one.fetch())
    return got


@pytest.mark.parametrize("prec", [lyap.PREC_PRODUCT, lyap.PREC_FP64])
def test_flythrough_batch(ctx, prec):
    side = 40
    prm, lights, n = scene(side)
    ld = device_lights(lights, n)
    ts = lyap.flythrough_times(0.01)[::20]            # 6 of the 101 frames
    frames = []
    for t in ts:
        cam = lyap.flythrough_camera(lyap.ease_in_out_quart(t), side, side)
        frames.append(dict(cam=cam, prm=prm, seq="BCABA", w=side, h=side))
    mode = W.c2_mode(precision=prec, fp64_margin=0.0)
    got = batch_and_single(ctx, frames, ld, n, mode)
    assert np.isfinite(got[len(got) // 2]["depth"]).mean() > 0.3   # the box is in view
    if prec == lyap.PREC_FP64:
        for i in (0, 3):
            f = frames[i]
            ref = O.render(f["cam"], f["prm"], f["seq"], lights, n, side, side, mode=mode, lam=O.F64)
            assert np.array_equal(got[i]["evals"], ref["evals"])
            assert np.array_equal(got[i]["depth"], ref["depth"], equal_nan=True)


def test_d_sweep(ctx):
    side = 32
    cam, _, _, _, _ = W.render_config(side, side)
    ds = [2.0, 2.5, 3.0, 3.5, 3.9]
    frames, plain = [], []
    for d in ds:
        prm, lights, n = scene(side, d=d)
        frames.append(dict(cam=cam, prm=prm, seq="ABCD", w=side, h=side))
        plain.append(dict(cam=cam, prm=prm, seq="ABCA", w=side, h=side))
    ld = device_lights(lights, n)
    mode = W.c2_mode(precision=lyap.PREC_FP64, fp64_margin=0.0)
    got = batch_and_single(ctx, frames, ld, n, mode)
    # the slices differ (d enters through D) ...
    assert not np.array_equal(got[0]["lambda"], got[-1]["lambda"], equal_nan=True)
    for i in (0, 4):
        f = frames[i]
        ref = O.render(f["cam"], f["prm"], f["seq"], lights, n, side, side, mode=mode, lam=O.F64)
        assert np.array_equal(got[i]["depth"], ref["depth"], equal_nan=True)
        assert np.array_equal(got[i]["evals"], ref["evals"])
    # ... and without D, d is not read at all
    flat = batch_and_single(ctx, plain, ld, n, mode)
    for g in flat[1:]:
        same(g, flat[0])
