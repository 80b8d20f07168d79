"""The render's diagnostics on the GPU: the per-wave log and the launch timer.

The per-wave log doubles as a check of the persistent kernel's bookkeeping:
the rays its waves started in the first launch are exactly the pixels whose
ray meets the box (those, and only those, take at least one evaluation), each
later launch starts a subset of the rays the previous one did, and every
wave's clock split stays within its own run time.
"""
import os

import numpy as np
import pytest

import lyapunov3d_amd as lyap
from tests import workloads as W
from tests.gpu_util import RenderBuffers, device_lights, empty, to_numpy, torch_cuda

pytestmark = pytest.mark.gpu

WORDS, SLOTS, LEVELS = 16, 8192, 4


@pytest.fixture(scope="module")
def ctx():
    torch_cuda()
    c = lyap.Context(0)
    yield c
    c.close()


def logged_render(ctx, w, h, mode, seq=None):
    torch = torch_cuda()
    cam, prm, s, lights, n_lights = W.render_config(w, h)
    ld = device_lights(lights, n_lights)
    bufs = RenderBuffers(w, h)
    log = empty(8 * WORDS * SLOTS * LEVELS)
    log.zero_()
    torch.cuda.synchronize()
    assert ctx.lib.lyap_debug_wave_log(ctx.h, log.data_ptr()) == lyap.OK
    try:
        bufs.render(ctx, cam, prm, seq or s, ld, n_lights, mode)
        ctx.sync()
    finally:
        assert ctx.lib.lyap_debug_wave_log(ctx.h, None) == lyap.OK
    rec = to_numpy(log, np.uint64).reshape(LEVELS, SLOTS, WORDS).astype(np.int64)
    return bufs.fetch(), rec


@pytest.mark.parametrize("precision", [lyap.PREC_PRODUCT, lyap.PREC_PLAIN])
def test_wave_log_accounts_for_every_ray(ctx, precision):
    out, rec = logged_render(ctx, 192, 160, W.c2_mode(precision=precision, fp64_margin=0.0))
    evals = out["evals"]
    started = [int(rec[lev][:, 1].sum()) for lev in range(LEVELS)]
    # level 0 starts the rays that meet the box: the pixels with evaluations
    assert started[0] == int(np.count_nonzero(evals))
    assert started[0] > 0
    # a later launch resumes some of the rays the previous one handed over
    assert started[0] >= started[1] >= started[2] >= started[3] >= 0
    for lev in range(LEVELS):
        r = rec[lev]
        ran = r[r[:, 3] > 0]
        if len(ran) == 0:
            continue
        dur = ran[:, 3] - ran[:, 2]
        assert (dur >= 0).all()
        parts = ran[:, 4] + ran[:, 5] + ran[:, 6]
        assert (parts <= dur).all(), lev                    # eval + refill + combine
        assert (ran[:, 7] <= ran[:, 5]).all()               # tile claims are refill time
        assert (ran[:, 9] <= ran[:, 5]).all()               # ray starts are refill time
        assert (ran[:, 8] <= ran[:, 6]).all()               # hand-overs are combine time
        # the constant-rate clock brackets the wave too
        assert (ran[:, 11] >= ran[:, 10]).all() and (ran[:, 10] > 0).all()
        assert (ran[:, 12:] == 0).all()
        # a wave that started rays evaluated them
        assert (ran[ran[:, 1] > 0, 0] > 0).all()


def test_wave_log_off_leaves_buffer_untouched(ctx):
    torch = torch_cuda()
    cam, prm, seq, lights, n_lights = W.render_config(64, 64)
    ld = device_lights(lights, n_lights)
    bufs = RenderBuffers(64, 64)
    log = empty(8 * WORDS * SLOTS * LEVELS)
    log.fill_(0xAB)
    torch.cuda.synchronize()
    bufs.render(ctx, cam, prm, seq, ld, n_lights, W.c2_mode(fp64_margin=0.0))
    ctx.sync()
    assert bool((log == 0xAB).all())


def test_timing_counts_the_level_launches(ctx):
    cam, prm, seq, lights, n_lights = W.render_config(128, 128)
    ld = device_lights(lights, n_lights)
    bufs = RenderBuffers(128, 128)
    mode = W.c2_mode(fp64_margin=0.0)
    # the persistent schedule: four march levels and the refinement launch
    lv = W.c2_mode(fp64_margin=0.0)
    lv.schedule = lyap.SCHEDULE_LEVELS
    ctx.timing(True)
    try:
        bufs.render(ctx, cam, prm, seq, ld, n_lights, lv)
        ctx.sync()
        assert ctx.timing_collect()[1] == 5
    finally:
        ctx.timing(False)
    # a small frame marches in passes: the set-up launch, seven passes and
    # the refinement launch
    ctx.timing(True)
    try:
        for _ in range(3):
            bufs.render(ctx, cam, prm, seq, ld, n_lights, mode)
        ctx.sync()
        ms, launches = ctx.timing_collect()
        per_frame = 5 if os.environ.get("LYAP_MARCH_PASSES") == "0" else 9
        assert launches == 3 * per_frame and ms > 0.0
        # collecting resets the sums
        ms2, launches2 = ctx.timing_collect()
        assert launches2 == 0 and ms2 == 0.0
    finally:
        ctx.timing(False)
    # switching off pauses: a sample of frames can be timed
    ctx.timing_collect()
    try:
        for on in (True, False, True):
            ctx.timing(on)
            bufs.render(ctx, cam, prm, seq, ld, n_lights, mode)
        ctx.timing(False)
        bufs.render(ctx, cam, prm, seq, ld, n_lights, mode)
        ctx.sync()
        assert ctx.timing_collect()[1] == 2 * per_frame
    finally:
        ctx.timing(False)
    # direct schedule: one launch per frame
    d = W.c2_mode(fp64_margin=0.0)
    d.schedule = lyap.SCHEDULE_DIRECT
    ctx.timing(True)
    try:
        bufs.render(ctx, cam, prm, seq, ld, n_lights, d)
        ctx.sync()
        assert ctx.timing_collect()[1] == 1
    finally:
        ctx.timing(False)
