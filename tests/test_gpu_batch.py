"""Batched views, the autotuner and captured frames on the GPU.

 - a batch of views with different sizes, sequences (periods 3..12 and one
   of 20), iteration counts and cameras gives, view by view, exactly the
   bytes of rendering each view alone (scheduling never changes a ray's
   arithmetic);
 - FP64 batches reproduce the oracle exactly for sampled views;
 - a captured frame replayed several times, with other work and
   synchronisations between replays, matches the eager batch every time;
 - render calls issued back to back, without synchronising, keep their own
   view tables;
 - every launch shape (lanes per ray, samples per lane, tile length) gives
   the same bytes.
"""
import ctypes as C
import os

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


def small_batch(n=10, seed=7, sides=(24, 40, 64, 96, 17), accum=(60, 200)):
    views, lights, n_lights = W.batch_views(n, seed=seed, sides=sides, accum_range=accum, settle=40)
    if n > 3:
        views[3]["seq"] = "ABCABCABCABCABCABCAB"   # period 20: the 32-wide window
    return views, lights, n_lights


def alloc(views):
    return [RenderBuffers(v["w"], v["h"]) for v in views]


def as_api(views, bufs, ld, n_lights):
    out = []
    for v, b in zip(views, bufs):
        out.append(lyap.make_view(v["cam"], v["prm"], v["seq"], v["w"], v["h"], b.rgba.data_ptr(),
                                  b.points.data_ptr(), b.depth.data_ptr(), b.lam.data_ptr(),
                                  b.evals.data_ptr(), ld.data_ptr(), n_lights))
    return out


def same(a, b):
    for k in ("rgba", "depth", "lambda", "evals", "points"):
        assert np.array_equal(a[k], b[k], equal_nan=True), k


@pytest.mark.parametrize("prec", [lyap.PREC_PRODUCT, lyap.PREC_FP64])
def test_batch_equals_single_views(ctx, prec):
    views, lights, n_lights = small_batch()
    ld = device_lights(lights, n_lights)
    mode = W.c2_mode(precision=prec, fp64_margin=0.0)
    bufs = alloc(views)
    ctx.render_batch(as_api(views, bufs, ld, n_lights), mode)
    ctx.sync()
    batch_evals = ctx.last_evals()
    got = [b.fetch() for b in bufs]
    total = 0
    for v, g in zip(views, got):
        one = RenderBuffers(v["w"], v["h"])
        one.render(ctx, v["cam"], v["prm"], v["seq"], ld, n_lights, mode)
        ctx.sync()
        same(g, one.fetch())
        total += ctx.last_evals()
    assert batch_evals == total
    if prec == lyap.PREC_FP64:
        for i in (0, 3, 5):
            v = views[i]
            ref = O.render(v["cam"], v["prm"], v["seq"], lights, n_lights, v["w"], v["h"], mode=mode, lam=O.F64)
            assert np.array_equal(got[i]["evals"], ref["evals"])
            assert np.array_equal(got[i]["depth"], ref["depth"], equal_nan=True)


def test_back_to_back_calls_keep_their_tables(ctx):
    """Render calls do not wait for the previous frame: each stages its view
    table in one of two pinned buffers used in turn. Six different views
    rendered with no synchronisation between the calls (so the third call
    reuses the first buffer while frames may still be in flight) give the
    bytes of rendering each alone."""
    views, lights, n_lights = small_batch(6, seed=23)
    ld = device_lights(lights, n_lights)
    mode = W.c2_mode(fp64_margin=0.0)
    bufs = alloc(views)
    for v, b in zip(views, bufs):
        b.render(ctx, v["cam"], v["prm"], v["seq"], ld, n_lights, mode)
    ctx.sync()
    got = [b.fetch() for b in bufs]
    for v, g in zip(views, got):
        one = RenderBuffers(v["w"], v["h"])
        one.render(ctx, v["cam"], v["prm"], v["seq"], ld, n_lights, mode)
        ctx.sync()
        same(g, one.fetch())


def test_frame_replays_match_eager(ctx):
    torch = torch_cuda()
    views, lights, n_lights = small_batch(seed=11)
    ld = device_lights(lights, n_lights)
    mode = W.c2_mode(fp64_margin=0.0)
    eager_bufs = alloc(views)
    ctx.render_batch(as_api(views, eager_bufs, ld, n_lights), mode)
    ctx.sync()
    eager = [b.fetch() for b in eager_bufs]
    bufs = alloc(views)
    frame = ctx.frame_capture(as_api(views, bufs, ld, n_lights), mode)
    for rep in range(4):
        for b in bufs:                       # scrub the outputs between replays
            b.rgba.zero_(); b.depth.fill_(7.0); b.evals.zero_()
        torch.cuda.synchronize()
        frame.replay()
        ctx.sync()
        # unrelated eager work on the same context between replays
        other = RenderBuffers(16, 16)
        other.render(ctx, views[0]["cam"], views[0]["prm"], "AB", ld, n_lights, mode)
        ctx.sync()
        for b, e in zip(bufs, eager):
            same(b.fetch(), e)
    frame.close()


def test_autotune_keeps_results(ctx):
    views, lights, n_lights = small_batch(n=6, seed=3)
    ld = device_lights(lights, n_lights)
    mode = W.c2_mode(fp64_margin=0.0)
    bufs = alloc(views)
    api = as_api(views, bufs, ld, n_lights)
    ctx.render_batch(api, mode)
    ctx.sync()
    before = [b.fetch() for b in bufs]
    best = ctx.autotune(api, mode)
    # the built-in shape (zeros) or one of the candidates, one lane per ray included
    assert best.lanes_per_ray in (0, 1, 2, 4, 8, 16) and best.slots in (0, 1, 2, 4)
    assert all(b in (0, 16, 24, 32, 48, 64, 96, 128, 192, 256) for b in best.level_budget)
    ctx.render_batch(api, mode)
    ctx.sync()
    for b, e in zip(bufs, before):
        same(b.fetch(), e)
    ctx.set_tuning(None)


@pytest.mark.parametrize("margin", [0.0, 0.01])
@pytest.mark.parametrize("lanes,slots,tile,n", [(1, 1, 128, 3), (1, 1, 200, 3), (4, 1, 32, 3),
                                                (8, 2, 8, 3), (16, 1, 4, 3), (2, 4, 64, 3),
                                                (1, 1, 200, 5)])
def test_launch_shapes_same_bytes(ctx, lanes, slots, tile, n, margin):
    """Lanes per ray, samples per lane and tile length change only the
    schedule: a tile longer than a wave is set up in chunks of 64 rays, a
    short one leaves groups to take rays from the next claim. Five views
    include the 20-letter sequence (the 32-slot window). With an FP64
    margin, rays refined in place (one lane per ray) and rays parked for the
    refinement launch (two or more) re-evaluate the same bisection samples
    in FP64, so the bytes and the evaluation count agree as well."""
    views, lights, n_lights = small_batch(n=n, seed=11)
    ld = device_lights(lights, n_lights)
    mode = W.c2_mode(fp64_margin=margin)
    bufs = alloc(views)
    api = as_api(views, bufs, ld, n_lights)
    ctx.render_batch(api, mode)
    ctx.sync()
    ref_evals = ctx.last_evals()
    ref = [b.fetch() for b in bufs]
    t = lyap.Tuning()
    t.lanes_per_ray, t.slots, t.tile_rays = lanes, slots, tile
    ctx.render_batch(api, mode, t)
    ctx.sync()
    assert ctx.last_evals() == ref_evals
    for b, e in zip(bufs, ref):
        same(b.fetch(), e)


def test_batch_errors(ctx):
    views, lights, n_lights = small_batch(n=2)
    ld = device_lights(lights, n_lights)
    bufs = alloc(views)
    api = as_api(views, bufs, ld, n_lights)
    api[1].seq[0] = 9
    with pytest.raises(lyap.LyapError) as e:
        ctx.render_batch(api)
    assert e.value.status == lyap.ERR_BAD_SEQUENCE
    api = as_api(views, bufs, ld, n_lights)
    api[0].rgba = None
    with pytest.raises(lyap.LyapError) as e:
        ctx.render_batch(api)
    assert e.value.status == lyap.ERR_INVALID_ARGUMENT


@pytest.mark.parametrize("prec,margin", [(lyap.PREC_PLAIN, 0.0), (lyap.PREC_PRODUCT, 0.0),
                                         (lyap.PREC_FP64, 0.0), (lyap.PREC_PRODUCT, 0.01)])
def test_direct_schedule_same_bytes(ctx, prec, margin):
    """The reference's launch shape (one lane per ray, no refill, refinement
    in place) and the batched schedule (rays parked for the refinement
    launch) compute every ray identically, with and without FP64
    re-evaluation of bisection samples near the threshold."""
    views, lights, n_lights = small_batch(6, seed=3)
    ld = device_lights(lights, n_lights)
    batched = W.c2_mode(precision=prec, fp64_margin=margin)
    direct = W.c2_mode(precision=prec, fp64_margin=margin)
    direct.schedule = lyap.SCHEDULE_DIRECT
    for v in views:
        a = RenderBuffers(v["w"], v["h"])
        b = RenderBuffers(v["w"], v["h"])
        a.render(ctx, v["cam"], v["prm"], v["seq"], ld, n_lights, batched)
        ctx.sync()
        ea = ctx.last_evals()
        b.render(ctx, v["cam"], v["prm"], v["seq"], ld, n_lights, direct)
        ctx.sync()
        assert ctx.last_evals() == ea
        same(a.fetch(), b.fetch())
    # a batch in the direct shape too
    bufs = alloc(views)
    ctx.render_batch(as_api(views, bufs, ld, n_lights), direct)
    ctx.sync()
    for v, g in zip(views, bufs):
        one = RenderBuffers(v["w"], v["h"])
        one.render(ctx, v["cam"], v["prm"], v["seq"], ld, n_lights, batched)
        ctx.sync()
        same(g.fetch(), one.fetch())


@pytest.mark.parametrize("prec,margin", [(lyap.PREC_PLAIN, 0.0), (lyap.PREC_PRODUCT, 0.0),
                                         (lyap.PREC_FP64, 0.0), (lyap.PREC_PRODUCT, 0.01)])
def test_march_passes_same_bytes(ctx, prec, margin):
    """A small launch marches in passes (a set-up launch, then launches of
    one round each with 4 to 64 lanes per ray, the last running its rays to
    the end); the persistent levels (LYAP_SCHEDULE_LEVELS) and the
    reference's shape give the same bytes and evaluation counts, view by
    view (the 20-letter sequence takes the 32-slot window) and as one batch."""
    views, lights, n_lights = small_batch(6, seed=5)
    ld = device_lights(lights, n_lights)
    passes = W.c2_mode(precision=prec, fp64_margin=margin)
    levels = W.c2_mode(precision=prec, fp64_margin=margin)
    levels.schedule = lyap.SCHEDULE_LEVELS
    direct = W.c2_mode(precision=prec, fp64_margin=margin)
    direct.schedule = lyap.SCHEDULE_DIRECT
    for v in views:
        out, ev = [], []
        for m in (passes, levels, direct):
            b = RenderBuffers(v["w"], v["h"])
            b.render(ctx, v["cam"], v["prm"], v["seq"], ld, n_lights, m)
            ctx.sync()
            ev.append(ctx.last_evals())
            out.append(b.fetch())
        assert ev[0] == ev[1] == ev[2]
        same(out[0], out[1])
        same(out[0], out[2])
    ref = []
    for m in (passes, levels):
        bufs = alloc(views)
        ctx.render_batch(as_api(views, bufs, ld, n_lights), m)
        ctx.sync()
        ref.append((ctx.last_evals(), [b.fetch() for b in bufs]))
    assert ref[0][0] == ref[1][0]
    for a, b in zip(ref[0][1], ref[1][1]):
        same(a, b)


def test_march_passes_many_views_same_bytes(ctx):
    """Twelve views make 96 sub-lists (eight per view), more than a wave
    holds, so the passes and the refinement launch find their chunks by
    walking the counts: the batch marches in passes (nine launches) and gives
    the persistent levels' bytes and evaluation count."""
    views, lights, n_lights = small_batch(12, seed=9, sides=(24, 40, 17))
    ld = device_lights(lights, n_lights)
    passes = W.c2_mode(fp64_margin=0.0)
    levels = W.c2_mode(fp64_margin=0.0)
    levels.schedule = lyap.SCHEDULE_LEVELS
    ref = []
    for m in (passes, levels):
        bufs = alloc(views)
        ctx.timing_collect()
        ctx.timing(True)
        try:
            ctx.render_batch(as_api(views, bufs, ld, n_lights), m)
            ctx.sync()
            launches = ctx.timing_collect()[1]
        finally:
            ctx.timing(False)
        ref.append((ctx.last_evals(), [b.fetch() for b in bufs], launches))
    if os.environ.get("LYAP_MARCH_PASSES") != "0":
        assert ref[0][2] == 9
    assert ref[0][0] == ref[1][0]
    for a, b in zip(ref[0][1], ref[1][1]):
        same(a, b)


def test_march_passes_benched_frame(ctx):
    """The 512x512 frame in the benchmarked mode (product form, no FP64
    margin): passes and persistent levels give the same bytes."""
    cam, prm, seq, lights, n_lights = W.render_config(512, 512)
    ld = device_lights(lights, n_lights)
    out = []
    for sched in (lyap.SCHEDULE_BATCHED, lyap.SCHEDULE_LEVELS):
        m = W.c2_mode(precision=lyap.PREC_PRODUCT, fp64_margin=0.0)
        m.schedule = sched
        b = RenderBuffers(512, 512)
        b.render(ctx, cam, prm, seq, ld, n_lights, m)
        ctx.sync()
        out.append((ctx.last_evals(), b.fetch()))
    assert out[0][0] == out[1][0]
    same(out[0][1], out[1][1])


def test_bad_schedule(ctx):
    views, lights, n_lights = small_batch(1)
    ld = device_lights(lights, n_lights)
    m = W.c2_mode(fp64_margin=0.0)
    m.schedule = 7
    b = RenderBuffers(8, 8)
    v = views[0]
    with pytest.raises(lyap.LyapError) as e:
        b.render(ctx, v["cam"], v["prm"], v["seq"], ld, n_lights, m)
    assert e.value.status == lyap.ERR_INVALID_ARGUMENT


def test_last_evals_after_capture_and_destroy(ctx):
    """Capturing a frame runs nothing, so the evaluation count still reads
    the previous call's; a replay reports its own; after the frame is
    destroyed the count no longer reads the frame's freed counters."""
    views, lights, n_lights = small_batch(n=4, seed=5)
    ld = device_lights(lights, n_lights)
    mode = W.c2_mode(fp64_margin=0.0)
    bufs = alloc(views)
    api = as_api(views, bufs, ld, n_lights)
    ctx.render_batch(api, mode)
    ctx.sync()
    e0 = ctx.last_evals()
    assert e0 > 0
    frame = ctx.frame_capture(api, mode)
    assert ctx.last_evals() == e0
    frame.replay()
    ctx.sync()
    assert ctx.last_evals() == e0
    frame.close()
    assert ctx.last_evals() in (0, e0)
    ctx.render_batch(api, mode)
    ctx.sync()
    assert ctx.last_evals() == e0
