"""Views and volume blocks in one persistent launch (lyap_run_batch).

 - a batch of views (different sizes, sequences, iteration counts) and
   volume blocks (whole grids and z slabs of larger ones, a 20-letter
   sequence that needs the 32-slot window) gives each view the bytes of
   rendering it alone (lyap_calc_render) and each block the bytes of the
   same slices of lyap_calc_volume over the whole grid, with the FP64
   re-evaluation rules of the precision workloads; the evaluation count is
   the views' plus one per voxel;
 - a batch of blocks only, and blocks whose precision differs from the
   views' (run apart), give the same bytes;
 - every tuning (voxel tile, level budgets, lanes per ray) gives the same
   bytes, and the batch autotuner leaves the results unchanged.
"""
import numpy as np
import pytest

import lyapunov3d_amd as lyap
from tests import workloads as W
from tests.gpu_util import RenderBuffers, device_lights, empty, to_numpy, torch_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    torch_cuda()
    c = lyap.Context(0)
    yield c
    c.close()


def small_views(n=5, seed=7):
    views, lights, n_lights = W.batch_views(n, seed=seed, sides=(24, 40, 64, 17), accum_range=(60, 200),
                                            settle=40)
    return views, lights, n_lights


# (params, sequence, lo, hi, grid, z range) of the blocks
BLOCKS = [
    (dict(settle=60, accum=300), "ABBAC", (2.0, 2.0, 2.0), (4.0, 4.0, 4.0), (24, 24, 24), (0, 24)),
    (dict(settle=100, accum=500), "ABCABCCA", (2.5, 2.5, 2.5), (4.0, 4.0, 4.0), (40, 33, 21), (5, 16)),
    (dict(settle=37, accum=211), "ABCABCABCABCABCABCAB", (2.2, 2.4, 2.6), (3.9, 4.0, 3.8), (13, 9, 7), (0, 7)),
    (dict(settle=0, accum=50), "AB", (2.0, 2.0, 2.0), (4.0, 4.0, 4.0), (5, 3, 2), (0, 2)),
]


def alloc_blocks():
    out = []
    for prm_kw, seq, lo, hi, n, (z0, z1) in BLOCKS:
        buf = empty(4 * n[0] * n[1] * (z1 - z0))
        buf.fill_(0x55)
        out.append(buf)
    return out


def api_blocks(bufs):
    return [lyap.make_block(lyap.make_params(**kw), seq, lo, hi, n, b.data_ptr(), z0, z1)
            for (kw, seq, lo, hi, n, (z0, z1)), b in zip(BLOCKS, bufs)]


def block_alone(ctx, i, vmode):
    kw, seq, lo, hi, n, (z0, z1) = BLOCKS[i]
    full = empty(4 * n[0] * n[1] * n[2])
    ctx.calc_volume(full.data_ptr(), lyap.make_params(**kw), seq, lo, hi, n, vmode)
    ctx.sync()
    v = to_numpy(full, np.uint32)[:n[0] * n[1] * n[2]].reshape(n[2], n[1] * n[0])
    return v[z0:z1].reshape(-1)


def fetch_block(buf, i):
    kw, seq, lo, hi, n, (z0, z1) = BLOCKS[i]
    return to_numpy(buf, np.uint32)[:n[0] * n[1] * (z1 - z0)]


def same(a, b):
    for k in ("rgba", "depth", "lambda", "evals", "points"):
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k


MODES = [
    ("product_strict", W.c2_mode(fp64_margin=0.0), lyap.volume_mode(**W.C3_STRICT)),
    ("product_margins", W.c2_mode(fp64_margin=0.01), lyap.volume_mode(**W.C3_MODE)),
    ("product_plain_rule", W.c2_mode(fp64_margin=0.0), lyap.volume_mode(lyap.PREC_PRODUCT)),
    ("plain", W.c2_mode(precision=lyap.PREC_PLAIN, fp64_margin=0.0), lyap.volume_mode(lyap.PREC_PLAIN)),
    ("fp64", W.c2_mode(precision=lyap.PREC_FP64, fp64_margin=0.0), lyap.volume_mode(lyap.PREC_FP64)),
    ("walk_product", lyap.render_mode(precision=lyap.PREC_PRODUCT), lyap.volume_mode(**W.C3_MODE)),
    ("mismatch", W.c2_mode(fp64_margin=0.0), lyap.volume_mode(lyap.PREC_PLAIN)),
]


@pytest.mark.parametrize("name,mode,vmode", MODES, ids=[m[0] for m in MODES])
def test_mixed_batch_equals_each_alone(ctx, name, mode, vmode):
    views, lights, n_lights = small_views()
    ld = device_lights(lights, n_lights)
    vbufs = [RenderBuffers(v["w"], v["h"]) for v in views]
    api = [lyap.make_view(v["cam"], v["prm"], v["seq"], v["w"], v["h"], b.rgba.data_ptr(), b.points.data_ptr(),
                          b.depth.data_ptr(), b.lam.data_ptr(), b.evals.data_ptr(), ld.data_ptr(), n_lights)
           for v, b in zip(views, vbufs)]
    bbufs = alloc_blocks()
    ctx.run_batch(api, api_blocks(bbufs), mode, vmode)
    ctx.sync()
    batch_evals = ctx.last_evals()
    got_v = [b.fetch() for b in vbufs]
    got_b = [fetch_block(b, i) for i, b in enumerate(bbufs)]
    total = 0
    for v, g in zip(views, got_v):
        one = RenderBuffers(v["w"], v["h"])
        one.render(ctx, v["cam"], v["prm"], v["seq"], ld, n_lights, mode)
        ctx.sync()
        same(g, one.fetch())
        total += ctx.last_evals()
    voxels = 0
    for i, g in enumerate(got_b):
        alone = block_alone(ctx, i, vmode)
        assert np.array_equal(g, alone), f"block {i}"
        voxels += g.size
    if name != "mismatch":
        assert batch_evals == total + voxels
    # the blocks hold lambdas, the settle-0 block the 0.5 guard's zeros
    assert np.all(got_b[3].view(np.float32) == 0.0)
    assert np.isfinite(got_b[0].view(np.float32)).mean() > 0.9


def test_blocks_only(ctx):
    vmode = lyap.volume_mode(**W.C3_STRICT)
    bbufs = alloc_blocks()
    ctx.run_batch([], api_blocks(bbufs), None, vmode)
    ctx.sync()
    assert ctx.last_evals() == sum(fetch_block(b, i).size for i, b in enumerate(bbufs))
    for i, b in enumerate(bbufs):
        assert np.array_equal(fetch_block(b, i), block_alone(ctx, i, vmode)), f"block {i}"


@pytest.mark.parametrize("voxel_tile,budget,lanes", [(64, (16, 32, 96), 1), (192, (64, 96, 256), 4),
                                                     (1024, (0, 0, 0), 2), (100, (24, 0, 0), 16)])
def test_tunings_same_bytes(ctx, voxel_tile, budget, lanes):
    views, lights, n_lights = small_views(seed=9)
    ld = device_lights(lights, n_lights)
    mode, vmode = W.c2_mode(fp64_margin=0.01), lyap.volume_mode(**W.C3_STRICT)
    res = []
    for tuned in (False, True):
        vbufs = [RenderBuffers(v["w"], v["h"]) for v in views]
        api = [lyap.make_view(v["cam"], v["prm"], v["seq"], v["w"], v["h"], b.rgba.data_ptr(),
                              b.points.data_ptr(), b.depth.data_ptr(), b.lam.data_ptr(), b.evals.data_ptr(),
                              ld.data_ptr(), n_lights) for v, b in zip(views, vbufs)]
        bbufs = alloc_blocks()
        t = None
        if tuned:
            t = lyap.Tuning()
            t.voxel_tile, t.lanes_per_ray = voxel_tile, lanes
            for k in range(3):
                t.level_budget[k] = budget[k]
        ctx.run_batch(api, api_blocks(bbufs), mode, vmode, t)
        ctx.sync()
        res.append(([b.fetch() for b in vbufs], [fetch_block(b, i) for i, b in enumerate(bbufs)],
                    ctx.last_evals()))
    (va, ba, ea), (vb, bb, eb) = res
    assert ea == eb
    for x, y in zip(va, vb):
        same(x, y)
    for x, y in zip(ba, bb):
        assert np.array_equal(x, y)


def test_autotune_batch_keeps_results(ctx):
    views, lights, n_lights = small_views(seed=3)
    ld = device_lights(lights, n_lights)
    mode, vmode = W.c2_mode(fp64_margin=0.0), lyap.volume_mode(**W.C3_MODE)
    vbufs = [RenderBuffers(v["w"], v["h"]) for v in views]
    api = [lyap.make_view(v["cam"], v["prm"], v["seq"], v["w"], v["h"], b.rgba.data_ptr(), b.points.data_ptr(),
                          b.depth.data_ptr(), b.lam.data_ptr(), b.evals.data_ptr(), ld.data_ptr(), n_lights)
           for v, b in zip(views, vbufs)]
    bbufs = alloc_blocks()
    blocks = api_blocks(bbufs)
    ctx.run_batch(api, blocks, mode, vmode)
    ctx.sync()
    before_v = [b.fetch() for b in vbufs]
    before_b = [fetch_block(b, i) for i, b in enumerate(bbufs)]
    best = ctx.autotune_batch(api, blocks, mode, vmode)
    assert best.lanes_per_ray in (0, 1, 2, 4, 8, 16) and best.slots in (0, 1, 2, 4)
    assert best.voxel_tile in (0, 64, 128, 512, 1024)
    ctx.run_batch(api, blocks, mode, vmode)
    ctx.sync()
    for b, e in zip(vbufs, before_v):
        same(b.fetch(), e)
    for i, (b, e) in enumerate(zip(bbufs, before_b)):
        assert np.array_equal(fetch_block(b, i), e)
    # the context keeps the choice; clear it for the other tests
    assert ctx.get_tuning().as_dict() == best.as_dict()
    ctx.set_tuning(None)
    assert ctx.get_tuning().as_dict() == lyap.Tuning().as_dict()


def test_block_errors(ctx):
    buf = empty(64)
    bad = lyap.make_block(lyap.make_params(), "AB", (2, 2, 2), (4, 4, 4), (4, 4, 4), buf.data_ptr(), 3, 2)
    with pytest.raises(lyap.LyapError) as e:
        ctx.run_batch([], [bad])
    assert e.value.status == lyap.ERR_INVALID_ARGUMENT
    bad = lyap.make_block(lyap.make_params(), "AB", (2, 2, 2), (4, 4, 4), (4, 4, 4), 0)
    with pytest.raises(lyap.LyapError) as e:
        ctx.run_batch([], [bad])
    assert e.value.status == lyap.ERR_INVALID_ARGUMENT
    bad = lyap.make_block(lyap.make_params(), "AB", (2, 2, 2), (4, 4, 4), (4, 4, 4), buf.data_ptr())
    bad.seq[0] = 7
    with pytest.raises(lyap.LyapError) as e:
        ctx.run_batch([], [bad])
    assert e.value.status == lyap.ERR_BAD_SEQUENCE
