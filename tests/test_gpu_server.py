"""The serving loop (lyap_server_*) on the GPU.

 - views and volumes served together, with launches small enough that every
   view is cut into pieces spread over several launches and slots, give
   exactly the bytes of rendering or computing each request alone (a ray's
   or a voxel's arithmetic never depends on what shares its launch), both
   when launches carry views and volume slabs together and when volumes run
   on slots of their own;
 - resident outputs (no caller buffers) hold the same bytes as caller buffers;
 - cancellation, the HBM budget's admission rule and the error paths.
"""
import ctypes as C

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


def fetch_device(ctx, ptr, n_bytes, dtype):
    out = np.empty(n_bytes, dtype=np.uint8)
    rc = ctx.lib.lyap_copy_to_host(ctx.h, out.ctypes.data_as(C.c_void_p), ptr, n_bytes)
    assert rc == lyap.OK
    return out.view(dtype)


def served_views(n=8, seed=11):
    views, lights, n_lights = W.batch_views(n, seed=seed, sides=(24, 40, 64, 96, 17, 130),
                                            accum_range=(60, 300), settle=40)
    views[1]["seq"] = "ABCABCABCABCABCABCAB"   # period 20: the 32-wide window
    return views, lights, n_lights


VOLUMES = [
    # (n, lo, hi, seq, settle, accum)
    ((17, 9, 33), (2.0, 2.0, 2.0), (4.0, 4.0, 4.0), "ABBAC", 50, 200),
    ((32, 32, 32), (2.5, 3.0, 2.2), (3.5, 3.9, 3.0), "ABCABCCA", 100, 400),
    ((40, 24, 8), (3.0, 2.0, 3.5), (4.0, 3.0, 4.0), "AB", 60, 150),
]


def volume_args(spec):
    n, lo, hi, seq, settle, accum = spec
    prm = lyap.make_params(settle=settle, accum=accum)
    return n, lo, hi, seq, prm


@pytest.mark.parametrize("resident,flags,vrule", [(False, 0, "margin"), (True, 0, "strict"),
                                                 (False, lyap.SERVER_SEPARATE_VOLUMES, "margin"),
                                                 (True, lyap.SERVER_SEPARATE_VOLUMES, "strict")])
def test_mix_equals_alone(ctx, resident, flags, vrule):
    """Launches that carry views and volume slabs together (the default) and
    separate view and volume slots give the bytes of each request alone."""
    torch = torch_cuda()
    views, lights, n_lights = served_views()
    ld = device_lights(lights, n_lights)
    mode = W.c2_mode(fp64_margin=0.0)
    vmode = lyap.volume_mode(**(W.C3_MODE if vrule == "margin" else W.C3_STRICT))
    cfg = lyap.server_config(render_slots=2, volume_slots=1, slice_ms=0.02, chunk_rays=128,
                             mode=mode, vol_mode=vmode, flags=flags)
    srv = lyap.Server(ctx, cfg)
    bufs, vbufs, tickets = [], [], []
    # interleave views and volumes in submission order
    for i, v in enumerate(views):
        b = RenderBuffers(v["w"], v["h"])
        bufs.append(b)
        if resident:
            api = lyap.make_view(v["cam"], v["prm"], v["seq"], v["w"], v["h"], 0, 0, 0, 0, 0,
                                 ld.data_ptr(), n_lights)
        else:
            api = lyap.make_view(v["cam"], v["prm"], v["seq"], v["w"], v["h"], b.rgba.data_ptr(),
                                 b.points.data_ptr(), b.depth.data_ptr(), b.lam.data_ptr(),
                                 b.evals.data_ptr(), ld.data_ptr(), n_lights)
        tickets.append(("view", i, srv.submit(lyap.view_request(api, deadline_ms=5.0 * (i % 3)))))
        if i < len(VOLUMES):
            n, lo, hi, seq, prm = volume_args(VOLUMES[i])
            out = None if resident else empty(4 * n[0] * n[1] * n[2])
            vbufs.append(out)
            req = lyap.volume_request(prm, seq, lo, hi, n, out.data_ptr() if out is not None else 0)
            tickets.append(("vol", i, srv.submit(req)))
    torch.cuda.synchronize()
    srv.drain(120000.0)
    st = srv.stats()
    assert st.completed == len(tickets) and st.failed == 0 and st.queued == 0 and st.active == 0
    assert st.launches > len(tickets)          # pieces really were spread over launches
    total_rays = sum(v["w"] * v["h"] for v in views)
    assert st.rays == total_rays
    assert st.voxels == sum(int(np.prod(s[0])) for s in VOLUMES)
    if resident:
        assert st.resident_bytes == sum(8 * v["w"] * v["h"] for v in views) + 4 * st.voxels

    assert sorted(srv.finished(cap=4)) == sorted(t for _, _, t in tickets)
    assert srv.finished() == []
    served_evals = 0
    for kind, i, t in tickets:
        r = srv.poll(t)
        assert r.status == lyap.OK, (kind, i)
        assert r.finish_ms >= r.start_ms >= r.submit_ms
        if kind == "view":
            v = views[i]
            alone = RenderBuffers(v["w"], v["h"])
            alone.render(ctx, v["cam"], v["prm"], v["seq"], ld, n_lights, mode)
            ctx.sync()
            served_evals += ctx.last_evals()
            ref = alone.fetch()
            assert r.rays == v["w"] * v["h"]
            if resident:
                rgba = fetch_device(ctx, r.rgba, 4 * r.rays, np.uint8).reshape(v["h"], v["w"], 4)
                lam = fetch_device(ctx, r.lambda_, 4 * r.rays, np.float32).reshape(v["h"], v["w"])
                assert np.array_equal(rgba, ref["rgba"]), i
                assert np.array_equal(lam, ref["lambda"], equal_nan=True), i
            else:
                got = bufs[i].fetch()
                for k in ("rgba", "depth", "lambda", "evals", "points"):
                    assert np.array_equal(got[k], ref[k], equal_nan=True), (i, k)
        else:
            n, lo, hi, seq, prm = volume_args(VOLUMES[i])
            count = n[0] * n[1] * n[2]
            ref_buf = empty(4 * count)
            ctx.calc_volume(ref_buf.data_ptr(), prm, seq, lo, hi, n, vmode)
            ctx.sync()
            reeval = ctx.last_reevaluated()
            ref = to_numpy(ref_buf, np.float32)[:count]
            if resident:
                got = fetch_device(ctx, r.lambda_, 4 * count, np.float32)
            else:
                got = to_numpy(vbufs[i], np.float32)[:count]
            assert np.array_equal(got, ref, equal_nan=True), i
            assert r.voxels == count and r.evals >= count
            served_evals += count + reeval
    # every evaluation is counted once, whatever the launches were
    assert st.evals == served_evals
    for _, _, t in tickets:
        srv.release(t)
    assert srv.stats().resident_bytes == 0
    srv.close()


def test_cancel_and_errors(ctx):
    torch = torch_cuda()
    views, lights, n_lights = served_views(3, seed=5)
    ld = device_lights(lights, n_lights)
    cfg = lyap.server_config(slice_ms=0.02, chunk_rays=64, mode=W.c2_mode(fp64_margin=0.0))
    srv = lyap.Server(ctx, cfg)
    api = [lyap.make_view(v["cam"], v["prm"], v["seq"], v["w"], v["h"], 0, 0, 0, 0, 0,
                          ld.data_ptr(), n_lights) for v in views]
    torch.cuda.synchronize()
    # cancelled before anything was issued: ends at once
    t0 = srv.submit(lyap.view_request(api[0]))
    srv.cancel(t0)
    assert srv.poll(t0).status == lyap.ERR_CANCELLED
    # cancelled after its first pieces were issued: ends once they are done
    t1 = srv.submit(lyap.view_request(api[2]))
    srv.progress(0.0)
    srv.cancel(t1)
    srv.drain(60000.0)
    assert srv.poll(t1).status in (lyap.ERR_CANCELLED, lyap.OK)
    # a finished request cannot be cancelled into another state
    t2 = srv.submit(lyap.view_request(api[1]))
    srv.drain(60000.0)
    srv.cancel(t2)
    assert srv.poll(t2).status == lyap.OK
    # errors
    with pytest.raises(lyap.LyapError) as e:
        srv.poll(987654)
    assert e.value.status == lyap.ERR_NOT_FOUND
    t3 = srv.submit(lyap.view_request(api[1]))
    with pytest.raises(lyap.LyapError) as e:
        srv.release(t3)                       # still pending
    assert e.value.status == lyap.ERR_PENDING
    srv.drain(60000.0)
    bad = lyap.view_request(api[1])
    bad.view.width = 0
    with pytest.raises(lyap.LyapError) as e:
        srv.submit(bad)
    assert e.value.status == lyap.ERR_INVALID_ARGUMENT
    bad = lyap.view_request(api[1])
    bad.view.seq[0] = 7
    with pytest.raises(lyap.LyapError) as e:
        srv.submit(bad)
    assert e.value.status == lyap.ERR_BAD_SEQUENCE
    bad = lyap.Request()
    bad.kind = 5
    with pytest.raises(lyap.LyapError):
        srv.submit(bad)
    for t in (t0, t1, t2, t3):
        srv.release(t)
    srv.close()


def test_budget_admission(ctx):
    torch = torch_cuda()
    views, lights, n_lights = served_views(2, seed=9)
    ld = device_lights(lights, n_lights)
    mode = W.c2_mode(fp64_margin=0.0)
    side = 64
    one = 8 * side * side                      # RGBA + lambda of one resident view
    srv = lyap.Server(ctx, lyap.server_config(hbm_budget_bytes=one, mode=mode))
    api = [lyap.make_view(v["cam"], v["prm"], v["seq"], side, side, 0, 0, 0, 0, 0,
                          ld.data_ptr(), n_lights) for v in views]
    torch.cuda.synchronize()
    # larger than the whole budget: refused at submission
    big = lyap.make_view(views[0]["cam"], views[0]["prm"], views[0]["seq"], 2 * side, side, 0, 0, 0, 0, 0,
                         ld.data_ptr(), n_lights)
    with pytest.raises(lyap.LyapError) as e:
        srv.submit(lyap.view_request(big))
    assert e.value.status == lyap.ERR_BUDGET
    ta = srv.submit(lyap.view_request(api[0]))
    tb = srv.submit(lyap.view_request(api[1]))
    # b waits for a's outputs to be released
    with pytest.raises(lyap.LyapError) as e:
        srv.drain(60000.0)
    assert e.value.status == lyap.ERR_BUDGET
    assert srv.poll(ta).status == lyap.OK
    assert srv.poll(tb).status == lyap.ERR_PENDING and srv.stats().queued == 1
    assert srv.stats().resident_bytes == one
    srv.release(ta)
    srv.drain(60000.0)
    rb = srv.poll(tb)
    assert rb.status == lyap.OK and rb.start_ms >= srv.poll(tb).submit_ms
    assert srv.stats().peak_resident_bytes == one
    srv.release(tb)
    srv.close()
