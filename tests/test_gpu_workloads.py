"""The batched (C4) and served (C5) workloads at full size on the GPU.

 - C4: the 64-view batch (sides 128..1024, camera radius 3..6, sequences of
   3-12 letters, 100 + 200..2000 iterations) in one launch sequence, in the
   benchmarked mode (product form, no FP64 margin) and in FP64. On three
   seeded views (every k-th row, about 70k rays each) the FP64 batch equals
   the FP64 oracle exactly (hit, depth, sample count), and the product batch,
   pooled over the three, agrees with it on hits and on depth within one
   march step at least as often as the reference's own FP32 build (the
   oracle's FP32-as-written form) on the same rays, less 0.002;
 - C5: the first 2 s of the seeded Poisson stream (200 requests/s, 70 %
   views of side 256/512/1024, 30 % volume blocks of edge 32..256 inside
   [2,4]^3, 100 + 200..4000 iterations) paced through the server with the
   benchmark's configuration; every result equals the same request computed
   alone (lyap_calc_render / lyap_calc_volume) byte for byte, compared on
   the device, and two views and two volumes sit at the reference-FP32
   level against the FP64 oracle.
"""
import time

import numpy as np
import pytest

import lyapunov3d_amd as lyap
from oracle import oracle as O
from tests import workloads as W
from tests.gpu_util import RenderBuffers, device_lights, empty, to_numpy, torch_cuda

pytestmark = pytest.mark.gpu
SLACK = 0.002


@pytest.fixture(scope="module")
def ctx():
    torch_cuda()
    c = lyap.Context(0)
    yield c
    c.close()


def depth_agreement(cam, prm, got_depth, ref_depth, ys):
    hit_g, hit_r = ~np.isnan(got_depth), ~np.isnan(ref_depth)
    agree = float(np.mean(hit_g == hit_r))
    both = hit_g & hit_r
    yy, xx = np.nonzero(both)
    dt = np.empty(len(xx))
    for i, (x, y) in enumerate(zip(xx, yy)):
        _, V, t0, t1 = O.ray_setup(cam, prm, int(x), int(ys[y]))
        dt[i] = (t1 - t0) / prm.depth
    within = float(np.mean(np.abs(got_depth[both] - ref_depth[both]) <= dt * (1 + 1e-5))) if both.any() else 1.0
    return agree, within


def test_c4_full_batch(ctx):
    views, lights, n_lights = W.batch_views(64, seed=4)
    ld = device_lights(lights, n_lights)
    bufs = [RenderBuffers(v["w"], v["h"]) for v in views]
    api = [lyap.make_view(v["cam"], v["prm"], v["seq"], v["w"], v["h"], b.rgba.data_ptr(), b.points.data_ptr(),
                          b.depth.data_ptr(), b.lam.data_ptr(), b.evals.data_ptr(), ld.data_ptr(), n_lights)
           for v, b in zip(views, bufs)]
    picks = sorted(np.random.default_rng(41).choice(64, 3, replace=False).tolist())
    out = {}
    for name, mode in (("product", W.c2_mode(fp64_margin=0.0)),
                       ("fp64", W.c2_mode(precision=lyap.PREC_FP64, fp64_margin=0.0))):
        print(f"c4: 64-view batch, {name}", flush=True)
        t = time.perf_counter()
        ctx.render_batch(api, mode)
        ctx.sync()
        evals = ctx.last_evals()
        per_view = [b.fetch() for b in bufs]
        assert evals == sum(int(g["evals"].astype(np.int64).sum()) for g in per_view)
        out[name] = [per_view[i] for i in picks]
        print(f"c4 {name}: {time.perf_counter() - t:.2f} s, {evals} evaluations", flush=True)
    # pooled over the three views: a within-one-step fraction over a few
    # thousand rays moves by more than the slack from one FP32 realisation
    # to another, over the ~200k pooled rays by less
    tot = {"n": 0, "agree": 0.0, "both": 0, "within": 0.0, "ragree": 0.0, "rboth": 0, "rwithin": 0.0}
    for k, i in enumerate(picks):
        v = views[i]
        step = max(1, v["h"] // max(1, 70000 // v["w"]))
        ys = np.arange(0, v["h"], step)
        print(f"c4 view {i}: {v['w']}x{v['h']} seq {v['seq']} accum {v['prm'].accum}, rows 0::{step}", flush=True)
        ref64 = O.render(v["cam"], v["prm"], v["seq"], lights, n_lights, v["w"], v["h"],
                         mode=W.c2_mode(), lam=O.F64, rows=(0, step))
        ref32 = O.render(v["cam"], v["prm"], v["seq"], lights, n_lights, v["w"], v["h"],
                         mode=W.c2_mode(), lam=O.F32W, rows=(0, step))
        g64, gp = out["fp64"][k], out["product"][k]
        assert np.array_equal(np.isnan(g64["depth"][ys]), np.isnan(ref64["depth"][ys]))
        assert np.array_equal(g64["evals"][ys], ref64["evals"][ys])
        assert np.array_equal(g64["depth"][ys], ref64["depth"][ys], equal_nan=True)
        a, w = depth_agreement(v["cam"], v["prm"], gp["depth"][ys], ref64["depth"][ys], ys)
        ra, rw = depth_agreement(v["cam"], v["prm"], ref32["depth"][ys], ref64["depth"][ys], ys)
        n = ys.size * v["w"]
        both = int(np.sum(~np.isnan(gp["depth"][ys]) & ~np.isnan(ref64["depth"][ys])))
        rboth = int(np.sum(~np.isnan(ref32["depth"][ys]) & ~np.isnan(ref64["depth"][ys])))
        tot["n"] += n
        tot["agree"] += a * n
        tot["ragree"] += ra * n
        tot["both"] += both
        tot["within"] += w * both
        tot["rboth"] += rboth
        tot["rwithin"] += rw * rboth
        print(f"c4 view {i}: product hit {a:.4f} within {w:.4f}; reference FP32 {ra:.4f} {rw:.4f}")
    a, w = tot["agree"] / tot["n"], tot["within"] / tot["both"]
    ra, rw = tot["ragree"] / tot["n"], tot["rwithin"] / tot["rboth"]
    print(f"c4 pooled over {tot['n']} rays: product hit {a:.4f} within {w:.4f}; reference FP32 {ra:.4f} {rw:.4f}")
    assert a >= ra - SLACK and w >= rw - SLACK


def vol_agreement(got, ref, h1, h2):
    stable = (ref < -1e-3) & (np.abs(h1 - h2) < 1e-4)
    d = np.abs(got[stable].astype(np.float64) - ref[stable])
    within = float(np.mean(np.nan_to_num(d, nan=9.0) <= 1e-3)) if d.size else 1.0
    sign = float(np.mean((np.sign(got) == np.sign(ref)) | (np.isnan(got) & np.isnan(ref))))
    return within, sign


def test_c5_stream_slice(ctx):
    torch = torch_cuda()
    reqs, lights, n_lights = W.serving_stream(seconds=2.0)
    ld = device_lights(lights, n_lights)
    mode = W.c2_mode(fp64_margin=0.0)
    vmode = lyap.volume_mode(**W.C3_MODE)
    n_views = sum(q["kind"] == "view" for q in reqs)
    print(f"c5 slice: {len(reqs)} requests ({n_views} views)", flush=True)
    outs, api = [], []
    for q in reqs:
        if q["kind"] == "view":
            px = q["w"] * q["h"]
            o = (torch.empty(px, dtype=torch.int32, device="cuda"),
                 torch.empty(px, dtype=torch.float32, device="cuda"))
            v = lyap.make_view(q["cam"], q["prm"], q["seq"], q["w"], q["h"], o[0].data_ptr(), 0, 0,
                               o[1].data_ptr(), 0, ld.data_ptr(), n_lights)
            api.append(lyap.view_request(v, deadline_ms=50.0))
        else:
            o = (torch.empty(q["n"] ** 3, dtype=torch.float32, device="cuda"),)
            api.append(lyap.volume_request(q["prm"], q["seq"], q["lo"], q["hi"], q["n"], o[0].data_ptr(),
                                           deadline_ms=250.0))
        outs.append(o)
    torch.cuda.synchronize()
    srv = lyap.Server(ctx, lyap.server_config(hbm_budget_bytes=64 << 30, mode=mode, vol_mode=vmode))
    t0 = srv.clock_ms()
    i = done = 0
    tickets = {}
    while done < len(reqs):
        now = srv.clock_ms() - t0
        while i < len(reqs) and reqs[i]["t_ms"] <= now:
            tickets[srv.submit(api[i])] = i
            i += 1
        srv.progress(0.2)
        for t in srv.finished():
            assert srv.poll(t).status == lyap.OK, tickets[t]
            srv.release(t)
            done += 1
    st = srv.stats()
    print(f"c5 slice served in {(srv.clock_ms() - t0) / 1e3:.2f} s, {st.launches} launches", flush=True)
    srv.close()
    # each request alone, compared on the device
    for k, q in enumerate(reqs):
        if q["kind"] == "view":
            px = q["w"] * q["h"]
            a = (torch.empty(px, dtype=torch.int32, device="cuda"), torch.empty(px, dtype=torch.float32, device="cuda"))
            ctx.calc_render(a[0].data_ptr(), 0, q["cam"], q["prm"], q["seq"], ld.data_ptr(), n_lights, q["w"], q["h"],
                            0, a[1].data_ptr(), 0, mode)
            ctx.sync()
            assert torch.equal(a[0], outs[k][0]), k
            assert torch.equal(a[1].view(torch.int32), outs[k][1].view(torch.int32)), k
        else:
            a = torch.empty(q["n"] ** 3, dtype=torch.float32, device="cuda")
            ctx.calc_volume(a.data_ptr(), q["prm"], q["seq"], q["lo"], q["hi"], q["n"], vmode)
            ctx.sync()
            assert torch.equal(a.view(torch.int32), outs[k][0].view(torch.int32)), k
    # two views and two volumes against the FP64 oracle, at the reference-FP32 level
    vi = [k for k, q in enumerate(reqs) if q["kind"] == "view" and q["w"] == 256]
    vi = sorted(vi, key=lambda k: reqs[k]["prm"].accum)[:2]
    for k in vi:
        q = reqs[k]
        got = outs[k][1].cpu().numpy().reshape(q["h"], q["w"])
        ref64 = O.render(q["cam"], q["prm"], q["seq"], lights, n_lights, q["w"], q["h"], mode=W.c2_mode(), lam=O.F64)
        ref32 = O.render(q["cam"], q["prm"], q["seq"], lights, n_lights, q["w"], q["h"], mode=W.c2_mode(), lam=O.F32W)
        ys = np.arange(q["h"])
        hit = lambda lam_img: ~np.isnan(lam_img)
        a = float(np.mean(hit(got) == hit(ref64["lambda"])))
        ra = float(np.mean(hit(ref32["lambda"]) == hit(ref64["lambda"])))
        print(f"c5 view {k}: hit agreement {a:.4f} (reference FP32 {ra:.4f})")
        assert a >= ra - SLACK
    wi = [k for k, q in enumerate(reqs) if q["kind"] == "volume" and q["n"] == 32][:2]
    for k in wi:
        q = reqs[k]
        got = outs[k][0].cpu().numpy().reshape(32, 32, 32)
        prm = q["prm"]
        ref = O.volume(prm, q["seq"], q["lo"], q["hi"], 32, mode=O.F64).astype(np.float64)
        f32 = O.volume(prm, q["seq"], q["lo"], q["hi"], 32, mode=O.F32W)
        zs = [O.grid_coords(q["lo"][a], q["hi"][a], 32) for a in range(3)]
        zz, yy, xx = np.meshgrid(zs[2], zs[1], zs[0], indexing="ij")
        xyz = np.stack([xx, yy, zz], -1).reshape(-1, 3)
        h1, h2 = O.lyap_halves(xyz, prm.d, prm.settle, prm.accum, prm.accum // 2, q["seq"])
        h1, h2 = h1.reshape(ref.shape), h2.reshape(ref.shape)
        w, sgn = vol_agreement(got, ref, h1, h2)
        rw, rs = vol_agreement(f32, ref, h1, h2)
        print(f"c5 volume {k}: within {w:.4f} sign {sgn:.4f} (reference FP32 {rw:.4f} {rs:.4f})")
        assert w >= rw - SLACK and sgn >= rs - SLACK
