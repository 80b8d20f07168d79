#!/usr/bin/env python3
"""Throughput of the Lyapunov path on one MI355X.

Headline workload (one step = one frame): the plain render, a 512x512 image
of the [2,4]^3 Lyapunov box seen from distance 4 along a seeded unit
quaternion, 45 degree field of view, sequence AABAB, 100 settle + 300
accumulated iterations per lambda, up to 500 uniform steps per ray, surface
at the first lambda <= 0, 8 bisection steps, normal from six samples at
h = 1e-3, Phong shading, outputs RGBA8 + depth + lambda + LyapPoint per pixel.
`value` is rays/s of that frame loop with all buffers resident in HBM.

Also reported (same JSON line, "workloads"):
  c2  the same frame with FP64 re-evaluation near the threshold during
      bisection, in exact FP64, and in the plain FP32 per-step-log form;
  c1  64^3 volume over [2,4]^3, ABBAC, 100 + 400 (plain and product forms);
  c3  256^3 volume over [2.5,4]^3, ABCABCCA, 200 + 2000: plain per-step-log,
      product, and product with FP64 re-evaluation (three rules), with
      max/RMS |dlambda| on the stable voxels and sign agreement, over all
      voxels, against the FP64 form (the reference orbit bit for bit);
  c4  64 views of seeded size 128..1024 in one batched launch sequence
      (camera radius 3..6, sequences of 3-12 letters, 200..2000 iterations)
      with the built-in and the autotuned launch shapes, a captured graph,
      the depth agreement with the FP64 batch, against rendering the same
      views one request at a time in the plain FP32 per-step-log form;
      then the same views and eight 128^3 volume blocks in one launch
      sequence against views then volumes;
  c5  the serving stream: Poisson arrivals at 200 requests/s (70% views of
      side 256/512/1024, 30% volume blocks of edge 32..256 inside [2,4]^3,
      200..4000 iterations) paced through the server for --c5-seconds
      (p50/p99 latency, served vs offered rays/s), the head of the stream
      again with volumes under the rule that meets the FP64 tolerance, then
      the first --c5-capacity requests offered at once to the server (both
      rules, the autotuned shapes, separate volume slots) and to plain
      per-request calls, and the FP64 error of a sample of those requests.
The roofline uses algorithmic work: 3*settle + 6*accum FLOP per lambda
evaluation, against MI355X's FP32 vector peak of 157.3 TFLOP/s. The CPU
baseline is the oracle (the reference's algorithm restated for the host,
FP32 as written) on bounded samples of the workloads, timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3     # MI355X dense FP32 vector


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--precision", default="product", choices=["product", "plain", "fp64"])
    p.add_argument("--size", type=int, default=512)
    p.add_argument("--cpu-seconds", type=float, default=12.0,
                   help="approximate CPU budget of the baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-tune", action="store_true",
                   help="time the headline frame with the built-in launch shape")
    p.add_argument("--workloads", default="c2,c1,c3,c4,c5", help="extra workloads ('' for none)")
    p.add_argument("--c4-views", type=int, default=64)
    p.add_argument("--c5-seconds", type=float, default=60.0,
                   help="length of the paced request stream (the workload's 60 s)")
    p.add_argument("--c5-strict-seconds", type=float, default=20.0,
                   help="length of the paced leg with volumes under the tolerance rule")
    p.add_argument("--copy-calibration", action="store_true",
                   help="first run a plain 1 GiB device copy three times (the reference kernel "
                        "for scaling HBM counters)")
    p.add_argument("--c5-capacity", type=int, default=1000,
                   help="requests offered at once for the served vs plain capacity")
    return p.parse_args()


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (lines 'TRAFFIC <kernel> <bytes per launch>' and 'COMMIT <hash>'): the
    counters need a profiler run of their own, so the figure is not measured
    in this run; the source string says where and at which commit it was."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_render_c2.md")
    value, commit = None, "unknown"
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "TRAFFIC" and parts[1] == kernel:
                    value = float(parts[2])
                if len(parts) == 2 and parts[0] == "COMMIT":
                    commit = parts[1]
    except OSError:
        pass
    src = (f"profiles/hbm_traffic_render_c2.md (rocprofv3 FETCH_SIZE and WRITE_SIZE at commit {commit}); "
           "not measured in this run")
    return value, src


def flops(prm):
    return 3 * prm.settle + 6 * prm.accum


class Timer:
    """HIP events on the library's stream."""

    def __init__(self, torch, ctx):
        self.torch = torch
        self.ctx = ctx
        self.stream = torch.cuda.ExternalStream(ctx.stream)

    def run(self, fn, reps):
        t = self.torch
        fn()
        self.ctx.sync()
        s = t.cuda.Event(enable_timing=True)
        e = t.cuda.Event(enable_timing=True)
        s.record(self.stream)
        for _ in range(reps):
            fn()
        e.record(self.stream)
        self.ctx.sync()
        t.cuda.synchronize()
        return s.elapsed_time(e) / reps


def _host_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _median3(fn, warm=None):
    """Median of three timed calls (seconds) after an untimed warm-up call
    (of `warm` when given: a smaller piece of the same work)."""
    (warm or fn)()
    ts = []
    for _ in range(3):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts))


def cpu_baseline(cam, prm, seq, lights, n_lights, w, h, mode, budget_s, c5_requests):
    """The oracle (the reference's algorithm restated for the host, FP32 as
    written) timed on this host's cores: the headline frame on every k-th
    row; `detail` adds the FP64 frame, the 64^3 and 256^3 volumes, a seeded
    1/256 ray subsample of the 64-view batch and a seeded sample of the
    serving stream's requests. Every figure is the median of three runs
    after a warm-up."""
    from oracle import oracle as O
    from tests import workloads as W
    threads = O.threads()

    def rows_for(lam, budget):
        t = time.perf_counter()
        O.render(cam, prm, seq, lights, n_lights, w, h, mode=mode, lam=lam, rows=(h // 2, h), nthreads=threads)
        one_row = max(time.perf_counter() - t, 1e-4)
        n_rows = int(max(1, min(h, budget / one_row)))
        step = max(1, h // n_rows)
        rows = (step // 2, step)
        dt = _median3(lambda: O.render(cam, prm, seq, lights, n_lights, w, h, mode=mode, lam=lam, rows=rows,
                                       nthreads=threads))
        return rows, len(range(rows[0], h, rows[1])) * w, dt

    rows, sampled, dt = rows_for(O.F32W, budget_s / 4)
    detail = {"host": _host_model(), "os_cpu_count": os.cpu_count(), "omp_threads": threads,
              "timing": "median of 3 after a warm-up"}
    r64, s64, d64 = rows_for(O.F64, budget_s / 16)
    detail["c2_fp64_rays_per_s"] = s64 / d64
    detail["c2_fp64_sample"] = f"rows {r64[0]}::{r64[1]} ({s64} rays), {d64:.2f} s"
    vprm, vseq, lo, hi, n = W.volume_config("c1")
    for name, lam in (("fp32_as_written", O.F32W), ("fp64", O.F64)):
        t = _median3(lambda: O.volume(vprm, vseq, lo, hi, n, mode=lam, nthreads=threads))
        detail[f"c1_volume_{name}_evals_per_s"] = n ** 3 / t
    vprm, vseq, lo, hi, n = W.volume_config("c3")
    t = time.perf_counter()
    O.volume(vprm, vseq, lo, hi, n, mode=O.F32W, z_range=(n // 2, n // 2 + 1), nthreads=threads)
    one = max(time.perf_counter() - t, 1e-4)
    nz = int(max(1, min(n, (budget_s / 16) / one)))
    z0 = (n - nz) // 2
    dv = _median3(lambda: O.volume(vprm, vseq, lo, hi, n, mode=O.F32W, z_range=(z0, z0 + nz), nthreads=threads))
    detail["c3_volume_fp32_as_written_evals_per_s"] = nz * n * n / dv
    detail["c3_volume_sample"] = f"z-slices {z0}..{z0 + nz - 1} of 256 ({nz * n * n} voxels), {dv:.2f} s"
    # the 64-view batch: every 1024th row of the views stacked, from a
    # seeded offset per view (a row of each 1024-pixel view, every fourth
    # 256-pixel view, ...)
    views, blights, bn = W.batch_views(64, seed=4)
    rng = np.random.default_rng(12)
    picks = []
    for v in views:
        o = int(rng.integers(0, 1024))
        if o < v["h"]:
            picks.append((o, v))
    c4_rays = sum(v["w"] for _, v in picks)
    all_rays = sum(v["w"] * v["h"] for v in views)

    def c4(part=None):
        for o, v in (part or picks):
            O.render(v["cam"], v["prm"], v["seq"], blights, bn, v["w"], v["h"], mode=mode, lam=O.F32W,
                     rows=(o, 1 << 20), nthreads=threads)
    d4 = _median3(c4, lambda: c4(picks[:2]))
    detail["c4_batch_rays_per_s"] = c4_rays / d4
    detail["c4_batch_sample"] = (f"one seeded row in 1024 of the 64 views ({len(picks)} rows, {c4_rays} of "
                                 f"{all_rays} rays, {c4_rays / all_rays:.5f}), {d4:.2f} s; extrapolated "
                                 f"batch time {all_rays / (c4_rays / d4):.0f} s")
    # the serving stream: a seeded sample of the capacity set's requests,
    # every 64th row of 8 views and 1/256 of the voxels of 12 volumes
    vq = [q for q in c5_requests if q["kind"] == "view"]
    wq = [q for q in c5_requests if q["kind"] == "volume"]
    rng = np.random.default_rng(13)
    vs = [vq[i] for i in sorted(rng.choice(len(vq), min(8, len(vq)), replace=False))]
    ws = [wq[i] for i in sorted(rng.choice(len(wq), min(12, len(wq)), replace=False))]
    pts = []
    for q in ws:
        k = max(1, q["n"] ** 3 // 256)
        idx = rng.integers(0, q["n"], size=(k, 3))
        c = [O.grid_coords(q["lo"][a], q["hi"][a], q["n"]) for a in range(3)]
        pts.append(np.stack([c[0][idx[:, 0]], c[1][idx[:, 1]], c[2][idx[:, 2]]], -1).astype(np.float32))
    v_rays = sum(len(range(0, q["h"], 64)) * q["w"] for q in vs)
    n_pts = sum(len(p) for p in pts)

    def c5_views(part):
        for q in part:
            O.render(q["cam"], q["prm"], q["seq"], lights, n_lights, q["w"], q["h"], mode=mode, lam=O.F32W,
                     rows=(0, 64), nthreads=threads)
    dv5 = _median3(lambda: c5_views(vs), lambda: c5_views(vs[:1]))
    dw5 = _median3(lambda: [O.lyap_points(p, q["prm"].d, q["prm"].settle, q["prm"].accum, q["seq"], mode=O.F32W,
                                          nthreads=threads) for p, q in zip(pts, ws)])
    rays_all = sum(q["w"] * q["h"] for q in c5_requests if q["kind"] == "view")
    vox_all = sum(q["n"] ** 3 for q in c5_requests if q["kind"] == "volume")
    detail["c5_view_rays_per_s"] = v_rays / dv5
    detail["c5_volume_evals_per_s"] = n_pts / dw5
    detail["c5_sample"] = (f"8 seeded view requests of the {len(c5_requests)}-request capacity set, every 64th "
                           f"row ({v_rays} rays, {dv5:.2f} s), and 1/256 of the voxels of 12 seeded volume "
                           f"requests ({n_pts} voxels, {dw5:.2f} s); extrapolated time for the whole set "
                           f"{rays_all / (v_rays / dv5) + vox_all / (n_pts / dw5):.0f} s")
    return {
        "value": sampled / dt, "unit": "rays/s", "cores": threads, "kind": "port",
        "sample": f"oracle FP32-as-written render of rows {rows[0]}::{rows[1]} "
                  f"({sampled} rays) of the same frame, median {dt:.2f} s of 3 on {threads} OpenMP threads",
        "detail": detail,
    }


def volume_stats(got, ref, h1, h2):
    stable = (ref < -1e-3) & (np.abs(h1 - h2) < 1e-4)
    d = np.abs(got[stable].astype(np.float64) - ref[stable])
    d = np.nan_to_num(d, nan=9.0)
    sign = np.mean((np.sign(got) == np.sign(ref)) | (np.isnan(got) & np.isnan(ref)))
    return {"stable_fraction": float(stable.mean()), "max_abs": float(d.max()) if d.size else 0.0,
            "rms": float(np.sqrt(np.mean(d ** 2))) if d.size else 0.0,
            "within_1e-3": float(np.mean(d <= 1e-3)) if d.size else 1.0, "sign_agreement": float(sign)}


RULES_C3 = {
    "plain": "FP32 x-form, one log per step (the reference's method), no re-evaluation",
    "product": "FP32 product form, no re-evaluation",
    "product_fp64_margin": "product form; FP64 where |lambda| < 0.01 (the configured rule)",
    "product_fp64_strict": "product form; FP64 where |lambda| < 0.01 or the window halves differ by "
                           "> 1e-4 with the smaller half < 0.1",
    "product_fp64_strict_wide": "product form; FP64 where |lambda| < 0.01 or the window halves differ by "
                                "> 1e-4 with the smaller half < 0.2",
}


def bench_volumes(lyap, timer, ctx, which):
    from tests import workloads as W
    from tests.gpu_util import empty, to_numpy
    out = {}
    if "c1" in which:
        prm, seq, lo, hi, n = W.volume_config("c1")
        buf = empty(4 * n ** 3)
        res = {}
        for name, mode in [("plain", lyap.volume_mode(lyap.PREC_PLAIN)),
                           ("product", lyap.volume_mode(lyap.PREC_PRODUCT))]:
            ms = timer.run(lambda: ctx.calc_volume(buf.data_ptr(), prm, seq, lo, hi, n, mode), 20)
            res[name] = {"ms": ms, "lambda_evals_per_s": n ** 3 / (ms * 1e-3),
                         "tflops": n ** 3 * flops(prm) / (ms * 1e-3) / 1e12}
        out["c1_volume_64"] = res
        log("c1 done")
    if "c3" in which:
        prm, seq, lo, hi, n = W.volume_config("c3")
        buf = empty(4 * n ** 3)

        def fp64_volume(p):
            rb = empty(4 * n ** 3)
            ctx.calc_volume(rb.data_ptr(), p, seq, lo, hi, n, lyap.volume_mode(lyap.PREC_FP64))
            ctx.sync()
            v = to_numpy(rb, np.float32)[:n ** 3].astype(np.float64)
            del rb
            return v

        # References over all n^3 voxels from the FP64 form, the reference
        # orbit bit for bit: lambda, and lambda of the first half of the
        # accumulation window (same settle, accum/2 steps); the second half's
        # lambda follows from the two, h2 = (accum*lam - split*h1)/(accum - split).
        # Stable voxels: FP64 lambda < -1e-3 with the halves within 1e-4.
        ref_all = fp64_volume(prm)
        split = prm.accum // 2
        prm_half = W.volume_config("c3")[0]
        prm_half.accum = split
        h1 = fp64_volume(prm_half)
        h2 = (prm.accum * ref_all - split * h1) / (prm.accum - split)
        res = {}
        plain_ms = None
        for name, mode in [("plain", lyap.volume_mode(lyap.PREC_PLAIN)),
                           ("product", lyap.volume_mode(lyap.PREC_PRODUCT)),
                           ("product_fp64_margin", lyap.volume_mode(**W.C3_MODE)),
                           ("product_fp64_strict", lyap.volume_mode(**W.C3_STRICT)),
                           ("product_fp64_strict_wide", lyap.volume_mode(**W.C3_STRICT_WIDE))]:
            ms = timer.run(lambda: ctx.calc_volume(buf.data_ptr(), prm, seq, lo, hi, n, mode), 5)
            re = ctx.last_reevaluated()
            vol = to_numpy(buf, np.float32)[:n ** 3]
            st = volume_stats(vol, ref_all, h1, h2)
            r = {"ms": ms, "lambda_evals_per_s": n ** 3 / (ms * 1e-3),
                 "tflops": n ** 3 * flops(prm) / (ms * 1e-3) / 1e12,
                 "roofline_frac": n ** 3 * flops(prm) / (ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
                 "fp64_reevaluated_fraction": re / n ** 3, "vs_fp64": st}
            if name == "plain":
                plain_ms = ms
            else:
                r["speedup_vs_plain"] = plain_ms / ms
            r["rule"] = RULES_C3[name]
            res[name] = r
        # the headline C3 figure: the cheapest rule that meets both FP64 bars
        # (99.9 % of stable voxels within 1e-3, 99.9 % sign agreement); the
        # reference's own FP32 arithmetic reaches neither on this volume
        ok = [k for k, r in res.items() if r["vs_fp64"]["within_1e-3"] >= 0.999
              and r["vs_fp64"]["sign_agreement"] >= 0.999]
        res["headline"] = min(ok, key=lambda k: res[k]["ms"]) if ok else None
        out["c3_volume_256"] = res
        log("c3 done")
    return out


def depth_stats(lyap_depth_got, lyap_depth_ref, step):
    """Hit/miss agreement and, where both hit, the fraction with depth
    within one march step (arrays of equal shape; step per ray)."""
    hg, hr = ~np.isnan(lyap_depth_got), ~np.isnan(lyap_depth_ref)
    both = hg & hr
    within = np.abs(lyap_depth_got[both] - lyap_depth_ref[both]) <= step[both] * (1 + 1e-5)
    return {"hit_agreement": float(np.mean(hg == hr)), "hit_fraction": float(np.mean(hr)),
            "depth_within_one_step": float(np.mean(within)) if within.size else 1.0}


def ray_steps(v):
    """March step (t1 - t0) / depth of every pixel ray of a view (stepMethod 1),
    from the camera: NaN where the ray misses the box."""
    cam, prm, w, h = v["cam"], v["prm"], v["w"], v["h"]
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float32)
    S0 = np.array([cam.S0.x, cam.S0.y, cam.S0.z], np.float32)
    DX = np.array([cam.SDX.x, cam.SDX.y, cam.SDX.z], np.float32)
    DY = np.array([cam.SDY.x, cam.SDY.y, cam.SDY.z], np.float32)
    V = S0 + xs[..., None] * DX + ys[..., None] * DY
    V = V / np.linalg.norm(V, axis=-1, keepdims=True) / np.float32(cam.M)
    C = np.array([cam.C.x, cam.C.y, cam.C.z], np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        t_lo = (prm.lMin - C) / V
        t_hi = (prm.lMax - C) / V
    t0 = np.nanmax(np.minimum(t_lo, t_hi), axis=-1)
    t1 = np.nanmin(np.maximum(t_lo, t_hi), axis=-1)
    t0 = np.maximum(t0, 0.0)
    return ((t1 - t0) / prm.depth).astype(np.float64)


def bench_batch(lyap, timer, ctx, n_views):
    """The 64-view batch: default launch shapes, the autotuner's, a captured
    graph, plain per-request calls, and the accuracy of the benchmarked mode
    against the FP64 batch (exact against the FP64 oracle)."""
    from tests import workloads as W
    from tests.gpu_util import RenderBuffers, device_lights
    views, lights, n_lights = W.batch_views(n_views, seed=4)
    ld = device_lights(lights, n_lights)
    bufs = [RenderBuffers(v["w"], v["h"]) for v in views]
    api = [lyap.make_view(v["cam"], v["prm"], v["seq"], v["w"], v["h"], b.rgba.data_ptr(), b.points.data_ptr(),
                          b.depth.data_ptr(), b.lam.data_ptr(), b.evals.data_ptr(), ld.data_ptr(), n_lights)
           for v, b in zip(views, bufs)]
    rays = sum(v["w"] * v["h"] for v in views)
    mode = W.c2_mode(fp64_margin=0.0)
    ctx.render_batch(api, mode, lyap.Tuning())
    ctx.sync()
    default_ms = timer.run(lambda: ctx.render_batch(api, mode, lyap.Tuning()), 3)
    log("c4: autotuning the batch")
    t = time.perf_counter()
    tuned = ctx.autotune(api, mode)
    tune_s = time.perf_counter() - t
    ms = timer.run(lambda: ctx.render_batch(api, mode), 3)
    evals = ctx.last_evals()
    flop = 0
    # algorithmic FLOP: per-view evaluations x per-view FLOP per evaluation
    for v, b in zip(views, bufs):
        flop += int(b.evals.cpu().numpy().view(np.uint32).astype(np.int64).sum()) * flops(v["prm"])
    got = [b.depth.cpu().numpy().view(np.float32).reshape(v["h"], v["w"]).copy() for v, b in zip(views, bufs)]
    frame = ctx.frame_capture(api, mode)
    graph_ms = timer.run(lambda: frame.replay(), 3)
    frame.close()
    # accuracy: the same batch in FP64 (the reference orbit, exact against the FP64 oracle)
    ctx.render_batch(api, W.c2_mode(precision=lyap.PREC_FP64, fp64_margin=0.0))
    ctx.sync()
    ref = [b.depth.cpu().numpy().view(np.float32).reshape(v["h"], v["w"]) for v, b in zip(views, bufs)]
    steps = [ray_steps(v) for v in views]
    acc = depth_stats(np.concatenate([g.ravel() for g in got]), np.concatenate([r.ravel() for r in ref]),
                      np.concatenate([s.ravel() for s in steps]))
    # the same views one request at a time, plain FP32 per-step log, in the
    # reference's launch shape and in the batched schedule
    def plain_all(schedule):
        m = W.c2_mode(precision=lyap.PREC_PLAIN, fp64_margin=0.0)
        m.schedule = schedule
        for v, b in zip(views, bufs):
            b.render(ctx, v["cam"], v["prm"], v["seq"], ld, n_lights, m)
    plain_ms = timer.run(lambda: plain_all(lyap.SCHEDULE_DIRECT), 1)
    plain_batched_ms = timer.run(lambda: plain_all(lyap.SCHEDULE_BATCHED), 1)
    ctx.set_tuning(None)          # later workloads use the built-in shapes unless they tune
    log("c4 done")
    return {"views": n_views, "rays": rays, "ms": ms, "graph_replay_ms": graph_ms,
            "default_shape_ms": default_ms, "tuned_over_default": default_ms / ms,
            "tuning": tuned.as_dict(), "autotune_s": tune_s,
            "rays_per_s": rays / (ms * 1e-3), "lambda_evals_per_s": evals / (ms * 1e-3),
            "tflops": flop / (ms * 1e-3) / 1e12, "roofline_frac": flop / (ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
            "plain_per_request_ms": plain_ms, "plain_per_request_rays_per_s": rays / (plain_ms * 1e-3),
            "speedup_vs_plain_per_request": plain_ms / ms,
            "plain_batched_schedule_ms": plain_batched_ms,
            "speedup_vs_plain_batched_schedule": plain_batched_ms / ms,
            "vs_fp64": acc,
            "mode": "product form, 8 bisections, no FP64 margin; tuned shape from lyap_autotune on this batch"}


def bench_mixed(lyap, timer, ctx):
    """The 64-view batch and eight 128^3 volume blocks of the precision
    workload's parameters (the configured FP64 rule) in one persistent launch
    sequence, against the views' batch followed by the volumes' own kernels."""
    from tests import workloads as W
    from tests.gpu_util import RenderBuffers, device_lights, empty
    views, lights, n_lights = W.batch_views(64, seed=4)
    ld = device_lights(lights, n_lights)
    bufs = [RenderBuffers(v["w"], v["h"]) for v in views]
    api = [lyap.make_view(v["cam"], v["prm"], v["seq"], v["w"], v["h"], b.rgba.data_ptr(), b.points.data_ptr(),
                          b.depth.data_ptr(), b.lam.data_ptr(), b.evals.data_ptr(), ld.data_ptr(), n_lights)
           for v, b in zip(views, bufs)]
    prm, seq, lo, hi, _ = W.volume_config("c3")
    vmode = lyap.volume_mode(**W.C3_MODE)
    n = 128
    outs = [empty(4 * n ** 3) for _ in range(8)]
    rng = np.random.default_rng(8)
    boxes = []
    for _ in range(8):
        a = [float(rng.uniform(2.5, 3.25)) for _ in range(3)]
        boxes.append((tuple(a), tuple(x + 0.75 for x in a)))
    blocks = [lyap.make_block(prm, seq, b[0], b[1], n, o.data_ptr()) for b, o in zip(boxes, outs)]
    mode = W.c2_mode(fp64_margin=0.0)
    fused = timer.run(lambda: ctx.run_batch(api, blocks, mode, vmode), 3)

    def apart():
        ctx.render_batch(api, mode)
        for b, o in zip(boxes, outs):
            ctx.calc_volume(o.data_ptr(), prm, seq, b[0], b[1], n, vmode)
    separate = timer.run(apart, 3)
    voxels = 8 * n ** 3
    log("mixed batch done")
    return {"views": 64, "rays": sum(v["w"] * v["h"] for v in views), "volume_blocks": 8, "voxels": voxels,
            "fused_ms": fused, "views_then_volumes_ms": separate, "fused_speedup": separate / fused,
            "config": "64 views of the batched workload (product form, no FP64 margin) and eight 128^3 blocks "
                      "over seeded 0.75-wide cubes in [2.5,4]^3 with the precision workload's parameters "
                      "(ABCABCCA, 200 + 2000, product form, FP64 where |lambda| < 0.01)"}


def _pct(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if len(xs) else None


def _make_requests(lyap, reqs, ld, n_lights, view_deadline_ms, volume_deadline_ms):
    out = []
    for q in reqs:
        if q["kind"] == "view":
            v = lyap.make_view(q["cam"], q["prm"], q["seq"], q["w"], q["h"], 0, 0, 0, 0, 0,
                               ld.data_ptr(), n_lights)
            out.append(lyap.view_request(v, deadline_ms=view_deadline_ms))
        else:
            out.append(lyap.volume_request(q["prm"], q["seq"], q["lo"], q["hi"], q["n"],
                                           deadline_ms=volume_deadline_ms))
    return out


def _paced(lyap, srv, reqs, api, label, seconds):
    """Submits `api` at the stream's arrival times, drives the server until
    every request finished; latency from arrival to finish, per kind."""
    t0 = srv.clock_ms()
    arrival, index = {}, {}
    lat = {"view": [], "volume": [], "view_1024": [], "view_512": [], "view_256": []}
    i, last_finish, n_done, next_log = 0, t0, 0, 10000.0
    st0 = srv.stats()
    while n_done < len(api):
        now = srv.clock_ms() - t0
        while i < len(api) and reqs[i]["t_ms"] <= now:
            t = srv.submit(api[i])
            arrival[t] = t0 + reqs[i]["t_ms"]
            index[t] = i
            i += 1
        srv.progress(0.2)
        for t in srv.finished():
            r = srv.poll(t)
            if r.status != lyap.OK:
                raise RuntimeError(f"request {index[t]} ended with status {r.status}")
            q = reqs[index[t]]
            ms = r.finish_ms - arrival[t]
            lat[q["kind"]].append(ms)
            if q["kind"] == "view":
                lat[f"view_{q['w']}"].append(ms)
            last_finish = max(last_finish, r.finish_ms)
            srv.release(t)
            n_done += 1
        if now >= next_log:
            log(f"{label}: {now / 1e3:.0f} s, {n_done} of {len(api)} requests done")
            next_log += 10000.0
    st1 = srv.stats()
    wall_s = (last_finish - t0) * 1e-3
    view_rays = sum(q["w"] * q["h"] for q in reqs if q["kind"] == "view")
    voxels = sum(q["n"] ** 3 for q in reqs if q["kind"] == "volume")
    return {
        "requests": len(api), "stream_seconds": seconds, "wall_seconds": wall_s,
        "offered_rays_per_s": view_rays / seconds, "offered_voxels_per_s": voxels / seconds,
        "served_rays_per_s": view_rays / wall_s, "served_voxels_per_s": voxels / wall_s,
        "lambda_evals_per_s": (st1.evals - st0.evals) / wall_s,
        "device_busy_ms": st1.device_ms - st0.device_ms,
        "launches": st1.launches - st0.launches,
        "peak_resident_bytes": st1.peak_resident_bytes,
        "latency_ms": {k: {"n": len(v), "p50": _pct(v, 50), "p99": _pct(v, 99), "max": max(v) if v else None}
                       for k, v in lat.items()},
    }


def _offered_at_once(srv, api):
    """Seconds for the server to finish `api` submitted all at once, and the
    lambda evaluations it counted."""
    e0 = srv.stats().evals
    t = time.perf_counter()
    for r in api:
        srv.submit(r)
    srv.drain(600000.0)
    dt = time.perf_counter() - t
    for tk in srv.finished():
        srv.release(tk)
    return dt, srv.stats().evals - e0


def _c5_accuracy(lyap, ctx, reqs, ld, n_lights, mode, vmode, n_views, n_vols):
    """The served mode's results against FP64 on a seeded sample of the
    stream's requests, computed alone (the server gives the same bytes):
    views' hit/miss and depth within one step, volumes' C3-style stats."""
    import torch
    from tests import workloads as W
    from tests.gpu_util import RenderBuffers
    rng = np.random.default_rng(6)
    vi = [k for k, q in enumerate(reqs) if q["kind"] == "view"]
    wi = [k for k, q in enumerate(reqs) if q["kind"] == "volume"]
    vi = sorted(rng.choice(vi, min(n_views, len(vi)), replace=False).tolist())
    wi = sorted(rng.choice(wi, min(n_vols, len(wi)), replace=False).tolist())
    got, ref, steps = [], [], []
    for k in vi:
        q = reqs[k]
        b = RenderBuffers(q["w"], q["h"])
        for m, out in ((mode, got), (W.c2_mode(precision=lyap.PREC_FP64, fp64_margin=0.0), ref)):
            b.render(ctx, q["cam"], q["prm"], q["seq"], ld, n_lights, m)
            ctx.sync()
            out.append(b.depth.cpu().numpy().view(np.float32).copy())
        steps.append(ray_steps(q).ravel())
    views = depth_stats(np.concatenate(got), np.concatenate(ref), np.concatenate(steps))
    gv, rv = [], []
    for k in wi:
        q = reqs[k]
        n3 = q["n"] ** 3
        a = torch.empty(n3, dtype=torch.float32, device="cuda")
        for m, out in ((vmode, gv), (lyap.volume_mode(lyap.PREC_FP64), rv)):
            ctx.calc_volume(a.data_ptr(), q["prm"], q["seq"], q["lo"], q["hi"], q["n"], m)
            ctx.sync()
            out.append(a.cpu().numpy().astype(np.float64))
    g, r = np.concatenate(gv), np.concatenate(rv)
    d = np.abs(g - r)
    sign = float(np.mean((np.sign(g) == np.sign(r)) | (np.isnan(g) & np.isnan(r))))
    settled = r < -1e-3
    within = float(np.mean(np.nan_to_num(d[settled], nan=9.0) <= 1e-3)) if settled.any() else 1.0
    return {"views": views, "views_sampled": len(vi),
            "volumes": {"within_1e-3_where_fp64_below_-1e-3": within, "sign_agreement": sign,
                        "max_abs": float(np.nanmax(d)), "rms": float(np.sqrt(np.nanmean(d ** 2))),
                        "voxels": int(g.size)},
            "volumes_sampled": len(wi)}


def bench_serving(lyap, timer, ctx, seconds, strict_seconds, sat_n):
    """The serving workload: a paced Poisson stream through the server
    (latency), the same stream's head with volumes under the rule that meets
    the FP64 tolerance, then the first `sat_n` requests offered at once to
    the server (launches carrying views and volume slabs together, with the
    default and the autotuned shapes, and with volumes on slots of their own)
    and to plain per-request calls (capacity)."""
    import torch
    from tests import workloads as W
    from tests.gpu_util import RenderBuffers, device_lights, empty
    reqs, lights, n_lights = W.serving_stream(seconds=seconds)
    ld = device_lights(lights, n_lights)
    mode = W.c2_mode(fp64_margin=0.0)
    rules = {"margin": lyap.volume_mode(**W.C3_MODE), "strict_wide": lyap.volume_mode(**W.C3_STRICT_WIDE)}
    api = _make_requests(lyap, reqs, ld, n_lights, 50.0, 250.0)
    torch.cuda.synchronize()

    def server(rule, flags=0):
        cfg = lyap.server_config(hbm_budget_bytes=64 << 30, mode=mode, vol_mode=rules[rule], flags=flags)
        srv = lyap.Server(ctx, cfg)
        for r in api[:12]:                 # warm-up: calibrates the cost rates
            srv.submit(r)
        srv.drain(60000.0)
        for t in srv.finished():
            srv.release(t)
        return srv

    out = {}
    log(f"c5 paced: {len(api)} requests over {seconds:.0f} s")
    srv = server("margin")
    out["paced"] = _paced(lyap, srv, reqs, api, "c5 paced", seconds)
    srv.close()
    log(f"c5 paced done in {out['paced']['wall_seconds']:.1f} s")
    head = [q for q in reqs if q["t_ms"] < strict_seconds * 1000.0]
    log(f"c5 paced, tolerance rule: {len(head)} requests over {strict_seconds:.0f} s")
    srv = server("strict_wide")
    out["paced_strict_wide"] = _paced(lyap, srv, head, api[:len(head)], "c5 paced strict", strict_seconds)
    srv.close()

    # ---- capacity: the first sat_n requests offered at once
    sub = api[:sat_n]
    sub_q = reqs[:sat_n]
    sub_rays = sum(q["w"] * q["h"] for q in sub_q if q["kind"] == "view")
    sub_vox = sum(q["n"] ** 3 for q in sub_q if q["kind"] == "volume")
    cap = {"requests": len(sub), "view_rays": sub_rays, "voxels": sub_vox}
    for name, rule, flags in (("served", "margin", 0), ("served_strict_wide", "strict_wide", 0),
                              ("served_separate_volume_slots", "margin", lyap.SERVER_SEPARATE_VOLUMES)):
        srv = server(rule, flags)
        dt, ev = _offered_at_once(srv, sub)
        srv.close()
        cap[f"{name}_s"] = dt
        cap[f"{name}_rays_per_s"] = sub_rays / dt
        cap[f"{name}_lambda_evals_per_s"] = ev / dt
        log(f"c5 capacity, {name}: {dt:.2f} s")
    # autotuned shapes: the batch autotuner on the first 48 requests as one
    # launch (views and volume blocks), then the server with the tuned context
    tq = sub_q[:48]
    vb = [RenderBuffers(q["w"], q["h"]) for q in tq if q["kind"] == "view"]
    tviews, vi = [], 0
    for q in tq:
        if q["kind"] == "view":
            b = vb[vi]
            vi += 1
            tviews.append(lyap.make_view(q["cam"], q["prm"], q["seq"], q["w"], q["h"], b.rgba.data_ptr(), 0, 0,
                                         b.lam.data_ptr(), 0, ld.data_ptr(), n_lights))
    tvol = [q for q in tq if q["kind"] == "volume"]
    vbufs = [empty(4 * q["n"] ** 3) for q in tvol]
    tblocks = [lyap.make_block(q["prm"], q["seq"], q["lo"], q["hi"], q["n"], b.data_ptr())
               for q, b in zip(tvol, vbufs)]
    t = time.perf_counter()
    tuned = ctx.autotune_batch(tviews, tblocks, mode, rules["margin"])
    cap["tuning"] = tuned.as_dict()
    cap["autotune_s"] = time.perf_counter() - t
    del vb, vbufs
    srv = server("margin")
    dt, ev = _offered_at_once(srv, sub)
    srv.close()
    ctx.set_tuning(None)
    cap["served_tuned_s"] = dt
    cap["served_tuned_rays_per_s"] = sub_rays / dt
    log(f"c5 capacity, tuned shapes: {dt:.2f} s")

    # plain per-request: each request its own call, one after another, in
    # the plain FP32 per-step-log form; views in the reference's launch shape
    # (one lane per ray) and, for comparison, in the batched schedule
    vol_out = empty(4 * 256 ** 3)
    bufs = {}

    def plain(schedule):
        m = W.c2_mode(precision=lyap.PREC_PLAIN, fp64_margin=0.0)
        m.schedule = schedule
        vm = lyap.volume_mode(lyap.PREC_PLAIN)
        evals = 0
        torch.cuda.synchronize()
        t = time.perf_counter()
        for q in sub_q:
            if q["kind"] == "view":
                b = bufs.get(q["w"])
                if b is None:
                    b = bufs[q["w"]] = RenderBuffers(q["w"], q["h"])
                b.render(ctx, q["cam"], q["prm"], q["seq"], ld, n_lights, m)
                ctx.sync()
                evals += ctx.last_evals()
            else:
                ctx.calc_volume(vol_out.data_ptr(), q["prm"], q["seq"], q["lo"], q["hi"], q["n"], vm)
                ctx.sync()
                evals += q["n"] ** 3
        return time.perf_counter() - t, evals

    direct_s, direct_evals = plain(lyap.SCHEDULE_DIRECT)
    log(f"c5 plain per-request (direct): {direct_s:.2f} s")
    batched_s, batched_evals = plain(lyap.SCHEDULE_BATCHED)
    log(f"c5 plain per-request (batched schedule): {batched_s:.2f} s")
    cap.update({
        "plain_per_request_s": direct_s, "plain_per_request_rays_per_s": sub_rays / direct_s,
        "plain_per_request_lambda_evals_per_s": direct_evals / direct_s,
        "served_over_plain_rays_per_s": direct_s / cap["served_s"],
        "served_strict_wide_over_plain_rays_per_s": direct_s / cap["served_strict_wide_s"],
        "plain_batched_schedule_s": batched_s,
        "plain_batched_schedule_rays_per_s": sub_rays / batched_s,
        "served_over_plain_batched_schedule": batched_s / cap["served_s"],
    })
    out["capacity"] = cap
    log("c5 accuracy sample")
    out["accuracy_vs_fp64"] = {name: _c5_accuracy(lyap, ctx, sub_q, ld, n_lights, mode, rules[name], 24, 24)
                               for name in rules}
    p99 = lambda leg: (leg["latency_ms"]["view_1024"]["p99"] or 0.0)
    out["targets"] = {
        "served_over_plain_at_least_3x": cap["served_over_plain_rays_per_s"] >= 3.0,
        "served_strict_wide_over_plain_at_least_3x": cap["served_strict_wide_over_plain_rays_per_s"] >= 3.0,
        "p99_1024_le_50ms": p99(out["paced"]) <= 50.0,
        "p99_1024_le_50ms_strict_wide": p99(out["paced_strict_wide"]) <= 50.0,
    }
    out["config"] = ("views: product form, 8 bisections, no FP64 margin; volumes: product form with FP64 "
                     "re-evaluation where |lambda| < 0.01 ('margin'), or also where the window halves differ "
                     "by > 1e-4 below 0.2 ('strict_wide', the rule that meets the FP64 tolerance on the 256^3 "
                     "volume); launches carry views and volume slabs together unless named 'separate'; "
                     "deadlines 50 ms (views) / 250 ms (volumes); 64 GiB budget; plain: per-step-log FP32, "
                     "one synchronous call per request, views in the reference's launch shape (one lane per "
                     "ray, 8x8 blocks)")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; volumes and views shard by z slab "
                                   "or by request across GPUs, which is not run here",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    import lyapunov3d_amd as lyap
    from tests import workloads as W
    from tests.gpu_util import RenderBuffers, device_lights

    prec = {"product": lyap.PREC_PRODUCT, "plain": lyap.PREC_PLAIN, "fp64": lyap.PREC_FP64}[args.precision]
    w = h = args.size
    cam, prm, seq, lights, n_lights = W.render_config(w, h)
    mode = W.c2_mode(precision=prec, fp64_margin=0.0)
    ctx = lyap.Context(0)
    timer = Timer(torch, ctx)
    bufs = RenderBuffers(w, h)
    ld = device_lights(lights, n_lights)

    if args.copy_calibration:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        log("copy calibration: 3 x 1 GiB device copies")

    # the frame with the built-in launch shape, then the autotuner's choice
    # for this frame (kept in the context for the timed loop)
    default_frame_ms = timer.run(lambda: bufs.render(ctx, cam, prm, seq, ld, n_lights, mode), 20)
    tuned = None
    if not args.no_tune:
        view = lyap.make_view(cam, prm, seq, w, h, bufs.rgba.data_ptr(), bufs.points.data_ptr(),
                              bufs.depth.data_ptr(), bufs.lam.data_ptr(), bufs.evals.data_ptr(),
                              ld.data_ptr(), n_lights)
        tuned = ctx.autotune([view], mode)
        log(f"autotuned frame shape: {tuned.as_dict()}")
    log(f"warmup {args.warmup} frames")
    for _ in range(args.warmup):
        bufs.render(ctx, cam, prm, seq, ld, n_lights, mode)
    ctx.sync()
    evals_per_frame = ctx.last_evals()

    torch.cuda.synchronize()
    # Every frame does its full work; one in four is bracketed by two events
    # on the library stream (the per-frame percentiles). The throughput is the
    # wall clock over all of them. The roofline's launch durations come from a
    # sample of frames after the timed region, with the library's events
    # around each launch (each event costs the device a few microseconds, so
    # they stay out of the timed frames).
    sampled = [i for i in range(args.steps) if i % 4 == 2] or [0]
    starts = {i: torch.cuda.Event(enable_timing=True) for i in sampled}
    ends = {i: torch.cuda.Event(enable_timing=True) for i in sampled}
    t0 = time.perf_counter()
    for i in range(args.steps):
        if i in starts:
            starts[i].record(timer.stream)
        bufs.render(ctx, cam, prm, seq, ld, n_lights, mode)
        if i in ends:
            ends[i].record(timer.stream)
    ctx.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    log(f"timed {args.steps} frames in {wall:.3f} s")
    per = np.array([starts[i].elapsed_time(ends[i]) for i in sampled if i < args.steps])
    ms = float(wall * 1000.0 / args.steps)
    timed_frames = max(10, args.steps // 4)
    ctx.timing_collect()                      # start from an empty launch record
    ctx.timing(True)
    for _ in range(timed_frames):
        bufs.render(ctx, cam, prm, seq, ld, n_lights, mode)
    ctx.timing(False)
    kernel_ms, launches = ctx.timing_collect()

    rays = w * h
    frame_ms = float(per.mean())
    # the render kernels: the march (a set-up launch and seven one-round
    # passes on this frame, or four persistent levels) and the refinement
    # launch; algorithmic FLOP of a frame over the summed duration of its
    # launches (HIP events on the library stream around each launch)
    launch_ms = kernel_ms / max(launches, 1)
    launches_per_frame = launches / timed_frames
    achieved = evals_per_frame * flops(prm) / launches_per_frame / (launch_ms * 1e-3) / 1e12
    traffic, traffic_src = profiled_traffic("render_frame")
    line = {
        "metric": "lambda-evals/s & rays/s on 1 MI355X; p50/p99 latency, served vs plain",
        "value": rays / (ms * 1e-3),
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec != lyap.PREC_FP64 else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"plain render {w}x{h}, seq {seq}, settle {prm.settle} accum {prm.accum}, "
                        f"500 uniform steps, 8 bisections, h=1e-3, Phong, camera seed 1",
            "precision": args.precision,
        },
        "lambda_evals_per_s": evals_per_frame / (ms * 1e-3),
        "lambda_evals_per_frame": int(evals_per_frame),
        "launch_shape": tuned.as_dict() if tuned is not None else "built-in",
        "frame_ms_builtin_shape": default_frame_ms,
        "frame_ms_events_mean": frame_ms,
        "frame_ms_p50": float(np.percentile(per, 50)),
        "frame_ms_p99": float(np.percentile(per, 99)),
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "flop_per_lambda_eval": flops(prm),
            "kernel": ("ray_start + march_pass x 7 + refine_rays" if round(launches_per_frame) == 9
                       else "render_persistent (4 levels) + refine_rays"),
            "launches_per_frame": launches_per_frame,
            "avg_launch_us": launch_ms * 1e3,
            "flop_per_launch": evals_per_frame * flops(prm) / launches_per_frame,
            "note": "achieved = algorithmic FLOP per launch (lambda evaluations x (3 settle + 6 accum)) / "
                    "mean launch duration from HIP events on the library stream, over the frame's "
                    f"{launches_per_frame:.0f} launches; traffic = HBM bytes per launch (mean over the "
                    "frame's launches) from rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a 1 GiB copy",
        },
    }
    ctx.set_tuning(None)          # the other workloads tune for themselves or use built-in shapes
    which = [x for x in args.workloads.split(",") if x]
    if which:
        extra = {}
        if "c2" in which:
            # the same frame in the other arithmetic forms
            c2 = {}
            for name, m in [("fp32_product_fp64_bisect_margin", W.c2_mode(fp64_margin=0.01)),
                            ("fp64_exact", W.c2_mode(precision=lyap.PREC_FP64, fp64_margin=0.0)),
                            ("fp32_plain", W.c2_mode(precision=lyap.PREC_PLAIN, fp64_margin=0.0))]:
                fms = timer.run(lambda: bufs.render(ctx, cam, prm, seq, ld, n_lights, m), 10)
                c2[name] = {"ms": fms, "rays_per_s": rays / (fms * 1e-3),
                            "lambda_evals_per_s": ctx.last_evals() / (fms * 1e-3)}
            # a light edit: the frame's points re-shaded instead of re-rendered
            bufs.render(ctx, cam, prm, seq, ld, n_lights, mode)
            rms = timer.run(lambda: ctx.shade_points(bufs.rgba.data_ptr(), bufs.points.data_ptr(), cam,
                                                     ld.data_ptr(), n_lights, w, h), 50)
            c2["reshade_for_new_lights"] = {"ms": rms, "pixels_per_s": rays / (rms * 1e-3)}
            extra["c2_variants"] = c2
            log("c2 variants done")
        extra.update(bench_volumes(lyap, timer, ctx, which))
        if "c4" in which:
            extra["c4_batched"] = bench_batch(lyap, timer, ctx, args.c4_views)
            extra["mixed_views_and_volumes"] = bench_mixed(lyap, timer, ctx)
        if "c5" in which:
            extra["c5_serving"] = bench_serving(lyap, timer, ctx, args.c5_seconds, args.c5_strict_seconds,
                                                args.c5_capacity)
        line["workloads"] = extra
    if not args.no_cpu:
        log("cpu baseline")
        c5_requests = W.serving_stream(seconds=args.c5_seconds)[0][:args.c5_capacity]
        line["cpu_baseline"] = cpu_baseline(cam, prm, seq, lights, n_lights, w, h, mode, args.cpu_seconds,
                                            c5_requests)
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
