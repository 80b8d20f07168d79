#!/usr/bin/env python3
"""Throughput of the Lyapunov raymarch on one MI355X.

Workload (one step = one frame): the plain render, a 512x512 image of the
[2,4]^3 Lyapunov box seen from distance 4 along a seeded unit quaternion,
45 degree field of view, sequence AABAB, 100 settle + 300 accumulated
iterations per lambda, up to 500 uniform steps per ray, surface at the first
lambda <= 0, 8 bisection steps, normal from six samples at h = 1e-3, Phong
shading, outputs RGBA8 + depth + lambda + LyapPoint per pixel.

The JSON line reports rays/s (value) and lambda-evals/s, the roofline of the
render kernel (algorithmic FLOP per lambda-eval = 3*settle + 6*accum, against
MI355X's FP32 vector peak) and the oracle's FP32 CPU path timed on a bounded
sample of rows on this host.
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

FP32_PEAK_TFLOPS = 157.3     # MI355X dense FP32 vector (packed)
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--precision", default="product", choices=["product", "plain", "fp64"])
    p.add_argument("--size", type=int, default=512)
    p.add_argument("--cpu-seconds", type=float, default=12.0,
                   help="approximate CPU budget of the baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    return p.parse_args()


def cpu_baseline(cam, prm, seq, lights, n_lights, w, h, mode, budget_s):
    """The oracle's FP32 as-written render on every k-th row, timed here."""
    from oracle import oracle as O
    threads = O.threads()
    # probe one row to size the sample
    t = time.perf_counter()
    O.render(cam, prm, seq, lights, n_lights, w, h, mode=mode, lam=O.F32W, rows=(h // 2, h), nthreads=threads)
    one_row = max(time.perf_counter() - t, 1e-4)
    n_rows = int(max(1, min(h, budget_s / one_row)))
    step = max(1, h // n_rows)
    rows = (step // 2, step)
    t = time.perf_counter()
    O.render(cam, prm, seq, lights, n_lights, w, h, mode=mode, lam=O.F32W, rows=rows, nthreads=threads)
    dt = time.perf_counter() - t
    sampled = len(range(rows[0], h, rows[1])) * w
    return {
        "value": sampled / dt, "unit": "rays/s", "cores": threads, "kind": "port",
        "sample": f"oracle FP32-as-written render of rows {rows[0]}::{rows[1]} "
                  f"({sampled} rays) of the same frame, {dt:.1f} s on {threads} OpenMP threads",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    import lyapunov3d_amd as lyap
    from tests import workloads as W
    from tests.gpu_util import RenderBuffers, device_lights

    prec = {"product": lyap.PREC_PRODUCT, "plain": lyap.PREC_PLAIN, "fp64": lyap.PREC_FP64}[args.precision]
    w = h = args.size
    cam, prm, seq, lights, n_lights = W.render_config(w, h)
    mode = W.c2_mode(precision=prec)
    ctx = lyap.Context(0)
    bufs = RenderBuffers(w, h)
    ld = device_lights(lights, n_lights)
    stream = torch.cuda.ExternalStream(ctx.stream)

    print(f"bench: warmup {args.warmup} frames", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        bufs.render(ctx, cam, prm, seq, ld, n_lights, mode)
    ctx.sync()
    print("bench: warmup done", file=sys.stderr, flush=True)
    evals_per_frame = ctx.last_evals()

    torch.cuda.synchronize()
    ctx.sync()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        bufs.render(ctx, cam, prm, seq, ld, n_lights, mode)
        ends[i].record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    print(f"bench: timed {args.steps} frames in {wall:.3f} s", file=sys.stderr, flush=True)
    per = np.array([s.elapsed_time(e) for s, e in zip(starts, ends)])   # ms
    ms = float(wall * 1000.0 / args.steps)

    rays = w * h
    flop_per_eval = 3 * prm.settle + 6 * prm.accum
    kernel_ms = float(per.mean())
    achieved = evals_per_frame * flop_per_eval / (kernel_ms * 1e-3) / 1e12
    out_bytes = rays * (4 + 36 + 4 + 4 + 4)
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
                        f"500 uniform steps, 8 bisections, h=1e-3, Phong",
            "precision": args.precision,
            "camera_seed": 1,
        },
        "lambda_evals_per_s": evals_per_frame / (ms * 1e-3),
        "lambda_evals_per_frame": int(evals_per_frame),
        "kernel_ms_mean": kernel_ms,
        "kernel_ms_p50": float(np.percentile(per, 50)),
        "kernel_ms_p99": float(np.percentile(per, 99)),
        "roofline": {
            "bound": "valu-fp32",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": None,
            "flop_per_lambda_eval": flop_per_eval,
            "output_bytes_per_frame": out_bytes,
        },
    }
    if not args.no_cpu:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(cam, prm, seq, lights, n_lights, w, h, mode, args.cpu_seconds)
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
