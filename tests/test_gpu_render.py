"""Render on the GPU against the oracle.

 - LYAP_PREC_FP64 renders reproduce the oracle's FP64 orbit at every
   sample, so march decisions, sample counts, hit positions and depths agree
   exactly for both refinement methods (the reference walk with jitter and
   the near/far step, and the configured bisection);
 - the FP32 forms on the plain-render workload (512x512, AABAB, 100 + 300
   iterations, 500 uniform steps, 8 bisections, h = 1e-3) are held to the
   reference's own FP32 build on the same rays: hit/miss agreement and depth
   within one march step of the FP64 oracle at least the oracle's
   FP32-as-written form's, less 0.002;
 - the analytic one-letter scene puts the surface at x = 2 + e^-0.5 with
   normal +x;
 - the device's shading of its own hit records matches the oracle's shade()
   to one RGBA step.
"""
import math

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


def gpu_render(ctx, cam, prm, seq, lights, n_lights, w, h, mode):
    bufs = RenderBuffers(w, h)
    ld = device_lights(lights, n_lights)
    bufs.render(ctx, cam, prm, seq, ld, n_lights, mode)
    ctx.sync()
    out = bufs.fetch()
    out["total_evals"] = ctx.last_evals()
    return out


def assert_exact(got, ref):
    hit_g = ~np.isnan(got["depth"])
    hit_r = ~np.isnan(ref["depth"])
    assert np.array_equal(hit_g, hit_r)
    assert np.array_equal(got["evals"], ref["evals"])
    assert np.array_equal(got["depth"][hit_g], ref["depth"][hit_r])
    assert np.array_equal(got["points"][..., :3][hit_g], ref["points"][..., :3][hit_r])
    assert np.allclose(got["points"][..., 6:8][hit_g], ref["points"][..., 6:8][hit_r], atol=1e-5, equal_nan=True)
    assert np.allclose(got["lambda"][hit_g], ref["lambda"][hit_r], atol=1e-6)
    assert np.allclose(got["points"][..., 3:6][hit_g], ref["points"][..., 3:6][hit_r], atol=1e-3, equal_nan=True)
    assert np.max(np.abs(got["rgba"].astype(int) - ref["rgba"].astype(int))) <= 2
    assert got["total_evals"] == int(got["evals"].sum())


def test_fp64_walk_matches_oracle(ctx):
    """Reference method and defaults (jitter 0.5, near step, refine 32, normal
    offset step*gradient) on the scripted camera, fewer iterations."""
    prm, cam, lights, n_lights = lyap.params_init()
    prm.settle = 18
    prm.accum = 120
    prm.depth = 256.0
    w = h = 40
    lyap.cam_recalculate(cam, w, h, 1)
    lyap.lights_recalculate(lights, n_lights)
    seq = lyap.default_sequence()
    got = gpu_render(ctx, cam, prm, seq, lights, n_lights, w, h, lyap.render_mode(precision=lyap.PREC_FP64))
    ref = O.render(cam, prm, seq, lights, n_lights, w, h, mode=lyap.render_mode(), lam=O.F64)
    assert np.mean(~np.isnan(ref["depth"])) > 0.2
    assert_exact(got, ref)


def test_fp64_bisect_matches_oracle(ctx):
    cam, prm, seq, lights, n_lights = W.render_config(64, 64)
    mode = W.c2_mode(precision=lyap.PREC_FP64)
    got = gpu_render(ctx, cam, prm, seq, lights, n_lights, 64, 64, mode)
    ref = O.render(cam, prm, seq, lights, n_lights, 64, 64, mode=mode, lam=O.F64)
    assert np.mean(~np.isnan(ref["depth"])) > 0.2
    assert_exact(got, ref)


def test_zero_step_matches_oracle(ctx):
    cam, prm, seq, lights, n_lights = W.render_config(32, 32)
    prm.depth = float("inf")
    for mode in [W.c2_mode(precision=lyap.PREC_FP64), lyap.render_mode(precision=lyap.PREC_FP64)]:
        got = gpu_render(ctx, cam, prm, seq, lights, n_lights, 32, 32, mode)
        ref = O.render(cam, prm, seq, lights, n_lights, 32, 32, mode=mode, lam=O.F64)
        assert_exact(got, ref)


def step_of(cam, prm, w, h, ys, xs):
    dt = np.empty(len(xs))
    for i, (x, y) in enumerate(zip(xs, ys)):
        _, V, t0, t1 = O.ray_setup(cam, prm, int(x), int(y))
        dt[i] = (t1 - t0) / prm.depth if prm.stepMethod == 1 else np.linalg.norm(V) / prm.depth
    return dt


@pytest.fixture(scope="module")
def c2_oracle():
    w = h = 512
    cam, prm, seq, lights, n_lights = W.render_config(w, h)
    return O.render(cam, prm, seq, lights, n_lights, w, h, mode=W.c2_mode(), lam=O.F64)


# Slack below the reference's own FP32 agreement that an FP32 form may sit
# at: a different FP32 realisation of the same chaotic transients.
FP32_SLACK = 0.002


def depth_agreement(cam, prm, got, ref):
    """(hit/miss agreement, fraction of rays both hit with depth within one
    march step) of `got` against `ref`."""
    hit_g, hit_r = ~np.isnan(got["depth"]), ~np.isnan(ref["depth"])
    agree = float(np.mean(hit_g == hit_r))
    both = hit_g & hit_r
    yy, xx = np.nonzero(both)
    dt = step_of(cam, prm, 0, 0, yy, xx)
    within = float(np.mean(np.abs(got["depth"][both] - ref["depth"][both]) <= dt * (1 + 1e-5)))
    return agree, within


@pytest.fixture(scope="module")
def c2_reference_fp32(c2_oracle):
    """The reference's FP32 build (the oracle's FP32-as-written form, whose
    map and derivative promote to double through kernel.cu:135,143-144's
    literals) against its FP64 build on the same frame: on this frame
    hit/miss agreement 0.9986 and depth within one step 0.8721."""
    w = h = 512
    cam, prm, seq, lights, n_lights = W.render_config(w, h)
    f32 = O.render(cam, prm, seq, lights, n_lights, w, h, mode=W.c2_mode(), lam=O.F32W)
    return depth_agreement(cam, prm, f32, c2_oracle)


@pytest.mark.parametrize("prec,margin", [(lyap.PREC_FP64, 0.0), (lyap.PREC_PRODUCT, 0.01),
                                         (lyap.PREC_PRODUCT, 0.0), (lyap.PREC_PLAIN, 0.0)])
def test_c2_depth_within_one_step(ctx, prec, margin, c2_oracle, c2_reference_fp32):
    """The 512x512 plain render against the FP64 oracle, every ray. FP64:
    every ray agrees (hit, depth, sample count). FP32 forms, including the
    benchmarked one (product, no FP64 margin): a finite-window lambda near a
    chaotic transient is a different realisation in FP32 than in FP64, and a
    ray samples ~26 points, so ~13% of rays stop a step or more apart; the
    floor is the reference's own FP32 build on the same rays (the oracle's
    FP32-as-written form against FP64) less FP32_SLACK."""
    w = h = 512
    cam, prm, seq, lights, n_lights = W.render_config(w, h)
    got = gpu_render(ctx, cam, prm, seq, lights, n_lights, w, h, W.c2_mode(precision=prec, fp64_margin=margin))
    ref = c2_oracle
    agree, within = depth_agreement(cam, prm, got, ref)
    ref_agree, ref_within = c2_reference_fp32
    hit_r = ~np.isnan(ref["depth"])
    print(f"c2 prec {prec} margin {margin}: hit fraction {hit_r.mean():.3f} agree {agree:.4f} "
          f"within-one-step {within:.4f} (reference FP32: {ref_agree:.4f} / {ref_within:.4f})")
    assert hit_r.mean() > 0.1
    if prec == lyap.PREC_FP64:
        assert agree == 1.0 and within == 1.0
        assert np.array_equal(got["evals"], ref["evals"])
    else:
        assert agree >= ref_agree - FP32_SLACK
        assert within >= ref_within - FP32_SLACK
    # misses are transparent black with a NaN record
    miss = np.isnan(got["depth"])
    assert np.all(got["rgba"][miss] == 0)
    assert np.all(np.isnan(got["points"][miss][:, 6]))
    assert got["total_evals"] == int(got["evals"].sum())


def test_shading_matches_oracle(ctx):
    w = h = 96
    cam, prm, seq, lights, n_lights = W.render_config(w, h, seed=2)
    got = gpu_render(ctx, cam, prm, seq, lights, n_lights, w, h, W.c2_mode())
    _, rgba = O.shade(got["points"].reshape(-1, 9), cam, lights, n_lights)
    diff = np.abs(rgba.astype(int) - got["rgba"].reshape(-1, 4).astype(int))
    assert diff.max() <= 1


@pytest.mark.parametrize("mode", [W.c2_mode(), lyap.render_mode(precision=lyap.PREC_PRODUCT)])
def test_analytic_plane(ctx, mode):
    """Sequence 'A' in the box [2,3]^3: lambda = ln|2 - x| on x in (2,3), so
    with opaque -0.5 the first opaque point seen from +x is x = 2 + e^-0.5."""
    w = h = 32
    prm, _, lights, n_lights = lyap.params_init()
    prm.settle, prm.accum = 200, 200
    prm.lMin, prm.lMax = 2.0, 3.0
    prm.opaqueThreshold = -0.5
    prm.chaosThreshold = 0.5
    prm.nearThreshold = -1.0
    prm.stepMethod = 1
    prm.depth = 400.0
    prm.jitter = 0.0
    prm.nearMultiplier = 1.0
    cam = lyap.cam_look_at((4.0, 2.5, 2.5), (2.0, 2.5, 2.5), 0.2, w, h)
    lyap.lights_recalculate(lights, n_lights)
    got = gpu_render(ctx, cam, prm, "A", lights, n_lights, w, h, mode)
    x_plane = 2.0 + math.exp(-0.5)
    P = got["points"][..., :3].reshape(-1, 3)
    N = got["points"][..., 3:6].reshape(-1, 3)
    assert not np.any(np.isnan(got["depth"]))
    ys, xs = np.divmod(np.arange(w * h), w)
    dt = step_of(cam, prm, w, h, ys, xs)
    V = np.array([O.ray_setup(cam, prm, int(x), int(y))[1] for x, y in zip(xs, ys)])
    assert np.all(np.abs(P[:, 0] - x_plane) <= dt * np.abs(V[:, 0]) * 1.001 + 1e-6)
    assert np.all(N[:, 0] > 0.99)
