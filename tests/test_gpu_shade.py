"""Re-shading a rendered frame for new lights (lyap_shade_points).

Lights enter only the shading of a finished ray, so re-shading a frame's
LyapPoint buffer under new lights must give exactly the RGBA of rendering
the frame again with them, and rendering again must leave the points as they
were.
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


def edited_lights(lights, n):
    """The viewer's light edits: colour, position and range of the first
    light, and a second, coloured light."""
    size = C.sizeof(lyap.LyapLight)
    new = (lyap.LyapLight * lyap.MAX_LIGHTS)()
    C.memmove(new, lights, size * lyap.MAX_LIGHTS)
    C.memmove(C.byref(new, size), C.byref(new, 0), size)     # light 1 starts as light 0
    L = new[0]                                                # views into the array
    L.diffuseColor.x, L.diffuseColor.y, L.diffuseColor.z = 0.2, 0.9, 0.4
    L.specularColor = L.diffuseColor
    L.C.x += 0.7
    L.C.y -= 0.4
    L.lightRange *= 1.5
    L1 = new[1]
    L1.diffuseColor.x, L1.diffuseColor.y = 0.8, 0.1
    L1.C.x, L1.C.y, L1.C.z = 1.0, 5.0, 1.0
    m = max(n, 2)
    lyap.lights_recalculate(new, m)
    return new, m


@pytest.mark.parametrize("precision", [lyap.PREC_PRODUCT, lyap.PREC_FP64])
def test_reshade_equals_render_with_new_lights(ctx, precision):
    torch = torch_cuda()
    w, h = 96, 80
    cam, prm, seq, lights, n = W.render_config(w, h)
    mode = W.c2_mode(precision=precision, fp64_margin=0.0)
    first = RenderBuffers(w, h)
    first.render(ctx, cam, prm, seq, device_lights(lights, n), n, mode)
    ctx.sync()
    before = first.fetch()
    hits = ~np.isnan(before["depth"])
    assert 0.2 < hits.mean() < 1.0                      # hits and misses both present

    new, m = edited_lights(lights, n)
    ld2 = device_lights(new, m)
    again = RenderBuffers(w, h)
    again.render(ctx, cam, prm, seq, ld2, m, mode)
    ctx.sync()
    ref = again.fetch()
    assert np.array_equal(ref["points"], before["points"], equal_nan=True)
    assert not np.array_equal(ref["rgba"], before["rgba"])   # the edit shows

    out = empty(4 * w * h)
    torch.cuda.synchronize()
    ctx.shade_points(out.data_ptr(), first.points.data_ptr(), cam, ld2.data_ptr(), m, w, h)
    ctx.sync()
    got = to_numpy(out, np.uint8, (h, w, 4))
    assert np.array_equal(got, ref["rgba"])
    assert np.all(got[~hits] == 0)


def test_shade_points_errors(ctx):
    cam, prm, seq, lights, n = W.render_config(8, 8)
    buf = empty(4 * 64)
    pts = empty(36 * 64)
    ld = device_lights(lights, n)
    with pytest.raises(lyap.LyapError) as e:
        ctx.shade_points(0, pts.data_ptr(), cam, ld.data_ptr(), n, 8, 8)
    assert e.value.status == lyap.ERR_INVALID_ARGUMENT
    with pytest.raises(lyap.LyapError) as e:
        ctx.shade_points(buf.data_ptr(), pts.data_ptr(), cam, ld.data_ptr(), lyap.MAX_LIGHTS + 1, 8, 8)
    assert e.value.status == lyap.ERR_INVALID_ARGUMENT
    with pytest.raises(lyap.LyapError) as e:
        ctx.shade_points(buf.data_ptr(), pts.data_ptr(), cam, 0, 1, 8, 8)
    assert e.value.status == lyap.ERR_INVALID_ARGUMENT
    ctx.shade_points(0, 0, cam, 0, 0, 0, 8)             # empty image: nothing to do
