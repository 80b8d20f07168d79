"""The allocation-free volume entry against lyap_calc_volume, on the GPU.

lyap_calc_volume_noalloc runs the same kernels as lyap_calc_volume with its
FP64 re-evaluation list and counter in a fixed per-device array, in z slabs
the list holds entirely, so the two give the same bytes. Two calls in a row
(the second smaller) check that nothing left in the shared array from the
first leaks into the second; the context's own device memory does not grow.
"""
import numpy as np
import pytest

import lyapunov3d_amd as lyap
from tests import workloads as W
from tests.gpu_util import empty, to_numpy, torch_cuda

pytestmark = pytest.mark.gpu


def run(ctx, fn, prm, seq, lo, hi, n, mode):
    nx, ny, nz = n
    out = empty(4 * nx * ny * nz)
    out.fill_(0x7F)
    fn(out.data_ptr(), prm, seq, lo, hi, n, mode)
    ctx.sync()
    return to_numpy(out, np.float32)[:nx * ny * nz].copy()


# (params, sequence, box, size, mode): the first case needs three slabs of
# the scratch list (1024 x 1024 voxels per slice, 4 slices per slab), the
# second is smaller and uses one
CASES = [
    (dict(settle=200, accum=600), "ABCABCCA", (2.5, 2.5, 2.5), (4.0, 4.0, 4.0), (1024, 1024, 10),
     lyap.volume_mode(**W.C3_STRICT)),
    (dict(settle=100, accum=400), "ABBAC", (2.0, 2.0, 2.0), (4.0, 4.0, 4.0), (64, 64, 64),
     lyap.volume_mode(**W.C3_MODE)),
]


@pytest.mark.parametrize("extra", [None, lyap.volume_mode(lyap.PREC_PLAIN), lyap.volume_mode(lyap.PREC_FP64)])
def test_noalloc_equals_calc_volume_twice(extra):
    torch_cuda()
    ctx = lyap.Context(0)
    try:
        for prm_kw, seq, lo, hi, n, mode in CASES:
            if extra is not None:
                mode = extra
            prm = lyap.make_params(**prm_kw)
            used = ctx.bytes_in_use()
            a = run(ctx, ctx.calc_volume_noalloc, prm, seq, lo, hi, n, mode)
            assert ctx.bytes_in_use() == used
            b = run(ctx, ctx.calc_volume, prm, seq, lo, hi, n, mode)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
            assert np.isfinite(a).mean() > 0.5
            if mode.precision == lyap.PREC_PRODUCT and mode.fp64_margin > 0:
                assert ctx.last_reevaluated() > 0      # the rule selected voxels
    finally:
        ctx.close()


def test_noalloc_rejects_a_slice_beyond_the_scratch():
    torch_cuda()
    ctx = lyap.Context(0)
    try:
        out = empty(64)
        with pytest.raises(lyap.LyapError) as e:
            ctx.calc_volume_noalloc(out.data_ptr(), lyap.make_params(), "AB", (2, 2, 2), (4, 4, 4),
                                    (4096, 1025, 1), lyap.volume_mode(**W.C3_MODE))
        assert e.value.status == lyap.ERR_UNSUPPORTED
    finally:
        ctx.close()
