"""culamp_lse_rows_noalloc (the row LSE that allocates nothing) against
culamp_lse_rows through the C ABI: the same kernels and launch shape, so the
same bits, on two calls in a row with the second on a smaller input (the
shared scratch array then holds the first call's partials), in both
precisions and in the split layout whose partials live in the scratch array;
and a launch too large for the array is an error, not a fault."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(rng, n, m):
    x = rng.uniform(size=(n, 3))
    y = rng.uniform(size=(m, 3))
    th = rng.uniform(-0.5, 0.5, m)
    return x, y, th


def _run(ctx, fn, x, y, th, reg, prec):
    import culamp_amd
    from culamp_amd._lib import check
    n, m = len(x), len(y)
    dx = culamp_amd.DeviceArray.from_host(ctx, x)
    dy = culamp_amd.DeviceArray.from_host(ctx, y)
    dth = culamp_amd.DeviceArray.from_host(ctx, th)
    out = culamp_amd.DeviceArray(ctx, n)
    check(getattr(ctx.lib, fn)(ctx.handle, dx.ptr, n, dy.ptr, m, dth.ptr, C.c_double(reg),
                               C.c_double(1.0), C.c_double(2.0), C.c_double(2.0), prec, out.ptr))
    check(ctx.lib.culamp_ctx_synchronize(ctx.handle))
    return out.get()


@pytest.mark.parametrize("layout", [0, 1])
def test_noalloc_matches_lse_rows(gpu_ctx, layout):
    from culamp_amd import FP64, MIXED
    from culamp_amd._lib import check
    rng = np.random.default_rng(71)
    check(gpu_ctx.lib.culamp_ctx_set_layout(gpu_ctx.handle, layout))
    try:
        for prec in (FP64, MIXED):
            for (n, m) in ((3000, 5000), (700, 1300)):  # second call smaller
                x, y, th = _case(rng, n, m)
                say("noalloc layout", layout, "prec", prec, "n", n, "m", m)
                a = _run(gpu_ctx, "culamp_lse_rows_noalloc", x, y, th, 0.05, prec)
                b = _run(gpu_ctx, "culamp_lse_rows", x, y, th, 0.05, prec)
                assert np.all(np.isfinite(a))
                assert np.array_equal(a, b)
    finally:
        check(gpu_ctx.lib.culamp_ctx_set_layout(gpu_ctx.handle, 0))


def test_noalloc_rejects_launch_beyond_scratch(gpu_ctx):
    """A forced split count whose partials exceed the 2^21-double array is
    refused on the host before any launch."""
    import culamp_amd
    from culamp_amd import FP64
    from culamp_amd._lib import CulampError, check
    rng = np.random.default_rng(72)
    n, m = 40000, 8192
    x, y, th = _case(rng, n, m)
    check(gpu_ctx.lib.culamp_ctx_set_launch(gpu_ctx.handle, 1, 64))
    try:
        with pytest.raises(CulampError, match="scratch"):
            _run(gpu_ctx, "culamp_lse_rows_noalloc", x, y, th, 0.05, FP64)
    finally:
        check(gpu_ctx.lib.culamp_ctx_set_launch(gpu_ctx.handle, 0, 0))
    # the automatic launch of the same size fits and matches
    a = _run(gpu_ctx, "culamp_lse_rows_noalloc", x, y, th, 0.05, FP64)
    b = _run(gpu_ctx, "culamp_lse_rows", x, y, th, 0.05, FP64)
    assert np.array_equal(a, b)
