"""GPU: the allocation-free WRP entry point (oph_wrp_noalloc), called through
the C ABI, against oph_wrp on the same device arrays: bit-identical planes on
two calls in a row (the second with fewer points on a smaller plane), with
accumulate and asynchronous flags, and clear errors beyond its limits."""
import ctypes as C

import numpy as np
import pytest

from openhologram_amd import scenes
from openhologram_amd.lib import ACCUMULATE, ASYNC, DEVICE_PTR, OphError, Points, ptr

pytestmark = pytest.mark.gpu


def _device_points(ctx, x, y, z, a):
    d = [ctx.to_device(np.ascontiguousarray(v, np.float64)) for v in (x, y, z, a)]
    return Points(ptr(d[0]), ptr(d[1]), ptr(d[2]), ptr(d[3]), None, len(x)), d


def _call(ctx, fn, n, pts, out, flags, zw, mode=0):
    return fn(ctx.h, n, n, 8e-6, 532e-9, zw, mode, C.byref(pts), ptr(out), flags)


def test_noalloc_matches_oph_wrp(ctx):
    lib = ctx.lib
    for n_pts, side, seed, mode in ((120_000, 2048, 2, 0), (5_000, 512, 3, 1)):
        print(f"case: wrp_noalloc {n_pts} points on {side}^2 mode {mode}", flush=True)
        x, y, z, a, zw = scenes.wrp_points(n_pts, seed=seed)
        if side < 2048:
            x, y = x * side / 2048, y * side / 2048
        pts, keep = _device_points(ctx, x, y, z, a)
        ref = ctx.empty((side, side), np.complex64)
        out = ctx.empty((side, side), np.complex64)
        ctx.check(_call(ctx, lib.oph_wrp, side, pts, ref, DEVICE_PTR, zw, mode))
        ctx.check(_call(ctx, lib.oph_wrp_noalloc, side, pts, out, DEVICE_PTR, zw, mode))
        r, o = ref.numpy(), out.numpy()
        assert np.any(r != 0)
        assert np.array_equal(o, r)
        # accumulate and asynchronous forms
        ctx.check(_call(ctx, lib.oph_wrp_noalloc, side, pts, out, DEVICE_PTR | ACCUMULATE | ASYNC, zw, mode))
        ctx.sync()
        ctx.check(_call(ctx, lib.oph_wrp, side, pts, ref, DEVICE_PTR | ACCUMULATE, zw, mode))
        assert np.array_equal(out.numpy(), ref.numpy())
        del keep


def test_noalloc_limits(ctx):
    lib = ctx.lib
    x, y, z, a, zw = scenes.wrp_points(10, seed=4)
    pts, keep = _device_points(ctx, x, y, z, a)
    out = ctx.empty((2112, 64), np.complex64)
    # plane of more than 1024 tiles, host pointers, too many points: errors, no launch
    assert _call(ctx, lib.oph_wrp_noalloc, 2112, pts, out, DEVICE_PTR, zw) == -1
    with pytest.raises(OphError):
        ctx.check(_call(ctx, lib.oph_wrp_noalloc, 64, pts, out, DEVICE_PTR, zw, mode=4))
    host = np.zeros((64, 64), np.complex64)
    hp = Points(*[v.ctypes.data_as(C.c_void_p).value for v in (x, y, z, a)], None, len(x))
    assert lib.oph_wrp_noalloc(ctx.h, 64, 64, 8e-6, 532e-9, zw, 0, C.byref(hp), ptr(host), 0) == -1
    big = Points(pts.x, pts.y, pts.z, pts.amp, None, 131_073)
    assert _call(ctx, lib.oph_wrp_noalloc, 64, big, out, DEVICE_PTR, zw) == -1
    del keep
