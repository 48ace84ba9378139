"""clk_filter_3d_noalloc: the allocation-free 3D filter gives the bits of
the allocating entry points (clk_gaussian_blur_3d, clk_*_box_3d,
clk_median_box_3d) through the C-ABI, on two calls in a row on one stream
with the second volume smaller than the first (the fixed device array keeps
the first call's contents), and rejects a job whose temporaries exceed the
array."""
import ctypes

import numpy as np
import pytest

from helpers import empty_device, from_device, nanzero_volume, random_volume, same_float_bits, to_device

pytestmark = pytest.mark.gpu

KINDS = {"gaussian_blur_3d": 0, "mean_box_3d": 1, "minimum_box_3d": 2, "maximum_box_3d": 3, "median_box_3d": 4}


def c_image(img):
    from clkmi import _lib
    return _lib.ClkImage(img.ptr, int(img.dtype), 0, int(img.width), int(img.height), int(img.depth))


def c_job(kind, dst, src, sigma=(0.0, 0.0, 0.0), radius=(0, 0, 0)):
    from clkmi import _lib
    j = _lib.ClkFilterJob()
    j.kind = KINDS[kind]
    j.reserved = 0
    j.dst = c_image(dst)
    j.src = c_image(src)
    for k in range(3):
        j.sigma[k] = float(sigma[k])
        j.radius[k] = int(radius[k])
    return j


def allocating(L, stream, kind, dst, src, sigma, radius):
    d, s = c_image(dst), c_image(src)
    if kind == "gaussian_blur_3d":
        return L.clk_gaussian_blur_3d(stream, ctypes.byref(d), ctypes.byref(s), *[float(v) for v in sigma])
    if kind == "median_box_3d":
        return L.clk_median_box_3d(stream, ctypes.byref(d), ctypes.byref(s), *[2 * int(r) + 1 for r in radius])
    return getattr(L, "clk_" + kind)(stream, ctypes.byref(d), ctypes.byref(s), *[int(r) for r in radius])


CASES = [
    ("gaussian_blur_3d", np.uint16, np.float32, (2.0, 1.5, 1.0), (0, 0, 0)),
    ("gaussian_blur_3d", np.float32, np.uint16, (0.0, 3.0, 0.7), (0, 0, 0)),
    ("mean_box_3d", np.uint8, np.uint8, (0, 0, 0), (2, 1, 3)),
    ("minimum_box_3d", np.int16, np.int16, (0, 0, 0), (3, 3, 3)),
    ("maximum_box_3d", np.float32, np.float32, (0, 0, 0), (1, 0, 2)),
    ("median_box_3d", np.uint8, np.uint8, (0, 0, 0), (1, 1, 1)),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s-%s" % (c[0], np.dtype(c[1]).name, np.dtype(c[2]).name))
def test_noalloc_matches_allocating_entry_point_twice(gpu, case):
    import torch
    from clkmi import _lib
    L = _lib.lib()
    kind, st, dt, sigma, radius = case
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    # the second call is smaller than the first: stale contents of the
    # fixed array must not leak into its result
    for i, shape in enumerate([(70, 90, 130), (33, 41, 57)]):
        src = random_volume(shape, st, seed=100 + i)
        ts, si = to_device(src)
        ta, da = empty_device(shape, dt)
        tb, db = empty_device(shape, dt)
        assert allocating(L, stream, kind, da, si, sigma, radius) == 0, L.clk_last_error_string()
        job = c_job(kind, db, si, sigma, radius)
        rc = L.clk_filter_3d_noalloc(stream, ctypes.byref(job))
        assert rc == 0, L.clk_last_error_string()
        want = from_device(ta, shape, dt)
        got = from_device(tb, shape, dt)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), "%s call %d %s" % (kind, i, shape)


def test_noalloc_rejects_temporaries_beyond_the_array(gpu):
    import torch
    from clkmi import _lib
    L = _lib.lib()
    shape = (257, 256, 256)  # two float32 temporaries: just over 128 MiB
    ts, si = empty_device(shape, np.uint16)
    td, di = empty_device(shape, np.float32)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    job = c_job("gaussian_blur_3d", di, si, sigma=(1.0, 1.0, 1.0))
    assert L.clk_filter_3d_noalloc(stream, ctypes.byref(job)) == -30
    assert b"fixed array" in L.clk_last_error_string()
    # two axes need one temporary: the same volume fits
    job = c_job("gaussian_blur_3d", di, si, sigma=(1.0, 0.0, 1.0))
    assert L.clk_filter_3d_noalloc(stream, ctypes.byref(job)) == 0, L.clk_last_error_string()
    torch.cuda.synchronize()


def test_noalloc_float_median_nan_negzero(gpu):
    """Float -> float medians with NaN and -0: the brick flags of the fix-up
    live in the fixed array too; two calls, the second smaller, bit for bit
    (sign of zero included) against the allocating entry point."""
    import torch
    from clkmi import _lib
    L = _lib.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    radius = (1, 2, 1)
    for i, shape in enumerate([(40, 50, 60), (9, 11, 13)]):
        a = nanzero_volume(shape, "ties", seed=40 + i)
        ts, si = to_device(a)
        ta, da = empty_device(shape, np.float32)
        tb, db = empty_device(shape, np.float32)
        assert allocating(L, stream, "median_box_3d", da, si, (0.0, 0.0, 0.0), radius) == 0, L.clk_last_error_string()
        job = c_job("median_box_3d", db, si, (0.0, 0.0, 0.0), radius)
        assert L.clk_filter_3d_noalloc(stream, ctypes.byref(job)) == 0, L.clk_last_error_string()
        want = from_device(ta, shape, np.float32)
        got = from_device(tb, shape, np.float32)
        assert same_float_bits(got, want), "call %d %s" % (i, shape)
        assert np.any(np.signbit(got) & (got == 0)), "no -0 medians in the sample"
