"""Float medians over windows holding NaN or -0, bit for bit (the sign of
zero included) against the oracle's insertion sort: median_box_3d,
median_box_2d, median_slice_by_slice_box_3d, median_sphere_3d and the
batch engine (the allocation-free entry point: test_zz_no_alloc.py).  The volumes mix bricks with and
without NaN / -0, so both the recomputed bricks and the untouched ones are
checked."""
import numpy as np
import pytest

import clk_oracle as O
from helpers import empty_device, from_device, nanzero_volume, same_float_bits, to_device

pytestmark = pytest.mark.gpu

KINDS = ["ties", "sparse", "slab"]


def assert_bits(got, want, ctx):
    if not same_float_bits(got, want):
        gn, wn = np.isnan(got), np.isnan(want)
        bad = np.argwhere((gn != wn) | (~gn & ~wn & (got.view(np.uint32) != want.view(np.uint32))))
        i = tuple(bad[0])
        raise AssertionError("%s: %d mismatches, first at (z,y,x)=%s got %r (%08x) want %r (%08x)"
                             % (ctx, len(bad), i, got[i], got.view(np.uint32)[i], want[i], want.view(np.uint32)[i]))


@pytest.mark.parametrize("kind", KINDS)
def test_median_box_3d_nan_negzero(gpu, kind):
    import clkmi
    a = nanzero_volume((13, 19, 37), kind, seed=5)
    ts, si = to_device(a)
    for Ns in ((3, 3, 3), (5, 5, 5), (7, 3, 1), (1, 1, 1), (3, 1, 5), (2, 4, 3), (1, 1, 9)):
        print("median_box_3d", kind, Ns, flush=True)
        td, di = empty_device(a.shape, np.float32)
        clkmi.median_box_3d(di, si, *Ns)
        assert_bits(from_device(td, a.shape, np.float32), O.median_box_3d(a, np.float32, *Ns),
                    "median_box_3d %s %s" % (kind, Ns))


def test_median_all_negzero_is_negzero(gpu):
    import clkmi
    a = np.full((9, 10, 11), -0.0, dtype=np.float32)
    ts, si = to_device(a)
    for Ns in ((3, 3, 3), (5, 5, 5), (3, 3, 1)):
        td, di = empty_device(a.shape, np.float32)
        clkmi.median_box_3d(di, si, *Ns)
        got = from_device(td, a.shape, np.float32)
        assert np.all(got == 0) and np.all(np.signbit(got)), Ns
        td, di = empty_device(a.shape, np.float32)
        clkmi.median_sphere_3d(di, si, *Ns)
        got = from_device(td, a.shape, np.float32)
        assert np.all(got == 0) and np.all(np.signbit(got)), Ns


@pytest.mark.parametrize("kind", KINDS)
def test_median_2d_and_slices_nan_negzero(gpu, kind):
    import clkmi
    a = nanzero_volume((5, 21, 45), kind, seed=6)
    ts, si = to_device(a)
    t2, i2 = to_device(a[2:3])
    for Nx, Ny in ((3, 3), (5, 5), (1, 7), (3, 1), (4, 2)):
        print("median 2d/slice", kind, Nx, Ny, flush=True)
        td, di = empty_device(a.shape, np.float32)
        clkmi.median_slice_by_slice_box_3d(di, si, Nx, Ny)
        assert_bits(from_device(td, a.shape, np.float32), O.median_slice_by_slice_box_3d(a, np.float32, Nx, Ny),
                    "slice median %s %dx%d" % (kind, Nx, Ny))
        td, di = empty_device((1,) + a.shape[1:], np.float32)
        clkmi.median_box_2d(di, i2, Nx, Ny)
        assert_bits(from_device(td, (1,) + a.shape[1:], np.float32)[0], O.median_box_2d(a[2], np.float32, Nx, Ny),
                    "median_box_2d %s %dx%d" % (kind, Nx, Ny))


@pytest.mark.parametrize("kind", KINDS)
def test_median_sphere_3d_nan_negzero(gpu, kind):
    import clkmi
    a = nanzero_volume((17, 21, 40), kind, seed=7)
    ts, si = to_device(a)
    for Ns in ((3, 3, 3), (5, 5, 5), (9, 7, 3), (3, 5, 7), (1, 3, 1), (7, 1, 5)):
        print("median_sphere_3d", kind, Ns, flush=True)
        td, di = empty_device(a.shape, np.float32)
        clkmi.median_sphere_3d(di, si, *Ns)
        assert_bits(from_device(td, a.shape, np.float32), O.sphere_3d("median_sphere_3d", a, np.float32, *Ns),
                    "median_sphere_3d %s %s" % (kind, Ns))


def test_batch_float_medians_nan_negzero(gpu):
    import clkmi
    jobs, keep, outs = [], [], []
    for i, (kind, shape, radius) in enumerate([("ties", (9, 14, 23), (1, 1, 1)), ("sparse", (17, 12, 40), (2, 2, 2)),
                                               ("slab", (12, 9, 30), (1, 0, 1)), ("ties", (6, 30, 8), (0, 1, 2)),
                                               ("sparse", (20, 20, 20), (1, 1, 1))]):
        a = nanzero_volume(shape, kind, seed=20 + i)
        ts, si = to_device(a)
        td, di = empty_device(shape, np.float32)
        keep += [ts, td]
        outs.append((td, a, radius))
        jobs.append(clkmi.FilterJob("median_box_3d", di, si, (0.0, 0.0, 0.0), radius))
    print("batch medians", len(jobs), flush=True)
    clkmi.batch_filter_3d(jobs)
    for td, a, radius in outs:
        want = O.median_box_3d(a, np.float32, *(2 * r + 1 for r in radius))
        assert_bits(from_device(td, a.shape, np.float32), want, "batch median %s %s" % (a.shape, radius))
