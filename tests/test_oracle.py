"""Pins the CPU restatement (oracle/clk_oracle.c) against facts that do not
come from it: scipy's rank filters and correlations, exact integer
arithmetic, known answers worked out by hand, and IEEE division.  A dropped
tap, a shifted window, a wrong border, a wrong conversion or a reordered sum
fails one of these."""
import math

import numpy as np
import pytest
from scipy import ndimage

import clk_oracle as O
import zlib

from helpers import DTYPES, nanzero_volume, random_volume, same_bits_or_value, same_float_bits

SHAPES = [(11, 13, 17), (5, 1, 9), (1, 1, 23)]


# ---- CONVERT_dst_PIXEL_TYPE -------------------------------------------------
# Known answers of OpenCL 1.2 convert_<T>_sat with round-toward-zero.
CONVERT_CASES = [
    # value, char, uchar, short, ushort
    (-1e9, -128, 0, -32768, 0),
    (-129.0, -128, 0, -129, 0),
    (-128.5, -128, 0, -128, 0),
    (-0.9, 0, 0, 0, 0),
    (-0.0, 0, 0, 0, 0),
    (0.5, 0, 0, 0, 0),
    (1.99, 1, 1, 1, 1),
    (127.7, 127, 127, 127, 127),
    (254.9, 127, 254, 254, 254),
    (255.5, 127, 255, 255, 255),
    (256.0, 127, 255, 256, 256),
    (32767.9, 127, 255, 32767, 32767),
    (65535.9, 127, 255, 32767, 65535),
    (1e9, 127, 255, 32767, 65535),
    (float("inf"), 127, 255, 32767, 65535),
    (float("-inf"), -128, 0, -32768, 0),
    (float("nan"), 0, 0, 0, 0),
]


@pytest.mark.parametrize("case", CONVERT_CASES)
def test_convert_known_answers(case):
    v = np.float32(case[0])
    for dt, want in zip([np.int8, np.uint8, np.int16, np.uint16], case[1:]):
        got = O.convert(np.array([v]), dt)[0]
        assert int(got) == want, (case[0], dt, got, want)
    f = O.convert(np.array([v], dtype=np.float32), np.float32)[0]
    assert (np.isnan(f) and np.isnan(v)) or f == v


def test_expf_matches_libm_within_half_ulp():
    xs = np.float32(-np.arange(0, 2000) / 37.0)
    for x in xs:
        got = O.expf(x)
        want = math.exp(float(x))
        assert abs(float(got) - want) <= 0.5 * float(np.spacing(np.float32(want))) + 1e-45


# ---- minimum / maximum: scipy rank filters, exactly ---------------------------

@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("dim", [0, 1, 2])
@pytest.mark.parametrize("N", [1, 3, 4, 7, 15])
def test_minmax_pass_equals_scipy(dt, dim, N):
    src = random_volume((11, 13, 17), dt, seed=N * 10 + dim)
    c = (N - 1) // 2
    axis = 2 - dim
    got_min = O.minimum_separable_3d(src, dt, dim, N)
    got_max = O.maximum_separable_3d(src, dt, dim, N)
    want_min = ndimage.minimum_filter1d(src, 2 * c + 1, axis=axis, mode="nearest")
    want_max = ndimage.maximum_filter1d(src, 2 * c + 1, axis=axis, mode="nearest")
    assert np.array_equal(got_min, want_min)
    assert np.array_equal(got_max, want_max)


@pytest.mark.parametrize("r", [(1, 1, 1), (3, 3, 3), (2, 0, 5)])
def test_minmax_chain_equals_3d_rank_filter(r):
    src = random_volume((12, 15, 18), np.uint16, seed=7)
    size = (2 * r[2] + 1, 2 * r[1] + 1, 2 * r[0] + 1)
    got = O.box_3d(O.OP_MAX, src, np.uint16, *r)
    assert np.array_equal(got, ndimage.maximum_filter(src, size=size, mode="nearest"))
    got = O.box_3d(O.OP_MIN, src, np.uint16, *r)
    assert np.array_equal(got, ndimage.minimum_filter(src, size=size, mode="nearest"))


def test_minmax_saturating_conversion():
    src = random_volume((6, 7, 9), np.int16, seed=3)
    got = O.maximum_separable_3d(src, np.uint8, 1, 5)
    want = np.clip(ndimage.maximum_filter1d(src, 5, axis=1, mode="nearest"), 0, 255).astype(np.uint8)
    assert np.array_equal(got, want)


# ---- mean: exact integer arithmetic --------------------------------------------

def _window_sum_int(src, dim, N):
    c = (N - 1) // 2
    axis = 2 - dim
    x = src.astype(np.int64)
    pad = [(0, 0)] * 3
    pad[axis] = (c, c)
    xp = np.pad(x, pad, mode="edge")
    acc = np.zeros_like(x)
    L = x.shape[axis]
    for v in range(2 * c + 1):
        sl = [slice(None)] * 3
        sl[axis] = slice(v, v + L)
        acc += xp[tuple(sl)]
    return acc, 2 * c + 1


@pytest.mark.parametrize("src_t", [np.int8, np.uint8, np.int16, np.uint16])
@pytest.mark.parametrize("dst_t", DTYPES)
@pytest.mark.parametrize("dim", [0, 1, 2])
def test_mean_integer_source_exact(src_t, dst_t, dim):
    N = 7
    src = random_volume((9, 10, 11), src_t, seed=dim + 100)
    got = O.mean_separable_3d(src, dst_t, dim, N)
    S, n = _window_sum_int(src, dim, N)
    if np.dtype(dst_t).kind == "f":
        want = (S.astype(np.float32) / np.float32(n)).astype(np.float32)  # IEEE divide of exact values
    else:
        q = np.trunc(S / n)  # |S| < 2^24: the float quotient truncates like integer division
        info = np.iinfo(dst_t)
        want = np.clip(q, info.min, info.max).astype(dst_t)
    assert np.array_equal(got, want)


def test_mean_float_sum_is_sequential():
    # Window (2^24, 1, 1): summed left to right each +1 is lost to rounding
    # (2^24 + 1 rounds to 2^24), so the mean is RN(2^24/3) = 5592405.5; any
    # other order keeps 2 and gives 5592406.
    src = np.array([[[16777216.0, 1.0, 1.0]]], dtype=np.float32)
    got = O.mean_separable_3d(src, np.float32, 0, 3)
    assert got[0, 0, 1] == np.float32(5592405.5)


def test_mean_float_source_close_to_fp64():
    src = random_volume((8, 9, 10), np.float32, seed=5)
    for dim in (0, 1, 2):
        got = O.mean_separable_3d(src, np.float32, dim, 9).astype(np.float64)
        want = ndimage.uniform_filter1d(src.astype(np.float64), 9, axis=2 - dim, mode="nearest")
        assert np.max(np.abs(got - want)) <= 1e-5 * np.max(np.abs(src))


# ---- Gaussian: FP64 reference, impulse response, hand-worked taps ----------------

@pytest.mark.parametrize("src_t", DTYPES)
@pytest.mark.parametrize("dim", [0, 1, 2])
@pytest.mark.parametrize("N,s", [(9, 1.0), (17, 2.0), (5, 0.7), (25, 3.0)])
def test_gaussian_float_dst_close_to_fp64(src_t, dim, N, s):
    src = random_volume((9, 14, 27), src_t, seed=N)
    got = O.gaussian_blur_separable_3d(src, np.float32, dim, N, s).astype(np.float64)
    want = O.gaussian_pass_fp64(src, dim, N, s)
    scale = max(1.0, float(np.max(np.abs(src.astype(np.float64)))))
    assert np.max(np.abs(got - want)) <= 1e-5 * scale


@pytest.mark.parametrize("dst_t", [np.uint8, np.uint16, np.int16])
def test_gaussian_integer_dst_matches_fp64_away_from_ties(dst_t):
    src = random_volume((10, 12, 14), np.uint8, seed=11)
    got = O.gaussian_blur_separable_3d(src, dst_t, 1, 9, 1.5).astype(np.int64)
    want = O.gaussian_pass_fp64(src, 1, 9, 1.5)
    info = np.iinfo(dst_t)
    wt = np.clip(np.trunc(want), info.min, info.max).astype(np.int64)
    frac = np.abs(want - np.round(want))
    clear = frac > 1e-3  # not within float rounding of an integer boundary
    assert clear.mean() > 0.9
    assert np.array_equal(got[clear], wt[clear])
    assert np.max(np.abs(got - wt)) <= 1


def test_gaussian_impulse_gives_normalised_weights():
    N, s = 13, 2.0
    src = np.zeros((1, 1, 41), dtype=np.float32)
    src[0, 0, 20] = 1.0
    got = O.gaussian_blur_separable_3d(src, np.float32, 0, N, s)[0, 0]
    v = np.arange(-6, 7)
    w = np.exp(-(v * v) / (2 * s * s))
    w /= w.sum()
    np.testing.assert_allclose(got[14:27], w[::-1], rtol=2e-7, atol=1e-9)
    assert np.all(got[:14] == 0) and np.all(got[27:] == 0)


def test_gaussian_three_taps_by_hand():
    # N=3, s=1: weights e^-0.5, 1, e^-0.5 (float32), hsum = (h + 1) + h.
    h = np.float32(math.exp(-0.5))
    hsum = np.float32(np.float32(h + np.float32(1.0)) + h)
    src = np.array([[[10.0, 20.0, 40.0]]], dtype=np.float32)
    got = O.gaussian_blur_separable_3d(src, np.float32, 0, 3, 1.0)[0, 0]
    # centre voxel: fma chain h*10 -> +1*20 -> +h*40, exact in float64 then rounded per step
    r = np.float32(np.float64(h) * 10.0)
    r = np.float32(np.float64(r) + 20.0)
    r = np.float32(np.float64(h) * 40.0 + np.float64(r))
    assert got[1] == np.float32(r / hsum)
    # clamped left border: taps read (10, 10, 20)
    r0 = np.float32(np.float64(h) * 10.0)
    r0 = np.float32(np.float64(r0) + 10.0)
    r0 = np.float32(np.float64(h) * 20.0 + np.float64(r0))
    assert got[0] == np.float32(r0 / hsum)


def test_gaussian_constant_volume():
    src = np.full((5, 6, 7), 123.0, dtype=np.float32)
    for dim in (0, 1, 2):
        got = O.gaussian_blur_separable_3d(src, np.float32, dim, 11, 2.0)
        np.testing.assert_allclose(got, 123.0, rtol=1e-6)


# ---- degenerate parameters ---------------------------------------------------------

@pytest.mark.parametrize("op", [O.OP_GAUSS, O.OP_MEAN, O.OP_MIN, O.OP_MAX])
def test_n1_is_converting_copy(op):
    src = random_volume((4, 5, 6), np.int16, seed=2)
    got = O.separable(op, src, np.uint8, 1, 1, 1.0)
    assert np.array_equal(got, np.clip(src, 0, 255).astype(np.uint8))


def test_out_of_range_dim_reads_centre_only():
    src = random_volume((4, 5, 6), np.uint16, seed=4)
    for op in (O.OP_MEAN, O.OP_MIN, O.OP_MAX):
        assert np.array_equal(O.separable(op, src, np.uint16, 7, 9, 0.0), src)


def test_negative_n_gives_converted_nan_or_centre():
    src = random_volume((3, 4, 5), np.uint8, seed=9)
    assert np.all(O.mean_separable_3d(src, np.uint8, 0, -3) == 0)
    assert np.all(np.isnan(O.gaussian_blur_separable_3d(src, np.float32, 0, -1, 1.0)))
    assert np.array_equal(O.minimum_separable_3d(src, np.uint8, 0, -5), src)


def test_sigma_zero_gives_nan():
    src = random_volume((3, 4, 5), np.uint8, seed=9)
    assert np.all(np.isnan(O.gaussian_blur_separable_3d(src, np.float32, 2, 5, 0.0)))


# ---- median: scipy's median filter, exactly ------------------------------------------

@pytest.mark.parametrize("dt", [np.uint8, np.int8, np.uint16, np.int16, np.float32])
@pytest.mark.parametrize("Ns", [(3, 3, 3), (5, 5, 5), (3, 1, 5), (1, 1, 1), (4, 3, 2)])
def test_median_equals_scipy(dt, Ns):
    src = random_volume((7, 9, 11), dt, seed=sum(Ns))
    got = O.median_box_3d(src, dt, *Ns)
    size = tuple(2 * ((n - 1) // 2) + 1 for n in Ns[::-1])
    assert np.array_equal(got, ndimage.median_filter(src, size=size, mode="nearest"))


def test_median_converts_to_dst_before_ranking():
    src = random_volume((5, 6, 7), np.int16, seed=1)
    got = O.median_box_3d(src, np.uint8, 3, 3, 3)
    conv = np.clip(src, 0, 255).astype(np.uint8)
    assert np.array_equal(got, ndimage.median_filter(conv, size=3, mode="nearest"))


def test_median_rejects_empty_box():
    with pytest.raises(ValueError):
        O.median_box_3d(np.zeros((2, 2, 2), np.uint8), np.uint8, -1, 3, 3)


# ---- float medians over NaN and -0: the insertion sort's closed form --------------
# The reference sorts with `temp < a[j]` (median_box_3d_x.cl:3-14).  A NaN
# compares false both ways, so it never moves and nothing moves past it; -0
# and +0 compare equal, so they keep their order.  The sorted array is the
# window split at its NaNs, each run sorted stably.  These tests build that
# array with Python's stable sort, not with an insertion sort.

def _closed_form_median(vals):
    out, run = [], []
    for v in vals:
        if v != v:
            out += sorted(run)
            out.append(v)
            run = []
        else:
            run.append(v)
    out += sorted(run)
    return out[len(vals) // 2]


def _median_by_closed_form(a, offsets):
    """offsets: (dx, dy, dz) in the reference's gather order (x outer, z
    inner); clamped reads."""
    D, H, W = a.shape
    out = np.empty_like(a)
    for k in range(D):
        for j in range(H):
            for i in range(W):
                vals = [a[min(max(k + dz, 0), D - 1), min(max(j + dy, 0), H - 1), min(max(i + dx, 0), W - 1)]
                        for dx, dy, dz in offsets]
                out[k, j, i] = _closed_form_median(vals)
    return out


def _box_offsets(ex, ey, ez):
    return [(x, y, z) for x in range(-ex, ex + 1) for y in range(-ey, ey + 1) for z in range(-ez, ez + 1)]


@pytest.mark.parametrize("kind", ["ties", "sparse", "slab"])
@pytest.mark.parametrize("Ns", [(3, 3, 3), (5, 3, 1), (1, 1, 3), (4, 2, 5)])
def test_median_nan_negzero_closed_form(kind, Ns):
    a = nanzero_volume((6, 7, 9), kind, seed=sum(Ns))
    ex, ey, ez = ((n - 1) // 2 for n in Ns)
    want = _median_by_closed_form(a, _box_offsets(ex, ey, ez))
    assert same_float_bits(O.median_box_3d(a, np.float32, *Ns), want)
    want2 = _median_by_closed_form(a, _box_offsets(ex, ey, 0))
    assert same_float_bits(O.median_slice_by_slice_box_3d(a, np.float32, Ns[0], Ns[1]), want2)
    assert same_float_bits(O.median_box_2d(a[2], np.float32, Ns[0], Ns[1]), want2[2])


@pytest.mark.parametrize("Ns", [(3, 3, 3), (5, 5, 5), (5, 3, 1)])
def test_sphere_median_nan_negzero_closed_form(Ns):
    a = nanzero_volume((6, 7, 9), "ties", seed=7 + sum(Ns))
    ex, ey, ez = ((n - 1) // 2 for n in Ns)
    fp = _sphere_footprint(*Ns)
    offs = [(x, y, z) for x, y, z in _box_offsets(ex, ey, ez) if fp[z + ez, y + ey, x + ex]]
    want = _median_by_closed_form(a, offs)
    assert same_float_bits(O.sphere_3d("median_sphere_3d", a, np.float32, *Ns), want)


def test_median_signed_zero_known_answers():
    """Worked by hand: all -0 gives -0 everywhere; in a row of +0 with one
    -0 the -0 is the median only where it is the middle member of the
    window (ties keep gather order); NaNs stay where they are."""
    a = np.full((2, 3, 4), -0.0, dtype=np.float32)
    got = O.median_box_3d(a, np.float32, 3, 3, 3)
    assert np.all(np.signbit(got)) and np.all(got == 0)
    # row [+0, +0, -0, +0, +0], window 3x1x1 (members x-1, x, x+1; t = 1):
    # x=1: [+0, +0, -0] -> +0; x=2: [+0, -0, +0] -> -0; x=3: [-0, +0, +0] -> +0
    row = np.array([[[0.0, 0.0, -0.0, 0.0, 0.0]]], dtype=np.float32)
    got = O.median_box_3d(row, np.float32, 3, 1, 1)[0, 0]
    assert list(np.signbit(got)) == [False, False, True, False, False]
    # a NaN splits the window: [5, NaN, 1] stays as it is -> NaN at t = 1;
    # [3, 1, NaN, 2, 0] -> [1, 3, NaN, 0, 2] -> NaN; [2, 1, 0, NaN, 9] -> [0, 1, 2, NaN, 9] -> 2
    r = np.array([[[5.0, np.nan, 1.0]]], dtype=np.float32)
    assert np.isnan(O.median_box_3d(r, np.float32, 3, 1, 1)[0, 0, 1])
    r = np.array([[[3.0, 1.0, np.nan, 2.0, 0.0]]], dtype=np.float32)
    assert np.isnan(O.median_box_3d(r, np.float32, 5, 1, 1)[0, 0, 2])
    r = np.array([[[2.0, 1.0, 0.0, np.nan, 9.0]]], dtype=np.float32)
    assert O.median_box_3d(r, np.float32, 5, 1, 1)[0, 0, 2] == 2.0


# ---- the GPU's divide formula equals the IEEE divide -----------------------------------

def test_div_uniform_equals_ieee_divide():
    rng = np.random.default_rng(0)
    L = O.lib()
    bs = np.concatenate([np.arange(1, 200, dtype=np.float32),
                         rng.uniform(1, 60, 200).astype(np.float32)])
    for b in bs:
        r = np.float32(1.0) / np.float32(b)
        a_s = np.concatenate([rng.uniform(-70000, 70000, 40), rng.integers(-2**24, 2**24, 40),
                              [0.0, -0.0, 1e-40, -3e-39, 1e38, float("inf"), float("nan")]])
        for a in a_s.astype(np.float32):
            got = np.float32(L.clko_div_uniform(float(a), float(b), float(r)))
            want = np.float32(a) / np.float32(b)
            assert (np.isnan(got) and np.isnan(want)) or got.tobytes() == want.tobytes(), (a, b)


def test_div_uniform_whole_binade_of_numerators():
    """Every float numerator of a binade against the IEEE divide, for the
    Gaussian weight sums of sigma 0.5..4 (N = int(8 sigma) made odd, the
    host chain's rule), the mean's tap counts 1..33 and divisors whose
    mantissas lie next to 1 and 2 (all 2^23 numerators each; the formula is
    invariant under power-of-two scaling of either operand)."""
    L = O.lib()
    bs = [np.float32(n) for n in range(1, 34)]
    for sigma in (0.5, 0.8, 1.0, 1.5, 2.0, 2.5, 3.0, 3.7, 4.0):
        N = int(8 * sigma)
        N += 1 - N % 2
        c = (N - 1) // 2
        n = np.float32(-2) * np.float32(sigma) * np.float32(sigma)
        h = np.float32(0)
        for v in range(-c, c + 1):
            h = np.float32(h + L.clko_expf(np.float32(v * v) / n))
        bs.append(h)
    one = np.float32(1).view(np.uint32)
    bs += [np.uint32(one + k).view(np.float32) for k in (1, 2, 3, 0x7FFFFF, 0x7FFFFE, 0x400000, 0x2AAAAB)]
    for b in bs:
        assert L.clko_div_uniform_binade_mismatches(float(b)) == 0, b


# ---- the 2D siblings -----------------------------------------------------------

def _specials_2d(dt, shape, seed):
    """Seeded 2D data; float images also hold NaN, +-inf and +-0 so that the
    2D kernels' centre skip (min/max) meets order-sensitive values."""
    a = random_volume(shape, dt, seed)
    if np.dtype(dt).kind == "f":
        flat = a.reshape(-1)
        flat[::7] = np.nan
        flat[3::11] = np.inf
        flat[5::13] = -np.inf
        flat[2::17] = -0.0
        flat[4::19] = 0.0
    return a


@pytest.mark.parametrize("op", ["gaussian_blur_separable_2d", "mean_separable_2d", "minimum_separable_2d",
                                "maximum_separable_2d"])
@pytest.mark.parametrize("src_t", DTYPES)
@pytest.mark.parametrize("dst_t", [np.uint8, np.int16, np.float32])
def test_2d_separable_equals_3d_at_depth_one(op, src_t, dst_t):
    """The 2D kernels restated from their own text (dir = (dim==0, dim==1),
    min/max without the second centre tap) give the 3D kernels' results on a
    one-plane volume for every dim, N and value, NaN and infinities included:
    the equivalence the library's 2D entry points rely on."""
    a = _specials_2d(src_t, (9, 23), seed=zlib.crc32((op + str(src_t)).encode()))
    op3 = op.replace("_2d", "_3d")
    for dim in (0, 1, 2, 5, -1):
        for N, s in ((1, 1.0), (3, 0.8), (4, 1.3), (7, 2.0), (-3, 1.0)):
            got = O.separable_2d(op, a, dst_t, dim, N, s)
            want = O.separable(op3, a[None], dst_t, dim, N, s)[0]
            assert same_bits_or_value(got, want), (op, dim, N)


def test_2d_median_casts_integers_like_c():
    """median_box_2d stores (dst_t)READ(src).x: integers narrow to their low
    bits (numpy's astype is the same two's-complement cast), unlike the 3D
    kernel's saturating path."""
    a = np.array([[-5, 300, 7, -129], [70000 % 65536 - 32768, 2, 3, 255]], dtype=np.int16)
    got = O.median_box_2d(a, np.uint8, 1, 1)
    assert got.tolist() == a.astype(np.uint8).tolist()
    assert got[0, 0] == 251 and got[0, 1] == 44
    assert O.median_box_3d(a[None], np.uint8, 1, 1, 1)[0, 0, 0] == 0  # 3D saturates


@pytest.mark.parametrize("src_t", [np.int8, np.uint8, np.int16, np.uint16])
@pytest.mark.parametrize("dst_t", [np.int8, np.uint8, np.int16, np.uint16])
@pytest.mark.parametrize("Ns", [(3, 3), (5, 3), (1, 7), (4, 2)])
def test_2d_median_equals_cast_then_scipy(src_t, dst_t, Ns):
    """Integer 2D medians = scipy's median filter of the C-cast image
    (brute force by an independent library)."""
    a = random_volume((11, 17), src_t, seed=Ns[0] * 10 + Ns[1])
    Nx, Ny = Ns
    ex, ey = (Nx - 1) // 2, (Ny - 1) // 2
    want = ndimage.median_filter(a.astype(dst_t), size=(2 * ey + 1, 2 * ex + 1), mode="nearest")
    assert np.array_equal(O.median_box_2d(a, dst_t, Nx, Ny), want)


def test_2d_median_float_source_saturates_like_3d():
    a = np.array([[-5.7, 300.2, 7.9, np.nan]], dtype=np.float32)
    assert np.array_equal(O.median_box_2d(a, np.uint8, 1, 1), O.median_box_3d(a[None], np.uint8, 1, 1, 1)[0])


# ---- ellipsoid neighbourhoods, slice medians, copy -------------------------------

def _sphere_footprint(Nx, Ny, Nz):
    """The reference's mask (minimum_sphere_3d_x.cl:16-37) in float32: a
    (z, y, x) footprint for scipy."""
    ex, ey, ez = (Nx - 1) // 2, (Ny - 1) // 2, (Nz - 1) // 2
    f32 = np.float32

    def sq(e):
        a = f32(e * e)
        return a if a != 0 else f32(np.finfo(np.float32).tiny)
    a2, b2, c2 = sq(ex), sq(ey), sq(ez)
    fp = np.zeros((2 * ez + 1, 2 * ey + 1, 2 * ex + 1), dtype=bool)
    for x in range(-ex, ex + 1):
        for y in range(-ey, ey + 1):
            for z in range(-ez, ez + 1):
                t = f32(x * x) / a2 + f32(y * y) / b2
                t = t + f32(z * z) / c2
                fp[z + ez, y + ey, x + ex] = t <= f32(1.0)
    return fp


@pytest.mark.parametrize("Ns,count", [((3, 3, 3), 7), ((5, 5, 5), 33), ((3, 1, 1), 3), ((1, 1, 1), 1),
                                      ((5, 3, 1), 7)])
def test_sphere_neighbourhood_sizes(Ns, count):
    """Closed forms: the lattice points of the ball of radius 1 (7) and 2
    (33), a segment, a point, an ellipse with semi-axes 2 and 1 (7) -- via
    the maximum of an impulse, which marks exactly the neighbourhood."""
    a = np.zeros((9, 9, 9), dtype=np.float32)
    a[4, 4, 4] = 1
    assert int(O.sphere_3d("maximum_sphere_3d", a, np.float32, *Ns).sum()) == count
    assert int(_sphere_footprint(*Ns).sum()) == count


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("Ns", [(3, 3, 3), (5, 5, 5), (7, 5, 3), (1, 3, 5)])
def test_sphere_minmax_equals_scipy(dt, Ns):
    a = random_volume((9, 11, 13), dt, seed=sum(Ns))
    fp = _sphere_footprint(*Ns)
    for op, ref in (("minimum_sphere_3d", ndimage.minimum_filter), ("maximum_sphere_3d", ndimage.maximum_filter)):
        want = ref(a, footprint=fp, mode="nearest")
        assert np.array_equal(O.sphere_3d(op, a, dt, *Ns), want), op


@pytest.mark.parametrize("src_t", [np.int8, np.uint8, np.int16, np.uint16])
@pytest.mark.parametrize("dst_t", DTYPES)
def test_sphere_mean_integer_source_exact(src_t, dst_t):
    """Integer sums are exact in float32 here, so the mean is the float32
    quotient of scipy's sum by the neighbourhood size, then CONVERT."""
    a = random_volume((7, 9, 11), src_t, seed=5)
    Ns = (5, 3, 5)
    fp = _sphere_footprint(*Ns)
    s = ndimage.correlate(a.astype(np.float64), fp.astype(np.float64), mode="nearest")
    q = (s.astype(np.float32) / np.float32(fp.sum())).astype(np.float32)
    want = O.convert(q, dst_t)
    assert same_bits_or_value(O.sphere_3d("mean_sphere_3d", a, dst_t, *Ns), want)


def test_sphere_mean_float_source_close_to_fp64():
    a = random_volume((7, 9, 11), np.float32, seed=6)
    Ns = (5, 5, 3)
    fp = _sphere_footprint(*Ns)
    want = ndimage.correlate(a.astype(np.float64), fp.astype(np.float64), mode="nearest") / fp.sum()
    got = O.sphere_3d("mean_sphere_3d", a, np.float32, *Ns)
    assert np.max(np.abs(got - want)) <= 1e-5 * np.max(np.abs(a))


@pytest.mark.parametrize("src_t", DTYPES)
@pytest.mark.parametrize("dst_t", [np.uint8, np.int16, np.float32])
@pytest.mark.parametrize("Ns", [(3, 3, 3), (5, 5, 5), (5, 3, 1)])
def test_sphere_median_equals_scipy(src_t, dst_t, Ns):
    """Rank count/2 of the dst-converted values in the neighbourhood."""
    a = random_volume((7, 9, 11), src_t, seed=11)
    fp = _sphere_footprint(*Ns)
    want = ndimage.median_filter(O.convert(a.astype(np.float32), dst_t), footprint=fp, mode="nearest")
    assert np.array_equal(O.sphere_3d("median_sphere_3d", a, dst_t, *Ns), want)


def test_sphere_degenerate_sizes():
    a = random_volume((4, 5, 6), np.int16, seed=2)
    # negative N: an empty neighbourhood; min/max keep the centre, the mean is 0/0
    assert np.array_equal(O.sphere_3d("minimum_sphere_3d", a, np.int16, -3, -3, -3), a)
    assert np.array_equal(O.sphere_3d("mean_sphere_3d", a, np.int16, -3, -3, -3), np.zeros_like(a))
    with pytest.raises(ValueError):
        O.sphere_3d("median_sphere_3d", a, np.int16, -1, 3, 3)


@pytest.mark.parametrize("src_t", [np.int16, np.uint16, np.float32])
@pytest.mark.parametrize("dst_t", [np.uint8, np.int8])
def test_slice_median_equals_cast_then_scipy(src_t, dst_t):
    a = random_volume((4, 9, 13), src_t, seed=9)
    cast = a.astype(dst_t) if np.dtype(src_t).kind != "f" else O.convert(a, dst_t)
    want = ndimage.median_filter(cast, size=(1, 3, 5), mode="nearest")
    assert np.array_equal(O.median_slice_by_slice_box_3d(a, dst_t, 5, 3), want)


@pytest.mark.parametrize("src_t", DTYPES)
@pytest.mark.parametrize("dst_t", DTYPES)
def test_copy_is_convert(src_t, dst_t):
    a = random_volume((3, 4, 5), src_t, seed=1)
    assert same_bits_or_value(O.copy_3d(a, dst_t), O.convert(a.astype(np.float32), dst_t))


# ---- consumers --------------------------------------------------------------------

def _detect_maxima_brute(a):
    """detect_maxima_3d by brute force in Python (small volumes): the first
    strict maximum over offsets in x, y, z order, from centre - 1, with the
    unclamped position."""
    D, H, W = a.shape
    out = np.zeros(a.shape, dtype=np.uint8)
    f = a.astype(np.float64)
    for k in range(D):
        for j in range(H):
            for i in range(W):
                best, pos = np.float32(f[k, j, i]) - np.float32(1), (0, 0, 0)
                for x in (-1, 0, 1):
                    for y in (-1, 0, 1):
                        for z in (-1, 0, 1):
                            v = f[min(max(k + z, 0), D - 1), min(max(j + y, 0), H - 1), min(max(i + x, 0), W - 1)]
                            if v > best:
                                best, pos = v, (x, y, z)
                out[k, j, i] = pos == (0, 0, 0)
    return out


@pytest.mark.parametrize("dt", DTYPES)
def test_detect_maxima_equals_brute_force(dt):
    a = random_volume((4, 5, 6), dt, seed=8, lo=0 if np.dtype(dt).kind == "f" else None,
                      hi=5 if np.dtype(dt).kind == "f" else None)
    if np.dtype(dt).kind != "f":
        a = (a % 4).astype(dt)  # many ties
    assert np.array_equal(O.detect_maxima_3d(a, np.uint8), _detect_maxima_brute(a))


def test_detect_maxima_known_answers():
    # a constant volume has no maxima: the first neighbour already equals it
    assert not O.detect_maxima_3d(np.full((3, 4, 5), 7, np.uint16), np.uint8).any()
    # one interior peak; a border voxel meets its own clamped value first
    a = np.zeros((5, 5, 5), np.float32)
    a[2, 2, 2] = 3
    got = O.detect_maxima_3d(a, np.float32)
    assert got[2, 2, 2] == 1 and got.sum() == 1
    b = np.zeros((5, 5, 5), np.float32)
    b[0, 0, 0] = 3
    assert O.detect_maxima_3d(b, np.uint8)[0, 0, 0] == 0
    # a NaN centre never sees a greater value: it stays a "maximum"
    c = np.zeros((3, 3, 3), np.float32)
    c[1, 1, 1] = np.nan
    assert O.detect_maxima_3d(c, np.uint8)[1, 1, 1] == 1


@pytest.mark.parametrize("src_t", DTYPES)
@pytest.mark.parametrize("dst_t", DTYPES)
def test_maximum_z_projection_equals_numpy(src_t, dst_t):
    a = random_volume((7, 5, 6), src_t, seed=3)
    want = O.convert(a.max(axis=0, keepdims=True).astype(np.float32), dst_t)
    assert same_bits_or_value(O.maximum_z_projection(a, dst_t), want)


@pytest.mark.parametrize("op", [O.OP_GAUSS, O.OP_MEAN, O.OP_MIN, O.OP_MAX])
def test_fast_build_gives_the_test_build_bits(op):
    """The -O3 x86-64-v3 build that bench.py times computes exactly what the
    test build computes (no contraction, no fast-math), specials
    included: the same bits everywhere except that a NaN may carry another
    sign or payload."""
    rng = np.random.default_rng(11 + op)
    srcs = [rng.integers(0, 65535, (12, 17, 29)).astype(np.uint16),
            rng.uniform(-1e3, 1e3, (12, 17, 29)).astype(np.float32)]
    srcs[1].flat[::97] = np.nan
    srcs[1].flat[5::89] = -0.0
    srcs[1].flat[7::83] = np.inf
    for src in srcs:
        for dst_t in (np.float32, np.uint16, np.int8):
            for dim in (0, 1, 2):
                for N, s in ((25, 3.0), (7, 0.0), (13, 1.5)):
                    a = O.separable(op, src, dst_t, dim, N, s)
                    b = O.separable(op, src, dst_t, dim, N, s, fast=True)
                    if a.dtype.kind == "f":
                        na, nb = np.isnan(a), np.isnan(b)
                        assert np.array_equal(na, nb), (op, src.dtype, dst_t, dim, N)
                        a, b = a[~na], b[~nb]
                    assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (op, src.dtype, dst_t, dim, N)
