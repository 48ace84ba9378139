"""GPU parity: every result of libclkmi on cuda:0 must equal the oracle's
(bit-exact for integer outputs; float outputs equal by value, i.e. +0 == -0
and NaN == NaN) on the same seeded inputs."""
import zlib

import numpy as np
import pytest

import clk_oracle as O
from helpers import DTYPES, DTYPE_NAMES, empty_device, from_device, random_volume, same_bits_or_value, to_device

pytestmark = pytest.mark.gpu

SEP_OPS = ["gaussian_blur_separable_3d", "mean_separable_3d", "minimum_separable_3d",
           "maximum_separable_3d"]


def gpu_separable(name, src, dst_t, dim, N, s):
    import clkmi
    ts, si = to_device(src)
    td, di = empty_device(src.shape, dst_t)
    getattr(clkmi, name)(di, si, dim, N, s)
    return from_device(td, src.shape, dst_t)


def gpu_median(src, dst_t, Nx, Ny, Nz):
    import clkmi
    ts, si = to_device(src)
    td, di = empty_device(src.shape, dst_t)
    clkmi.median_box_3d(di, si, Nx, Ny, Nz)
    return from_device(td, src.shape, dst_t)


def assert_parity(got, want, ctx):
    if not same_bits_or_value(got, want):
        bad = np.argwhere(~((got == want) | (np.isnan(got.astype(np.float64)) & np.isnan(want.astype(np.float64)))))
        i = tuple(bad[0])
        raise AssertionError("%s: %d mismatches, first at (z,y,x)=%s got %r want %r"
                             % (ctx, len(bad), i, got[i], want[i]))


@pytest.mark.parametrize("name", SEP_OPS)
@pytest.mark.parametrize("src_t", DTYPES, ids=lambda t: DTYPE_NAMES[t])
@pytest.mark.parametrize("dst_t", DTYPES, ids=lambda t: DTYPE_NAMES[t])
def test_separable_all_type_pairs(gpu, name, src_t, dst_t):
    src = random_volume((9, 13, 21), src_t, seed=zlib.crc32(("%s%s%s" % (name, np.dtype(src_t), np.dtype(dst_t))).encode()) & 0xffff)
    for dim in (0, 1, 2, 5):
        got = gpu_separable(name, src, dst_t, dim, 7, 1.3)
        want = O.separable(name, src, dst_t, dim, 7, 1.3)
        assert_parity(got, want, "%s dim=%d" % (name, dim))


@pytest.mark.parametrize("name", SEP_OPS)
@pytest.mark.parametrize("N", [-3, 0, 1, 2, 3, 5, 8, 9, 16, 17, 25, 33, 64, 131])
def test_separable_tap_counts(gpu, name, N):
    src = random_volume((19, 23, 37), np.uint16, seed=N + 500)
    srcf = random_volume((19, 23, 37), np.float32, seed=N + 600)
    for dim in (0, 1, 2):
        for s_, dt in ((src, np.float32), (src, np.uint16), (srcf, np.float32)):
            got = gpu_separable(name, s_, dt, dim, N, 2.2)
            want = O.separable(name, s_, dt, dim, N, 2.2)
            assert_parity(got, want, "%s N=%d dim=%d %s->%s" % (name, N, dim, s_.dtype, np.dtype(dt)))


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 1, 5), (3, 1, 1), (2, 3, 1), (1, 600, 3),
                                   (4, 4, 1030), (3, 2, 513), (70, 5, 8), (2, 300, 4)])
def test_separable_odd_shapes(gpu, shape):
    src = random_volume(shape, np.int16, seed=sum(shape))
    for name in SEP_OPS:
        for dim in (0, 1, 2):
            for N in (3, 11):
                got = gpu_separable(name, src, np.float32, dim, N, 1.7)
                want = O.separable(name, src, np.float32, dim, N, 1.7)
                assert_parity(got, want, "%s %s dim=%d N=%d" % (name, shape, dim, N))


def test_separable_large_halo_along_x(gpu):
    # c > 1024 along x takes the strided kernel with one row per lane.
    src = random_volume((2, 3, 2100), np.uint8, seed=3)
    for name in ("mean_separable_3d", "maximum_separable_3d"):
        got = gpu_separable(name, src, np.uint8, 0, 2301, 0.0)
        want = O.separable(name, src, np.uint8, 0, 2301, 0.0)
        assert_parity(got, want, name)


def test_float_special_values(gpu):
    src = random_volume((5, 6, 40), np.float32, seed=8)
    src[0, 0, 3] = np.inf
    src[1, 2, 10] = -np.inf
    src[2, 3, 20] = np.nan
    src[3, 4, 5] = -0.0
    src[4, 5, 6] = 1e-42
    for name in SEP_OPS:
        for dim in (0, 1, 2):
            got = gpu_separable(name, src, np.float32, dim, 5, 1.0)
            want = O.separable(name, src, np.float32, dim, 5, 1.0)
            assert_parity(got, want, "%s dim=%d" % (name, dim))
            got = gpu_separable(name, src, np.int16, dim, 5, 1.0)
            want = O.separable(name, src, np.int16, dim, 5, 1.0)
            assert_parity(got, want, "%s dim=%d int16" % (name, dim))


@pytest.mark.parametrize("src_t", DTYPES, ids=lambda t: DTYPE_NAMES[t])
@pytest.mark.parametrize("dst_t", DTYPES, ids=lambda t: DTYPE_NAMES[t])
def test_median_all_type_pairs(gpu, src_t, dst_t):
    src = random_volume((7, 9, 13), src_t, seed=77)
    for Ns in ((3, 3, 3), (5, 5, 5), (3, 1, 5), (1, 1, 1), (4, 2, 3)):
        got = gpu_median(src, dst_t, *Ns)
        want = O.median_box_3d(src, dst_t, *Ns)
        assert_parity(got, want, "median %s" % (Ns,))


def test_median_salt_and_pepper(gpu):
    rng = np.random.default_rng(3)
    src = rng.integers(40, 200, size=(16, 40, 50), endpoint=True).astype(np.uint8)
    m = rng.random(src.shape) < 0.05
    src[m] = np.where(rng.random(m.sum()) < 0.5, 0, 255)
    for r in (1, 2):
        n = 2 * r + 1
        got = gpu_median(src, np.uint8, n, n, n)
        want = O.median_box_3d(src, np.uint8, n, n, n)
        assert_parity(got, want, "median r=%d" % r)


@pytest.mark.parametrize("src_t", (np.uint8, np.int16, np.uint16, np.float32), ids=lambda t: DTYPE_NAMES[t])
def test_median_register_tiers(gpu, src_t):
    """Boxes within 7x7x1 or 5x5x5 and 16/32-bit dst keep the window in
    registers (one tier per box size), boxes just outside every tier and
    sparse ones (a quarter of the tier's box or less) use the generic search.
    Float data carries -0 and +0 side by side, and values beyond the dst
    range (saturating conversion)."""
    src = random_volume((7, 13, 70), src_t, seed=91)
    if src_t == np.float32:
        src[:, ::3, ::5] = 0.0
        src[:, 1::3, ::5] = -0.0
        src[2, 4:9, 10:60] = 1e6
    for Ns in ((3, 3, 1), (5, 5, 1), (7, 7, 1), (1, 3, 3), (3, 5, 3), (5, 5, 3), (5, 3, 5), (5, 5, 5),
               (7, 5, 1), (1, 1, 5), (7, 7, 3), (3, 1, 1), (6, 4, 2)):
        for dst_t in (np.int16, np.uint16, np.float32):
            got = gpu_median(src, dst_t, *Ns)
            assert_parity(got, O.median_box_3d(src, dst_t, *Ns),
                          "median %s %s->%s" % (Ns, DTYPE_NAMES[src_t], DTYPE_NAMES[dst_t]))


@pytest.mark.parametrize("src_t", (np.uint8, np.int8, np.float32), ids=lambda t: DTYPE_NAMES[t])
def test_median_hist8_wide_rows(gpu, src_t):
    # 8-bit dst with a window beyond 3x3x3 takes the histogram kernel: rows of
    # 150 pixels span three 64-lane blocks (interior, halo lanes 0/63, a
    # partial last block) and depth 70 spans three z segments
    src = random_volume((70, 5, 150), src_t, seed=29)
    for Ns in ((5, 5, 5), (3, 5, 7), (7, 3, 5), (9, 3, 3), (5, 5, 1)):
        for dst_t in (np.uint8, np.int8):
            got = gpu_median(src, dst_t, *Ns)
            want = O.median_box_3d(src, dst_t, *Ns)
            assert_parity(got, want, "median %s -> %s %s" % (DTYPE_NAMES[src_t], DTYPE_NAMES[dst_t], Ns))


@pytest.mark.parametrize("src_t,dst_t", [(np.uint8, np.uint8), (np.int8, np.int8), (np.uint16, np.uint8),
                                         (np.float32, np.int8)], ids=lambda t: DTYPE_NAMES[t])
def test_median_hist8_column_histograms(gpu, src_t, dst_t):
    """8-bit dst medians with x and y half widths up to 3 keep one histogram
    per window column (64 - 2 ex outputs per 64-lane block): rows narrower
    than a block, one and two pixels wide, spanning several blocks; depths
    of one plane and over several 64-plane segments; data that is constant
    at either end of the range (the median walk's bounds) and salt and
    pepper.  Equal to the per-output histograms and to the oracle."""
    import clkmi
    rng = np.random.default_rng(41)
    cases = []
    for shape in ((70, 5, 150), (3, 4, 1), (2, 3, 2), (1, 9, 61), (66, 3, 59), (5, 7, 121)):
        cases.append(random_volume(shape, src_t, seed=int(rng.integers(1 << 30))))
    lo, hi = (0, 255) if np.dtype(src_t).kind == "u" else (-128, 127)
    cases.append(np.full((9, 6, 70), lo, dtype=src_t))
    cases.append(np.full((9, 6, 70), hi, dtype=src_t))
    sp = rng.integers(40, 200, size=(20, 8, 90)).astype(src_t)
    m = rng.random(sp.shape) < 0.3
    sp[m] = np.where(rng.random(int(m.sum())) < 0.5, lo, hi).astype(src_t)
    cases.append(sp)
    old = clkmi.get_tuning("hist8_rows")
    try:
        for src in cases:
            for Ns in ((5, 5, 5), (3, 3, 5), (7, 7, 3), (5, 3, 7), (3, 7, 5), (7, 5, 1)):
                clkmi.set_tuning("hist8_rows", 0)
                got = gpu_median(src, dst_t, *Ns)
                clkmi.set_tuning("hist8_rows", 1)
                rows = gpu_median(src, dst_t, *Ns)
                tag = "median %s %s %s->%s" % (src.shape, Ns, DTYPE_NAMES[src_t], DTYPE_NAMES[dst_t])
                assert_parity(got, rows, tag + " columns vs rows")
                assert_parity(got, O.median_box_3d(src, dst_t, *Ns), tag)
    finally:
        clkmi.set_tuning("hist8_rows", old)


@pytest.mark.parametrize("src_t,dst_t", [(np.uint8, np.uint8), (np.int16, np.int16), (np.float32, np.uint16),
                                         (np.uint16, np.int8)], ids=lambda t: DTYPE_NAMES[t])
def test_median_3x3x3_packed_z_runs(gpu, src_t, dst_t):
    """The packed 3x3x3 median of single launches slides along z in runs of
    16 planes: depth 37 (three runs, the last partial), width 1030 (three
    512-wide tiles, the last partial), height 5."""
    a = random_volume((37, 5, 1030), src_t, seed=71)
    got = gpu_median(a, dst_t, 3, 3, 3)
    assert_parity(got, O.median_box_3d(a, dst_t, 3, 3, 3), "median 3x3x3 %s->%s" % (DTYPE_NAMES[src_t], DTYPE_NAMES[dst_t]))


@pytest.mark.parametrize("src_t", (np.float32, np.uint8, np.int16), ids=lambda t: DTYPE_NAMES[t])
def test_median_3x3x3_float_dst_z_runs(gpu, src_t):
    """The float-dst 3x3x3 median of single launches stages 32-bit keys and
    slides along z in runs of 16 planes: depth 37 (three runs, the last
    partial), width 600 (three 256-wide tiles, the last partial), depths 1
    and 2, and float data with -0 beside +0."""
    for shape in ((37, 5, 600), (1, 9, 300), (2, 3, 257), (17, 19, 33)):
        a = random_volume(shape, src_t, seed=73 + shape[0])
        if src_t == np.float32:
            a[:, ::2, ::3] = 0.0
            a[:, 1::2, ::3] = -0.0
        got = gpu_median(a, np.float32, 3, 3, 3)
        assert_parity(got, O.median_box_3d(a, np.float32, 3, 3, 3),
                      "median 3x3x3 %s %s->f32" % (shape, DTYPE_NAMES[src_t]))


def test_execute_by_name_matches_typed_call(gpu):
    import clkmi
    src = random_volume((6, 7, 8), np.uint16, seed=1)
    ts, si = to_device(src)
    for name in SEP_OPS:
        td, di = empty_device(src.shape, np.float32)
        clkmi.execute(name + "_x.cl", {"dst": di, "src": si, "dim": 1, "N": 5, "s": 1.0})
        got = from_device(td, src.shape, np.float32)
        assert_parity(got, O.separable(name, src, np.float32, 1, 5, 1.0), name)
    td, di = empty_device(src.shape, np.uint16)
    clkmi.execute("median_box_3d", {"src": si, "dst": di, "Nx": 3, "Ny": 3, "Nz": 3})
    assert_parity(from_device(td, src.shape, np.uint16), O.median_box_3d(src, np.uint16, 3, 3, 3), "median")


def test_host_chain_matches_oracle_chain(gpu):
    import clkmi
    src = random_volume((20, 31, 45), np.uint16, seed=12, hi=3000)
    ts, si = to_device(src)
    td, di = empty_device(src.shape, np.float32)
    clkmi.gaussian_blur_3d(si, di, 2.0, 2.0, 1.0)
    want = O.gaussian_blur_3d(src, np.float32, 2.0, 2.0, 1.0)
    assert_parity(from_device(td, src.shape, np.float32), want, "gaussian chain")
    td, di = empty_device(src.shape, np.float32)
    clkmi.gaussian_blur_3d(si, di, 0.0, 1.5, 0.0)
    want = O.gaussian_blur_3d(src, np.float32, 0.0, 1.5, 0.0)
    assert_parity(from_device(td, src.shape, np.float32), want, "gaussian chain, y only")
    for fn, op in (("mean_box_3d", O.OP_MEAN), ("minimum_box_3d", O.OP_MIN),
                   ("maximum_box_3d", O.OP_MAX)):
        td, di = empty_device(src.shape, np.uint16)
        getattr(clkmi, fn)(si, di, 3, 2, 1)
        want = O.box_3d(op, src, np.uint16, 3, 2, 1)
        assert_parity(from_device(td, src.shape, np.uint16), want, fn)


def test_host_chain_kept_temporaries_across_calls_and_streams(gpu):
    """The 3D chain keeps its temporaries per stream and reuses them: calls
    of shrinking, growing and repeated sizes queued back to back on two
    streams (no synchronisation between them) each equal the oracle."""
    import torch
    import clkmi
    shapes = [(9, 40, 70), (5, 12, 17), (14, 50, 90), (5, 12, 17), (14, 50, 90), (3, 7, 300)]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    keep, checks = [], []
    for i, shape in enumerate(shapes):
        src = random_volume(shape, np.uint16, seed=40 + i, hi=3000)
        ts, si = to_device(src)
        st = streams[i % 2]
        st.wait_stream(torch.cuda.current_stream())  # the upload
        tg, dg = empty_device(shape, np.float32)
        tm, dm = empty_device(shape, np.uint16)
        with torch.cuda.stream(st):
            clkmi.gaussian_blur_3d(si, dg, 1.5, 1.0, 0.8, stream=st)
            clkmi.maximum_box_3d(si, dm, 2, 1, 1, stream=st)
        keep += [ts, tg, tm]
        checks.append((src, shape, tg, tm))
    for st in streams:
        st.synchronize()
    for i, (src, shape, tg, tm) in enumerate(checks):
        assert_parity(from_device(tg, shape, np.float32), O.gaussian_blur_3d(src, np.float32, 1.5, 1.0, 0.8),
                      "gaussian call %d %s" % (i, shape))
        assert_parity(from_device(tm, shape, np.uint16), O.box_3d(O.OP_MAX, src, np.uint16, 2, 1, 1),
                      "maximum call %d %s" % (i, shape))


def test_full_size_c1_gaussian_chain(gpu):
    """The reference's CPU-sized case at its full size: float32 256x256x64,
    U[0, 1000) (seed 1), the 3D Gaussian sigma = (2, 2, 1) -> N = (17, 17, 9)
    through the host chain, every voxel against the oracle's chain."""
    import clkmi
    src = np.random.default_rng(1).uniform(0, 1000, size=(64, 256, 256)).astype(np.float32)
    ts, si = to_device(src)
    td, di = empty_device(src.shape, np.float32)
    clkmi.gaussian_blur_3d(si, di, 2.0, 2.0, 1.0)
    want = O.gaussian_blur_3d(src, np.float32, 2.0, 2.0, 1.0)
    assert_parity(from_device(td, src.shape, np.float32), want, "C1 gaussian chain")


def test_full_size_c3_median_crops(gpu):
    """median_box_3d r = 1 and r = 2 on the full uint8 512x512x256
    salt-and-pepper volume, checked on 24^3 crops: the four corners of the
    first and last planes (the volume's borders, where the clamp happens)
    and random interior crops, each compared inside a margin of r against
    the oracle run on the crop (an interior voxel's window lies in the crop;
    at a true border both clamp to the same edge)."""
    import torch
    import clkmi
    rng = np.random.default_rng(3)
    D, H, W = 256, 512, 512
    vol = rng.integers(40, 200, size=(D, H, W), endpoint=True).astype(np.uint8)
    m = rng.random(vol.shape) < 0.05
    vol[m] = np.where(rng.random(int(m.sum())) < 0.5, 0, 255).astype(np.uint8)
    ts, si = to_device(vol)
    C = 24
    starts = [(z, y, x) for z in (0, D - C) for y in (0, H - C) for x in (0, W - C)]
    starts += [tuple(int(rng.integers(0, n - C)) for n in (D, H, W)) for _ in range(8)]
    for r in (1, 2):
        td, di = empty_device(vol.shape, np.uint8)
        clkmi.median_box_3d(di, si, 2 * r + 1, 2 * r + 1, 2 * r + 1)
        out = from_device(td, vol.shape, np.uint8)
        for (z, y, x) in starts:
            crop = vol[z:z + C, y:y + C, x:x + C]
            want = O.median_box_3d(crop, np.uint8, 2 * r + 1, 2 * r + 1, 2 * r + 1)
            # keep the part whose windows do not reach a crop face that is
            # not also a volume face
            sl = []
            for a, n in zip((z, y, x), (D, H, W)):
                lo = 0 if a == 0 else r
                hi = C if a + C == n else C - r
                sl.append(slice(lo, hi))
            sl = tuple(sl)
            assert_parity(out[z:z + C, y:y + C, x:x + C][sl], want[sl], "C3 median r=%d crop %s" % (r, (z, y, x)))
    del ts
    torch.cuda.empty_cache()


def test_errors_are_opencl_codes(gpu):
    import clkmi
    src = random_volume((4, 5, 6), np.uint8, seed=0)
    ts, si = to_device(src)
    td, di = empty_device((4, 5, 7), np.uint8)
    with pytest.raises(clkmi.ClkError) as e:
        clkmi.mean_separable_3d(di, si, 0, 3)
    assert e.value.code == -30
    with pytest.raises(clkmi.ClkError) as e:
        clkmi.mean_separable_3d(si, si, 0, 3)
    assert e.value.code == -8
    bad = clkmi.Image(si.ptr, 9, 6, 5, 4)
    with pytest.raises(clkmi.ClkError) as e:
        clkmi.mean_separable_3d(bad, si, 0, 3)
    assert e.value.code == -30
    with pytest.raises(clkmi.ClkError) as e:
        clkmi.execute("no_such_kernel", {"dst": si, "src": si})
    assert e.value.code == -46
    td, di = empty_device((4, 5, 6), np.uint8)
    with pytest.raises(clkmi.ClkError) as e:
        clkmi.execute("mean_separable_3d", {"dst": di, "src": si, "dim": 0})
    assert e.value.code == -52
    with pytest.raises(clkmi.ClkError) as e:
        clkmi.gaussian_blur_separable_3d(di, si, 0, 2001, 100.0)
    assert e.value.code == -30
    with pytest.raises(clkmi.ClkError) as e:
        clkmi.median_box_3d(di, si, 3, -1, 3)
    assert e.value.code == -30


def _sample_lines(vol_t, dim, idx_a, idx_b):
    """Lines along axis `dim` of a (z, y, x) torch volume at the given
    coordinates of the other two axes, as a numpy (n, 1, L) array laid out
    so the filtered axis is x."""
    import torch
    if dim == 0:
        lines = vol_t[idx_a, idx_b, :]          # (z, y) -> x lines
    elif dim == 1:
        lines = vol_t[idx_a, :, idx_b]          # (z, x) -> y lines
    else:
        lines = vol_t[:, idx_a, idx_b].T        # (y, x) -> z lines
    return lines.contiguous().cpu().numpy()[:, None, :]


@pytest.mark.parametrize("name,N,s,dst", [
    ("gaussian_blur_separable_3d", 25, 3.0, np.float32),
    ("mean_separable_3d", 7, 0.0, np.uint16),
    ("minimum_separable_3d", 7, 0.0, np.uint16),
    ("maximum_separable_3d", 7, 0.0, np.uint16),
])
def test_full_size_sampled_lines(gpu, name, N, s, dst):
    """1024x1024x512 uint16: each single-axis launch checked exactly on 512
    random full lines against the oracle."""
    import torch
    import clkmi
    D, H, W = 512, 1024, 1024
    g = torch.Generator(device="cuda").manual_seed(2)
    # uint16 data held in int16 storage (values < 4096 read the same)
    src = torch.randint(0, 4096, (D, H, W), generator=g, device="cuda", dtype=torch.int16)
    si = clkmi.Image(src.data_ptr(), 3, W, H, D)
    tdt = torch.float32 if dst is np.float32 else torch.int16
    out = torch.empty((D, H, W), device="cuda", dtype=tdt)
    do = clkmi.Image(out.data_ptr(), 4 if dst is np.float32 else 3, W, H, D)
    rng = np.random.default_rng(0)
    for dim in (0, 1, 2):
        getattr(clkmi, name)(do, si, dim, N, s)
        torch.cuda.synchronize()
        sizes = {0: (D, H), 1: (D, W), 2: (H, W)}[dim]
        a = torch.from_numpy(rng.integers(0, sizes[0], 512)).cuda()
        b = torch.from_numpy(rng.integers(0, sizes[1], 512)).cuda()
        src_lines = _sample_lines(src, dim, a, b).view(np.uint16)
        got = _sample_lines(out, dim, a, b)
        if dst is not np.float32:
            got = got.view(np.uint16)
        want = O.separable(name, src_lines, dst, 0, N, s)
        assert_parity(got, want, "%s dim=%d" % (name, dim))
    del src, out
    torch.cuda.empty_cache()


@pytest.mark.parametrize("shape", [(3, 5, 8), (2, 3, 16), (3, 3, 40), (2, 4, 64), (2, 5, 256), (3, 2, 520),
                                   (2, 3, 1024), (1, 2, 1544), (37, 41, 8), (20, 300, 12), (70, 9, 6)])
@pytest.mark.parametrize("pair", [(np.uint16, np.uint16), (np.uint16, np.float32), (np.float32, np.float32),
                                  (np.int8, np.int8), (np.uint8, np.float32), (np.int16, np.int16)],
                         ids=lambda p: "%s-%s" % (DTYPE_NAMES[p[0]], DTYPE_NAMES[p[1]]))
def test_static_width_paths(gpu, shape, pair):
    """Half widths with compile-time kernels: the register-window x pass
    (W % 8 == 0, c = 1..8, Gaussian up to 16 on rows of at most 256 pixels,
    rows packed across a wave, rows shorter than the halo) and the static y/z
    pass (c = 1..8, Gaussian up to 12) -- every half width, short rows,
    partial 512-pixel segments, blocks spanning several slabs, both axis
    ends."""
    src_t, dst_t = pair
    src = random_volume(shape, src_t, seed=shape[2] + 7)
    for name in SEP_OPS:
        cmax = 16 if name.startswith("gaussian") else 8
        for c in range(1, cmax + 1):
            N = 2 * c + 1
            for dim in (0, 1, 2):
                got = gpu_separable(name, src, dst_t, dim, N, 0.4 * c + 0.3)
                want = O.separable(name, src, dst_t, dim, N, 0.4 * c + 0.3)
                assert_parity(got, want, "%s %s dim=%d c=%d" % (name, shape, dim, c))


KNOB_SETTINGS = [("strided_rounds", 2), ("strided_rounds", 8), ("strided_order", 1), ("strided_order", 2),
                 ("strided_order", 3), ("ystatic_cols", 1), ("ystatic_cols", 2), ("x_sub", 1), ("x_sub", 2),
                 ("x_sub", 4), ("no_xreg", 1), ("xreg_maxc", 16), ("strided_variant", 2), ("xreg_units", 2),
                 ("xreg_units", 4), ("x_dynamic", 1), ("x_tasks", 2)]


@pytest.mark.parametrize("name,N,dst_t", [("gaussian_blur_separable_3d", 25, np.float32),
                                          ("maximum_separable_3d", 7, np.uint16)])
def test_launch_shapes_and_autotune_do_not_change_results(gpu, name, N, dst_t):
    """Every launch-shape knob and the first-launch autotuner (on a volume
    large enough to be tuned, 2^22 voxels) give the default launch's bits."""
    import clkmi
    src = random_volume((64, 256, 256), np.uint16, seed=N)
    ref = [gpu_separable(name, src, dst_t, dim, N, 2.5) for dim in (0, 1, 2)]
    for key, val in KNOB_SETTINGS + [("autotune", 1)]:
        old = clkmi.get_tuning(key)
        try:
            clkmi.set_tuning(key, val)
            for rep in range(2 if key == "autotune" else 1):  # tuning launch, then the tuned one
                for dim in (0, 1, 2):
                    got = gpu_separable(name, src, dst_t, dim, N, 2.5)
                    assert same_bits_or_value(got, ref[dim]), "%s=%d dim=%d rep=%d" % (key, val, dim, rep)
        finally:
            clkmi.set_tuning(key, old)


@pytest.mark.parametrize("units", [2, 4])
@pytest.mark.parametrize("shape", [(3, 5, 8), (7, 9, 40), (5, 33, 256), (2, 3, 1024), (4, 37, 64)])
def test_register_x_pass_several_units_per_wave(gpu, shape, units):
    """Waves filtering 2 or 4 consecutive units of 512 outputs (the next
    unit's loads in flight), including grids whose last wave has fewer units."""
    import clkmi
    src = random_volume(shape, np.uint16, seed=units * 100 + shape[2])
    old = clkmi.get_tuning("xreg_units")
    try:
        clkmi.set_tuning("xreg_units", units)
        for name, cs in (("gaussian_blur_separable_3d", (1, 5, 12, 16)), ("mean_separable_3d", (1, 3, 8)),
                         ("maximum_separable_3d", (2, 8))):
            for c in cs:
                if c > 8 and shape[2] > 256:
                    continue  # wider windows take the LDS pass on long rows
                dst_t = np.float32 if name.startswith("g") else np.uint16
                got = gpu_separable(name, src, dst_t, 0, 2 * c + 1, 0.5 * c)
                want = O.separable(name, src, dst_t, 0, 2 * c + 1, 0.5 * c)
                assert_parity(got, want, "%s %s c=%d units=%d" % (name, shape, c, units))
    finally:
        clkmi.set_tuning("xreg_units", old)


@pytest.mark.parametrize("tasks", [1, 2])
@pytest.mark.parametrize("sub", [1, 2])
@pytest.mark.parametrize("shape", [(3, 5, 520), (2, 7, 1544), (4, 3, 1030), (1, 1, 2100)])
def test_static_lds_x_pass(gpu, shape, sub, tasks):
    """The LDS x pass with a compile-time half width (Gaussian c = 9..16 on
    rows longer than 256): 1 or 2 segments of 512 per task, 1 or 2 tasks per
    wave, unaligned row lengths, partial last segments."""
    import clkmi
    keys = {"x_sub": sub, "x_tasks": tasks}
    old = {k: clkmi.get_tuning(k) for k in keys}
    try:
        for k, v in keys.items():
            clkmi.set_tuning(k, v)
        for src_t in (np.uint16, np.float32):
            src = random_volume(shape, src_t, seed=shape[2] + sub)
            for c in (9, 12, 16):
                got = gpu_separable("gaussian_blur_separable_3d", src, np.float32, 0, 2 * c + 1, 0.4 * c)
                want = O.separable("gaussian_blur_separable_3d", src, np.float32, 0, 2 * c + 1, 0.4 * c)
                assert_parity(got, want, "%s %s c=%d sub=%d tasks=%d" % (shape, DTYPE_NAMES[src_t], c, sub, tasks))
    finally:
        for k, v in old.items():
            clkmi.set_tuning(k, v)


# ---- the 2D siblings (images of depth 1) ----------------------------------------

SEP_OPS_2D = ["gaussian_blur_separable_2d", "mean_separable_2d", "minimum_separable_2d", "maximum_separable_2d"]


def _image_2d(a, dtype=None):
    """Device image of depth 1 from a (H, W) array (or an empty one)."""
    if dtype is None:
        return to_device(a[None])
    return empty_device((1,) + tuple(a), dtype)


def _specials_2d(dt, shape, seed):
    a = random_volume(shape, dt, seed)
    if np.dtype(dt).kind == "f":
        flat = a.reshape(-1)
        flat[::7] = np.nan
        flat[3::11] = np.inf
        flat[2::17] = -0.0
    return a


@pytest.mark.parametrize("name", SEP_OPS_2D)
@pytest.mark.parametrize("src_t", DTYPES, ids=lambda t: DTYPE_NAMES[t])
@pytest.mark.parametrize("dst_t", DTYPES, ids=lambda t: DTYPE_NAMES[t])
def test_2d_separable_all_type_pairs(gpu, name, src_t, dst_t):
    """clk_*_separable_2d against the oracle's restatement of the 2D kernels
    (every dim incl. out of range, odd/even N, wide Gaussians)."""
    import clkmi
    a = _specials_2d(src_t, (37, 70), seed=zlib.crc32((name + DTYPE_NAMES[src_t]).encode()))
    ts, si = _image_2d(a)
    for dim in (0, 1, 2, -1):
        for N, s in ((3, 0.8), (4, 1.2), (7, 2.0), (25, 3.0)):
            td, di = _image_2d(a.shape, dst_t)
            getattr(clkmi, name)(di, si, dim, N, s)
            got = from_device(td, a.shape, dst_t)
            want = O.separable_2d(name, a, dst_t, dim, N, s)
            assert_parity(got, want, "%s dim=%d N=%d" % (name, dim, N))


@pytest.mark.parametrize("src_t", DTYPES, ids=lambda t: DTYPE_NAMES[t])
@pytest.mark.parametrize("dst_t", DTYPES, ids=lambda t: DTYPE_NAMES[t])
def test_2d_median_all_type_pairs(gpu, src_t, dst_t):
    """clk_median_box_2d (C-cast narrowing of integers) against the oracle,
    through the packed 3x3 path, the histogram path and the generic one."""
    import clkmi
    a = random_volume((29, 77), src_t, seed=41)
    ts, si = _image_2d(a)
    for Nx, Ny in ((3, 3), (5, 5), (1, 7), (3, 1), (4, 2)):
        td, di = _image_2d(a.shape, dst_t)
        clkmi.median_box_2d(di, si, Nx, Ny)
        got = from_device(td, a.shape, dst_t)
        want = O.median_box_2d(a, dst_t, Nx, Ny)
        assert_parity(got, want, "median_box_2d %dx%d" % (Nx, Ny))


@pytest.mark.parametrize("src_t,dst_t", [(np.uint8, np.uint8), (np.uint16, np.uint16), (np.int16, np.int8),
                                         (np.float32, np.uint16)], ids=lambda t: DTYPE_NAMES[t])
def test_median_3x3_packed_y_runs(gpu, src_t, dst_t):
    """The packed 3x3 median of single launches slides along y in runs of
    16 rows: height 37 (three runs, the last partial), width 1030 (three
    512-wide tiles, the last partial), for median_box_2d and for each plane
    of median_slice_by_slice_box_3d."""
    import clkmi
    a = random_volume((37, 1030), src_t, seed=73)
    ts, si = _image_2d(a)
    td, di = _image_2d(a.shape, dst_t)
    clkmi.median_box_2d(di, si, 3, 3)
    assert_parity(from_device(td, a.shape, dst_t), O.median_box_2d(a, dst_t, 3, 3), "median_box_2d 3x3")
    v = random_volume((3, 37, 1030), src_t, seed=74)
    tv, vi = to_device(v)
    tw, wi = empty_device(v.shape, dst_t)
    clkmi.median_slice_by_slice_box_3d(wi, vi, 3, 3)
    assert_parity(from_device(tw, v.shape, dst_t), O.median_slice_by_slice_box_3d(v, dst_t, 3, 3),
                  "slice median 3x3")


def test_2d_by_name_chains_and_errors(gpu):
    """execute() with the _2d names and program file names, the 2D host
    chains, and the depth-1 rule."""
    import clkmi
    a = random_volume((40, 90), np.uint16, seed=3)
    ts, si = _image_2d(a)
    td1, d1 = _image_2d(a.shape, np.uint8)
    td2, d2 = _image_2d(a.shape, np.uint8)
    clkmi.median_box_2d(d1, si, 3, 3)
    clkmi.execute("median_box_2d_x.cl", {"dst": d2, "src": si, "Nx": 3, "Ny": 3})
    assert np.array_equal(from_device(td1, a.shape, np.uint8), from_device(td2, a.shape, np.uint8))
    tf, df = _image_2d(a.shape, np.float32)
    clkmi.execute("gaussian_blur_separable_2d", {"dst": df, "src": si, "dim": 1, "N": 9, "s": 1.5})
    assert_parity(from_device(tf, a.shape, np.float32),
                  O.separable_2d("gaussian_blur_separable_2d", a, np.float32, 1, 9, 1.5), "execute 2d")
    # 2D chains: x then y with float temporaries (min/max: the source type)
    tg, dg = _image_2d(a.shape, np.float32)
    clkmi.gaussian_blur_2d(si, dg, 1.5, 0.7)
    n1, n2 = O.sigma_to_n(1.5), O.sigma_to_n(0.7)
    want = O.separable_2d("gaussian_blur_separable_2d",
                          O.separable_2d("gaussian_blur_separable_2d", a, np.float32, 0, n1, 1.5), np.float32, 1, n2, 0.7)
    assert_parity(from_device(tg, a.shape, np.float32), want, "gaussian_blur_2d")
    tm, dm = _image_2d(a.shape, np.uint16)
    clkmi.maximum_box_2d(si, dm, 2, 1)
    want = O.separable_2d("maximum_separable_2d", O.separable_2d("maximum_separable_2d", a, np.uint16, 0, 5), np.uint16, 1, 3)
    assert_parity(from_device(tm, a.shape, np.uint16), want, "maximum_box_2d")
    # a 2D kernel needs depth 1
    t3, s3 = to_device(random_volume((2, 8, 8), np.uint16, seed=1))
    t4, d4 = empty_device((2, 8, 8), np.uint16)
    with pytest.raises(clkmi.ClkError) as e:
        clkmi.median_box_2d(d4, s3, 3, 3)
    assert e.value.code == -30
    with pytest.raises(clkmi.ClkError) as e:
        clkmi.execute("mean_separable_2d", {"dst": d4, "src": s3, "dim": 0, "N": 3, "s": 0.0})
    assert e.value.code == -30
    with pytest.raises(clkmi.ClkError) as e:  # median_box_2d has no Nz
        clkmi.execute("median_box_2d", {"dst": d1, "src": si, "Nx": 3, "Ny": 3, "Nz": 3})
    assert e.value.code == -52


# ---- ellipsoid neighbourhoods, slice medians, copy -----------------------------

SPHERE_OPS = ["minimum_sphere_3d", "maximum_sphere_3d", "mean_sphere_3d", "median_sphere_3d"]


@pytest.mark.parametrize("name", SPHERE_OPS)
@pytest.mark.parametrize("src_t", DTYPES, ids=lambda t: DTYPE_NAMES[t])
@pytest.mark.parametrize("dst_t", DTYPES, ids=lambda t: DTYPE_NAMES[t])
def test_sphere_all_type_pairs(gpu, name, src_t, dst_t):
    import clkmi
    a = random_volume((6, 9, 70), src_t, seed=zlib.crc32((name + DTYPE_NAMES[src_t]).encode()))
    ts, si = to_device(a)
    for Ns in ((3, 3, 3), (5, 5, 5), (5, 3, 1), (1, 1, 1), (7, 5, 3), (4, 2, 6)):
        td, di = empty_device(a.shape, dst_t)
        getattr(clkmi, name)(di, si, *Ns)
        got = from_device(td, a.shape, dst_t)
        assert_parity(got, O.sphere_3d(name, a, dst_t, *Ns), "%s %s" % (name, Ns))


@pytest.mark.parametrize("src_t", DTYPES, ids=lambda t: DTYPE_NAMES[t])
@pytest.mark.parametrize("dst_t", DTYPES, ids=lambda t: DTYPE_NAMES[t])
def test_slice_median_and_copy_all_type_pairs(gpu, src_t, dst_t):
    import clkmi
    a = random_volume((5, 11, 37), src_t, seed=77)
    ts, si = to_device(a)
    for Nx, Ny in ((3, 3), (5, 1), (1, 7)):
        td, di = empty_device(a.shape, dst_t)
        clkmi.median_slice_by_slice_box_3d(di, si, Nx, Ny)
        assert_parity(from_device(td, a.shape, dst_t), O.median_slice_by_slice_box_3d(a, dst_t, Nx, Ny),
                      "slice median %dx%d" % (Nx, Ny))
    td, di = empty_device(a.shape, dst_t)
    clkmi.copy_3d(di, si)
    assert_parity(from_device(td, a.shape, dst_t), O.copy_3d(a, dst_t), "copy_3d")


def test_sphere_by_name_and_errors(gpu):
    import clkmi
    a = random_volume((8, 16, 40), np.uint16, seed=5)
    ts, si = to_device(a)
    t1, d1 = empty_device(a.shape, np.uint16)
    t2, d2 = empty_device(a.shape, np.uint16)
    clkmi.median_sphere_3d(d1, si, 5, 5, 3)
    clkmi.execute("median_sphere_3d_x.cl", {"dst": d2, "src": si, "Nx": 5, "Ny": 5, "Nz": 3})
    assert np.array_equal(from_device(t1, a.shape, np.uint16), from_device(t2, a.shape, np.uint16))
    clkmi.execute("copy_3d", {"dst": d2, "src": si})
    assert np.array_equal(from_device(t2, a.shape, np.uint16), a)
    clkmi.execute("median_slice_by_slice_box_3d", {"dst": d2, "src": si, "Nx": 3, "Ny": 3})
    assert np.array_equal(from_device(t2, a.shape, np.uint16), O.median_slice_by_slice_box_3d(a, np.uint16, 3, 3))
    # degenerate sizes as the reference: empty neighbourhood -> centre / 0/0
    clkmi.minimum_sphere_3d(d1, si, -1, -1, -1)
    assert np.array_equal(from_device(t1, a.shape, np.uint16), a)
    for bad in ((-1, 3, 3), (65, 3, 3)):
        with pytest.raises(clkmi.ClkError) as e:
            clkmi.median_sphere_3d(d1, si, *bad)
        assert e.value.code == -30
    with pytest.raises(clkmi.ClkError) as e:
        clkmi.execute("mean_sphere_3d", {"dst": d1, "src": si, "Nx": 3, "Ny": 3})
    assert e.value.code == -52


def test_copy_3d_unaligned_and_ragged(gpu):
    """The vector copy needs 4-pixel alignment; offset pointers take the
    element path, and a ragged tail is finished element by element (also
    after several blocks of the vector path)."""
    import clkmi
    import torch
    for st, dt in ((np.uint8, np.float32), (np.float32, np.int16), (np.uint16, np.uint8)):
        b = random_volume((3, 67, 129), st, seed=9)  # 25929 pixels: 7 blocks, 1 left over
        tb, bi = to_device(b)
        td, di = empty_device(b.shape, dt)
        clkmi.copy_3d(di, bi)
        assert_parity(from_device(td, b.shape, dt), O.copy_3d(b, dt), "copy %s->%s" % (DTYPE_NAMES[st], DTYPE_NAMES[dt]))
    a = random_volume((3, 7, 13), np.int16, seed=4)
    n = a.size
    src = torch.from_numpy(np.concatenate([np.zeros(1, np.int16), a.reshape(-1)])).cuda()
    dst = torch.zeros(n + 1, dtype=torch.float32, device="cuda")
    si = clkmi.Image(src.data_ptr() + 2, 2, 13, 7, 3)        # one pixel in: unaligned
    di = clkmi.Image(dst.data_ptr() + 4, 4, 13, 7, 3)
    clkmi.copy_3d(di, si)
    got = dst.cpu().numpy()[1:].reshape(a.shape)
    assert_parity(got, O.copy_3d(a, np.float32), "unaligned copy")
    assert dst.cpu().numpy()[0] == 0


@pytest.mark.parametrize("direct", [0, 1, 2])
def test_sphere_tiled_and_direct_paths(gpu, direct):
    """The ellipsoid kernels (0: z-sliding LDS ring where the neighbourhood
    fits it, else LDS row tiles; 1: per-voxel gathers; 2: LDS row tiles)
    against the oracle, including a neighbourhood too large for the tiles
    (r = 7)."""
    import clkmi
    a = random_volume((9, 12, 300), np.int16, seed=12 + direct)
    ts, si = to_device(a)
    old = clkmi.get_tuning("sphere_direct")
    try:
        clkmi.set_tuning("sphere_direct", direct)
        for name in SPHERE_OPS:
            for Ns in ((3, 3, 3), (7, 7, 7), (15, 15, 15), (9, 1, 3)):
                if name == "median_sphere_3d" and Ns == (15, 15, 15):
                    continue  # a 1400-value bisection per voxel: slow, covered by (7, 7, 7)
                td, di = empty_device(a.shape, np.float32)
                getattr(clkmi, name)(di, si, *Ns)
                assert_parity(from_device(td, a.shape, np.float32), O.sphere_3d(name, a, np.float32, *Ns),
                              "%s %s direct=%d" % (name, Ns, direct))
    finally:
        clkmi.set_tuning("sphere_direct", old)


@pytest.mark.parametrize("src_t,dst_t", [(np.uint8, np.uint8), (np.float32, np.float32), (np.int16, np.uint16),
                                         (np.uint16, np.float32), (np.int16, np.int8), (np.float32, np.uint8)],
                         ids=lambda t: DTYPE_NAMES[t])
def test_sphere_sliding_ring(gpu, src_t, dst_t):
    """The z-sliding forms over volumes whose blocks each slide across several
    planes (depth 37: ring wrap-around, a partial last run; two z segments of
    the byte-histogram median for 8-bit dst), with several x tiles (width
    300) and y tiles (height 21, not a multiple of the rows per block), for
    balls, ellipsoids and flat shapes up to 9 x 7 x 7."""
    import clkmi
    a = random_volume((37, 21, 300), src_t, seed=31)
    if src_t == np.float32:
        a.reshape(-1)[::97] = -0.0
        a.reshape(-1)[5::211] = np.nan
    ts, si = to_device(a)
    for name in SPHERE_OPS:
        for Ns in ((3, 3, 3), (5, 5, 5), (9, 7, 3), (3, 5, 7), (1, 3, 1), (7, 1, 5)):
            td, di = empty_device(a.shape, dst_t)
            getattr(clkmi, name)(di, si, *Ns)
            assert_parity(from_device(td, a.shape, dst_t), O.sphere_3d(name, a, dst_t, *Ns),
                          "%s %s %s->%s" % (name, Ns, DTYPE_NAMES[src_t], DTYPE_NAMES[dst_t]))


@pytest.mark.parametrize("pack", [0, 1])
@pytest.mark.parametrize("shape", [(3, 5, 7), (2, 9, 37), (4, 3, 144), (2, 5, 250), (1, 1, 3), (6, 7, 256)])
def test_packed_short_rows_x_pass(gpu, shape, pack):
    """Rows of at most 256 pixels several per wave (padded rows in LDS,
    halo positions dropped) against the oracle, and against the one-row-per-
    wave pass; half widths up to beyond the row length."""
    import clkmi
    old = clkmi.get_tuning("no_xpack")
    try:
        clkmi.set_tuning("no_xpack", 1 - pack)
        for src_t, dst_t in ((np.uint16, np.float32), (np.float32, np.float32), (np.int8, np.int8), (np.uint16, np.uint16)):
            a = random_volume(shape, src_t, seed=shape[2] + pack)
            for name in SEP_OPS:
                for c in (1, 3, 5, 16, 40):
                    got = gpu_separable(name, a, dst_t, 0, 2 * c + 1, 0.3 * c + 0.5)
                    want = O.separable(name, a, dst_t, 0, 2 * c + 1, 0.3 * c + 0.5)
                    assert_parity(got, want, "%s %s c=%d pack=%d" % (name, shape, c, pack))
    finally:
        clkmi.set_tuning("no_xpack", old)


# ---- consumers ------------------------------------------------------------------

@pytest.mark.parametrize("src_t", DTYPES, ids=lambda t: DTYPE_NAMES[t])
@pytest.mark.parametrize("dst_t", DTYPES, ids=lambda t: DTYPE_NAMES[t])
def test_detect_maxima_and_projection_all_type_pairs(gpu, src_t, dst_t):
    import clkmi
    a = random_volume((6, 17, 300), src_t, seed=31)
    if np.dtype(src_t).kind != "f":
        a = (a % 5).astype(src_t)  # plateaus and ties
    else:
        a.reshape(-1)[::97] = np.nan
    ts, si = to_device(a)
    td, di = empty_device(a.shape, dst_t)
    clkmi.detect_maxima_3d(si, di)
    assert_parity(from_device(td, a.shape, dst_t), O.detect_maxima_3d(a, dst_t), "detect_maxima_3d")
    tp, dp = empty_device((1,) + a.shape[1:], dst_t)
    clkmi.maximum_z_projection(dp, si)
    assert_parity(from_device(tp, (1,) + a.shape[1:], dst_t), O.maximum_z_projection(a, dst_t),
                  "maximum_z_projection")


@pytest.mark.parametrize("src_t", [np.int8, np.uint8, np.int16, np.uint16], ids=lambda t: DTYPE_NAMES[t])
def test_detect_maxima_integer_sliding(gpu, src_t):
    """Integer sources take the z-sliding kernel: depth 70 spans three z runs
    (the last partial), width 130 three 64-lane tiles with halo lanes, height
    9; plateau data (ties everywhere) and full-range data (the type's
    extremes, where centre - 1 leaves the type's range)."""
    import clkmi
    for data in ("plateau", "full"):
        a = random_volume((70, 9, 130), src_t, seed=41)
        if data == "plateau":
            a = (a % 3).astype(src_t)
        ts, si = to_device(a)
        for dst_t in (np.uint8, np.float32):
            td, di = empty_device(a.shape, dst_t)
            clkmi.detect_maxima_3d(si, di)
            assert_parity(from_device(td, a.shape, dst_t), O.detect_maxima_3d(a, dst_t),
                          "detect_maxima_3d %s %s" % (DTYPE_NAMES[src_t], data))


@pytest.mark.parametrize("shape", [(37, 5, 131), (1, 3, 260), (3, 4, 64), (130, 2, 300)])
def test_maximum_z_projection_depth_split(gpu, shape):
    """The depth is split over four waves: depths below and above 4, widths
    that are not a multiple of 4 (element loads) and a partial last
    256-pixel block.  Float data carries NaN in the first plane and at the
    start of a quarter, and columns of zeros that begin with -0 (the
    reference keeps the first of equal values)."""
    import clkmi
    for src_t in (np.int8, np.uint8, np.int16, np.uint16, np.float32):
        a = random_volume(shape, src_t, seed=shape[0] + 3)
        if src_t == np.float32:
            D = shape[0]
            a[0, :, ::7] = np.nan
            a[min((D + 3) // 4, D - 1), :, 3::11] = np.nan
            a[:, :, 5::13] = 0.0
            a[0, :, 5::13] = -0.0
            a[:, 1:, 6::17] = -5.0
            a[-1, 1:, 6::17] = np.nan
        ts, si = to_device(a)
        for dst_t in (np.uint8, np.int16, np.float32):
            tp, dp = empty_device((1,) + a.shape[1:], dst_t)
            clkmi.maximum_z_projection(dp, si)
            got, want = from_device(tp, (1,) + a.shape[1:], dst_t), O.maximum_z_projection(a, dst_t)
            assert_parity(got, want, "projection %s %s->%s" % (shape, DTYPE_NAMES[src_t], DTYPE_NAMES[dst_t]))
            if dst_t == np.float32:
                # the chain keeps the first of equal values: -0 and +0 differ in bits
                assert np.array_equal(np.signbit(got), np.signbit(want)), "projection %s signs" % (shape,)


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 1, 5), (3, 1, 1), (1, 4, 1), (2, 2, 2), (1, 2, 600)])
def test_neighbourhood_kernels_degenerate_shapes(gpu, shape):
    """Single pixels, rows, columns and planes through the sliding kernels
    (packed medians, ellipsoid filters, detect_maxima, projection), where
    every clamp collapses onto the volume's edges."""
    import clkmi
    a = random_volume(shape, np.uint8, seed=sum(shape))
    ts, si = to_device(a)
    got = gpu_median(a, np.uint8, 3, 3, 3)
    assert_parity(got, O.median_box_3d(a, np.uint8, 3, 3, 3), "median 3x3x3 %s" % (shape,))
    td, di = empty_device(a.shape, np.uint8)
    clkmi.median_slice_by_slice_box_3d(di, si, 3, 3)
    assert_parity(from_device(td, a.shape, np.uint8), O.median_slice_by_slice_box_3d(a, np.uint8, 3, 3),
                  "slice median %s" % (shape,))
    for name in SPHERE_OPS:
        for Ns in ((3, 3, 3), (5, 5, 5)):
            td, di = empty_device(a.shape, np.uint8)
            getattr(clkmi, name)(di, si, *Ns)
            assert_parity(from_device(td, a.shape, np.uint8), O.sphere_3d(name, a, np.uint8, *Ns),
                          "%s %s %s" % (name, Ns, shape))
    td, di = empty_device(a.shape, np.uint8)
    clkmi.detect_maxima_3d(si, di)
    assert_parity(from_device(td, a.shape, np.uint8), O.detect_maxima_3d(a, np.uint8), "detect_maxima %s" % (shape,))
    tp, dp = empty_device((1,) + a.shape[1:], np.uint8)
    clkmi.maximum_z_projection(dp, si)
    assert_parity(from_device(tp, (1,) + a.shape[1:], np.uint8), O.maximum_z_projection(a, np.uint8),
                  "projection %s" % (shape,))


def test_consumers_by_name_and_errors(gpu):
    import clkmi
    a = random_volume((4, 9, 33), np.uint16, seed=2)
    ts, si = to_device(a)
    td, di = empty_device(a.shape, np.uint8)
    clkmi.execute("detect_maxima_3d_x.cl", {"src": si, "dst": di})
    assert np.array_equal(from_device(td, a.shape, np.uint8), O.detect_maxima_3d(a, np.uint8))
    tp, dp = empty_device((1, 9, 33), np.float32)
    clkmi.execute("maximum_z_projection", {"dst_max": dp, "src": si})
    assert_parity(from_device(tp, (1, 9, 33), np.float32), O.maximum_z_projection(a, np.float32), "by name")
    bad_t, bad = empty_device((2, 9, 33), np.float32)   # depth must be 1
    with pytest.raises(clkmi.ClkError) as e:
        clkmi.maximum_z_projection(bad, si)
    assert e.value.code == -30
    with pytest.raises(clkmi.ClkError) as e:   # projection output is named dst_max
        clkmi.execute("maximum_z_projection", {"dst": dp, "src": si})
    assert e.value.code == -52


RING_SETTINGS = [
    (("yring", 2), ("ystatic_cols", 1)),
    (("yring", 2), ("ystatic_cols", 2), ("strided_order", 1)),
    (("yring", 2), ("ystatic_cols", 4), ("strided_rounds", 64), ("strided_order", 3)),
    (("yring", 2), ("ystatic_cols", 0), ("strided_rounds", 16), ("strided_order", 2)),
    (("yring", 2), ("yring_deep", 1), ("ystatic_cols", 1), ("strided_order", 1)),
    (("yring", 2), ("yring_deep", 1), ("ystatic_cols", 2)),
    (("yring", 2), ("yring_deep", 1), ("ystatic_cols", 4), ("strided_rounds", 64), ("strided_order", 3)),
]


@pytest.mark.parametrize("shape", [(37, 41, 8), (20, 300, 12), (70, 9, 6), (6, 70, 1030), (130, 33, 40)])
def test_lockstep_ring_pass(gpu, shape):
    """The float32 y/z pass in 1024-thread lockstep blocks: every op, half
    widths up to its limit, one, two or four columns per lane, blocks with
    idle lanes (fewer columns than 1024 per block), several segments per
    axis and every block order, along y and z; bit for bit as the oracle."""
    import clkmi
    src = random_volume(shape, np.float32, seed=shape[0] * 7 + shape[2])
    src.flat[::131] = -0.0
    for name in SEP_OPS:
        cmax = 12 if name.startswith("gaussian") else 8
        for c in (1, 2, 3, 6, 7, 8, 12):
            if c > cmax:
                continue
            N = 2 * c + 1
            for dim in (1, 2):
                want = O.separable(name, src, np.float32, dim, N, 0.4 * c + 0.3)
                for settings in RING_SETTINGS:
                    old = {k: clkmi.get_tuning(k) for k, _ in settings}
                    try:
                        for k, v in settings:
                            clkmi.set_tuning(k, v)
                        got = gpu_separable(name, src, np.float32, dim, N, 0.4 * c + 0.3)
                    finally:
                        for k, v in old.items():
                            clkmi.set_tuning(k, v)
                    assert_parity(got, want, "%s %s dim=%d c=%d %s" % (name, shape, dim, c, settings))
