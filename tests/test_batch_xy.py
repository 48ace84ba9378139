"""Fused x+y passes of the batch engine: a 3D filter whose x and y half
widths are at most 16 runs its x pass inside the y pass (no x temporary):
the LDS-staged form (any op; off by default), the register form of
minimum/maximum of 16-bit integers with half widths up to 8 (8-pixel chunks
per lane, packed 16-bit min/max) and the register form of the Gaussian and
the mean (x by DPP chunks, rows handed to the y ring through LDS); and the
whole 3D minimum/maximum of 16-bit integers in one pass (x and y through
LDS, z in a register ring).
Every such job must equal the same batch with fusion off (separate x and y
passes, already pinned against single calls) and the oracle chain, bit for
bit, over row lengths that exercise every column width (1, 2, 4 columns per
thread, rows shorter and longer than a block, several planes per block),
heights below one step and over several segments, half widths 0..16 on
each axis, and every pixel type the fused classes take."""
import numpy as np
import pytest

import clk_oracle as O
from helpers import empty_device, from_device, random_volume, same_bits_or_value, to_device

pytestmark = pytest.mark.gpu


def oracle_job(kind, src, dst_t, sigma, radius):
    if kind == "gaussian_blur_3d":
        return O.gaussian_blur_3d(src, dst_t, *sigma)
    op = {"mean_box_3d": O.OP_MEAN, "minimum_box_3d": O.OP_MIN, "maximum_box_3d": O.OP_MAX}[kind]
    return O.box_3d(op, src, dst_t, *radius)


def run_batch(jobs, **knobs):
    import clkmi
    old = {k: clkmi.get_tuning(k) for k in knobs}
    keep, descs, outs = [], [], []
    for kind, src, dst_t, sigma, radius in jobs:
        ts, si = to_device(src)
        td, di = empty_device(src.shape, dst_t)
        keep += [ts, td]
        outs.append((td, src.shape, dst_t))
        descs.append(clkmi.FilterJob(kind, di, si, sigma, radius))
    try:
        for k, v in knobs.items():
            clkmi.set_tuning(k, v)
        clkmi.batch_filter_3d(descs)
    finally:
        for k, v in old.items():
            clkmi.set_tuning(k, v)
    return [from_device(*o) for o in outs]


def sigma_for_c(c):
    """A sigma whose tap count int(8 sigma) (made odd) has half width c."""
    return (2 * c + 0.5) / 8.0 if c > 0 else 0.05


# (z, y, x) shapes: rows of 1..300 pixels (1, 2, 4 columns per thread, a
# block spanning up to 64 planes or a fraction of one row), heights below
# one 4-row step and over several 64-row segments
SHAPES = [(3, 5, 1), (2, 7, 3), (4, 3, 8), (3, 37, 31), (2, 2, 33), (1, 70, 64), (2, 150, 100),
          (2, 41, 144), (1, 9, 256), (2, 17, 257), (1, 33, 300), (40, 6, 5), (9, 1, 40), (5, 66, 13)]


def make_jobs(seed):
    rng = np.random.default_rng(seed)
    kinds = ["gaussian_blur_3d", "mean_box_3d", "minimum_box_3d", "maximum_box_3d"]
    gm_pairs = [(np.uint16, np.float32), (np.uint8, np.uint8), (np.int16, np.float32), (np.int8, np.int16),
                (np.float32, np.float32), (np.float32, np.uint16), (np.uint16, np.uint16)]
    mm_pairs = [(np.uint16, np.uint16), (np.uint8, np.uint8), (np.int16, np.int16), (np.int8, np.int8),
                (np.uint16, np.float32), (np.int16, np.uint8)]
    jobs = []
    for i, shape in enumerate(SHAPES * 2):
        kind = kinds[(i + seed) % 4]
        st, dt = (gm_pairs if kind in ("gaussian_blur_3d", "mean_box_3d") else mm_pairs)[
            int(rng.integers(len(gm_pairs if kind in ("gaussian_blur_3d", "mean_box_3d") else mm_pairs)))]
        src = random_volume(shape, st, seed=seed * 1000 + i,
                            lo=None if np.dtype(st).kind == "f" else (0 if np.dtype(st).kind == "u" else -3000),
                            hi=None if np.dtype(st).kind == "f" else (3000 if np.dtype(st).itemsize >= 2 else None))
        cs = [int(rng.integers(0, 17)) for _ in range(2)] + [int(rng.integers(1, 6))]
        if kind == "gaussian_blur_3d":
            sigma, radius = tuple(sigma_for_c(c) for c in cs), (0, 0, 0)
        else:
            sigma, radius = (0.0, 0.0, 0.0), tuple(max(c, 1) for c in cs)
        jobs.append((kind, src, dt, sigma, radius))
    return jobs


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_xy_fused_equals_separate_and_oracle(gpu, seed):
    jobs = make_jobs(seed)
    fused = run_batch(jobs, no_xy=0, no_xymm=0, no_xyg=1, no_xyzmm=1)
    separate = run_batch(jobs, no_xy=1, no_xymm=1, no_xyg=1, no_xyzmm=1, no_zmm=1)
    reg = run_batch(jobs, no_xy=1, no_xymm=0, no_xyg=0, no_xyzmm=0)
    for i, (job, f, s) in enumerate(zip(jobs, fused, separate)):
        tag = "job %d %s %s %s->%s sigma=%s r=%s" % (i, job[0], job[1].shape, job[1].dtype, np.dtype(job[2]),
                                                  job[3], job[4])
        assert same_bits_or_value(f, s), tag + " fused != separate"
        assert same_bits_or_value(f, oracle_job(*job)), tag + " fused != oracle"
    for i, (job, g, s_) in enumerate(zip(jobs, reg, separate)):
        assert same_bits_or_value(g, s_), "job %d %s %s register fused != separate" % (i, job[0], job[1].shape)


@pytest.mark.parametrize("seg", [4, 8, 20, 1000])
def test_xy_segment_knob(gpu, seg):
    """The y segment per block (4 = one step, 20 = not a power of two, 1000 =
    whole columns) changes no bit."""
    jobs = make_jobs(4)[:14]
    want = run_batch(jobs, no_xy=1, no_xymm=1, no_xyg=1, no_xyzmm=1, no_zmm=1)
    got = run_batch(jobs, no_xy=0, no_xymm=0, no_xyg=1, no_xyzmm=1, batch_xyseg=seg)
    reg = run_batch(jobs, no_xy=1, no_xymm=0, no_xyg=0, no_xyzmm=0, batch_xyseg=seg, batch_zseg=seg)
    for i, (g, r, w) in enumerate(zip(got, reg, want)):
        assert same_bits_or_value(g, w), "job %d seg %d" % (i, seg)
        assert same_bits_or_value(r, w), "job %d seg %d register fused" % (i, seg)


def test_xy_float_specials(gpu):
    """Float sources with NaN, +-inf, -0 and subnormals through the fused
    Gaussian and mean: the same values as the separate passes."""
    rng = np.random.default_rng(11)
    jobs = []
    for i, shape in enumerate([(3, 20, 37), (2, 9, 128), (4, 33, 7)]):
        a = rng.uniform(-1000, 1000, size=shape).astype(np.float32)
        flat = a.reshape(-1)
        idx = rng.choice(flat.size, 40, replace=False)
        flat[idx[:10]] = np.nan
        flat[idx[10:20]] = np.inf
        flat[idx[20:30]] = -np.inf
        flat[idx[30:35]] = -0.0
        flat[idx[35:]] = 1e-42
        for kind in ("gaussian_blur_3d", "mean_box_3d"):
            c = (3, 9, 16)[i]
            sigma = (sigma_for_c(c), sigma_for_c(16 - c), 1.0) if kind == "gaussian_blur_3d" else (0.0,) * 3
            radius = (0, 0, 0) if kind == "gaussian_blur_3d" else (c, max(1, 16 - c), 1)
            jobs.append((kind, a, np.float32, sigma, radius))
    fused = run_batch(jobs, no_xy=0, no_xymm=0, no_xyg=1)
    separate = run_batch(jobs, no_xy=1, no_xymm=1, no_xyg=1)
    reg = run_batch(jobs, no_xy=1, no_xymm=0, no_xyg=0)
    for i, (job, f, s, r) in enumerate(zip(jobs, fused, separate, reg)):
        assert same_bits_or_value(f, s), "job %d %s" % (i, job[0])
        assert same_bits_or_value(r, s), "job %d %s register fused" % (i, job[0])
        assert same_bits_or_value(f, oracle_job(*job)), "job %d %s oracle" % (i, job[0])


@pytest.mark.parametrize("seed", [5, 6])
def test_xymm_register_class(gpu, seed):
    """Minimum/maximum of uint16/int16 with x and y radii 1..8 (the register
    class, fusing x with y or, by default, with z; the last pass in the y/z
    ring classes or in the same kernel without an x filter) and 9 (just past
    it), on rows with partial last chunks, waves
    spanning many planes, the type extremes (key flip of int16), heights
    below one step and over several segments: equal to separate passes and
    the oracle."""
    rng = np.random.default_rng(seed)
    jobs = []
    for i, shape in enumerate(SHAPES + [(3, 90, 17), (2, 130, 9), (70, 4, 2)]):
        st = (np.uint16, np.int16)[i % 2]
        info = np.iinfo(st)
        src = rng.integers(info.min, info.max, size=shape, endpoint=True).astype(st)
        flat = src.reshape(-1)
        flat[rng.choice(flat.size, max(1, flat.size // 50), replace=True)] = info.min
        flat[rng.choice(flat.size, max(1, flat.size // 50), replace=True)] = info.max
        dt = (st, np.float32, np.uint8)[i % 3]
        kind = ("minimum_box_3d", "maximum_box_3d")[(i // 2) % 2]
        radius = (int(rng.integers(1, 10)), int(rng.integers(1, 10)), int(rng.integers(1, 4)))
        jobs.append((kind, src, dt, (0.0, 0.0, 0.0), radius))
    separate = run_batch(jobs, no_xymm=1, no_xy=1, no_xyg=1, no_xyzmm=1, no_zmm=1)
    # x with y, then z; x with z, then y; each with its last pass in the
    # y/z ring classes or in the register kernel (no x filter)
    for knobs in ({"xymm_z": 0}, {"xymm_z": 1}, {"xymm_z": 0, "no_zmm": 0}, {"xymm_z": 1, "no_ymm": 0}):
        fused = run_batch(jobs, no_xymm=0, no_xyzmm=1, **knobs)
        for i, (job, f, s_) in enumerate(zip(jobs, fused, separate)):
            tag = "job %d %s %s %s->%s r=%s %s" % (i, job[0], job[1].shape, job[1].dtype, np.dtype(job[2]),
                                                  job[4], knobs)
            assert same_bits_or_value(f, s_), tag + " fused != separate"
            assert same_bits_or_value(f, oracle_job(*job)), tag + " fused != oracle"


@pytest.mark.parametrize("seed", [7, 8])
def test_xyzmm_whole_3d_class(gpu, seed):
    """Whole-3D minimum/maximum of uint16/int16 (x, y radii 1..8, z radius
    1..5; the source type out): tiles of whole rows over heights of one to
    several tiles, rows of 1..300 pixels (partial last chunks, up to 38
    chunks a row), depths of one plane to several z segments (64 and the
    knob's 7), the type extremes.  Equal to separate passes and the
    oracle; radii just past the class take the other paths and must match
    too."""
    rng = np.random.default_rng(seed)
    shapes = [(3, 5, 1), (2, 7, 3), (1, 1, 1), (70, 9, 17), (5, 130, 40), (3, 41, 300), (9, 33, 64),
              (130, 6, 9), (4, 90, 144), (2, 3, 257)]
    jobs = []
    for i, shape in enumerate(shapes):
        st = (np.uint16, np.int16)[i % 2]
        info = np.iinfo(st)
        src = rng.integers(info.min, info.max, size=shape, endpoint=True).astype(st)
        flat = src.reshape(-1)
        flat[rng.choice(flat.size, max(1, flat.size // 40), replace=True)] = info.min
        flat[rng.choice(flat.size, max(1, flat.size // 40), replace=True)] = info.max
        kind = ("minimum_box_3d", "maximum_box_3d")[(i // 2) % 2]
        radius = (int(rng.integers(1, 9)), int(rng.integers(1, 9)), int(rng.integers(1, 7)))
        jobs.append((kind, src, st, (0.0, 0.0, 0.0), radius))
    separate = run_batch(jobs, no_xymm=1, no_xy=1, no_xyg=1, no_xyzmm=1, no_zmm=1)
    for knobs in ({}, {"batch_zseg": 7}):
        whole = run_batch(jobs, no_xyzmm=0, **knobs)
        for i, (job, w, s_) in enumerate(zip(jobs, whole, separate)):
            tag = "job %d %s %s %s r=%s %s" % (i, job[0], job[1].shape, job[1].dtype, job[4], knobs)
            assert same_bits_or_value(w, s_), tag + " whole-3D != separate"
            assert same_bits_or_value(w, oracle_job(*job)), tag + " whole-3D != oracle"


def test_xymm_partial_chunks_and_lane_layout(gpu):
    """The register min/max class's partial last chunks: rows of 491 pixels
    (62 chunks, so every 61st row ends at the first owning lane of a wave
    and takes its overlap from the halo lane before it), rows of 8..15
    pixels (one whole and one partial chunk), exactly 8 pixels, and rows
    under 8 pixels (separate passes); segments shorter than the volume so
    that rows past a segment are dropped stores.  Same bits as the separate
    passes and the oracle."""
    jobs = []
    shapes = [(2, 70, 491), (3, 130, 491), (4, 9, 9), (2, 17, 15), (5, 6, 8), (3, 11, 5), (2, 64, 497)]
    for i, shape in enumerate(shapes):
        for kind in ("minimum_box_3d", "maximum_box_3d"):
            st = (np.uint16, np.int16)[i % 2]
            src = random_volume(shape, st, seed=700 + 2 * i + (kind == "maximum_box_3d"))
            radius = (1 + i % 8, 1 + (3 * i) % 8, 1 + i % 5)
            jobs.append((kind, src, st, (0.0, 0.0, 0.0), radius))
    separate = run_batch(jobs, no_xy=1, no_xymm=1, no_xyg=1, no_xyzmm=1, no_zmm=1)
    for knobs in ({"batch_xyseg": 8, "xymm_z": 0}, {"batch_xyseg": 1000, "xymm_z": 0}, {"batch_xyseg": 8, "xymm_z": 1},
                  {"batch_xyseg": 1000, "xymm_z": 1}):
        reg = run_batch(jobs, no_xy=1, no_xymm=0, no_xyg=1, no_xyzmm=1, **knobs)
        for i, (job, r, s_) in enumerate(zip(jobs, reg, separate)):
            tag = "job %d %s %s r=%s %s" % (i, job[0], job[1].shape, job[4], knobs)
            assert same_bits_or_value(r, s_), tag + " register != separate"
            assert same_bits_or_value(r, oracle_job(*job)), tag + " != oracle"
