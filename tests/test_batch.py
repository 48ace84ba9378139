"""Batched execution (clk_batch_filter_3d): every job's result equals the
single call's and the oracle chain's, bit for bit, whatever the mix of
volume sizes, pixel types, ops and tuning."""
import numpy as np
import pytest

import clk_oracle as O
from helpers import empty_device, from_device, random_volume, same_bits_or_value, to_device

pytestmark = pytest.mark.gpu

SINGLE = {"gaussian_blur_3d": None, "mean_box_3d": O.OP_MEAN, "minimum_box_3d": O.OP_MIN,
          "maximum_box_3d": O.OP_MAX}


def oracle_job(kind, src, dst_t, sigma, radius):
    if kind == "gaussian_blur_3d":
        return O.gaussian_blur_3d(src, dst_t, *sigma)
    if kind == "median_box_3d":
        return O.median_box_3d(src, dst_t, *(2 * r + 1 for r in radius))
    return O.box_3d(SINGLE[kind], src, dst_t, *radius)


def make_jobs(rng, n, kinds, pairs, lo=5, hi=40):
    jobs = []
    for i in range(n):
        shape = tuple(int(v) for v in rng.integers(lo, hi + 1, 3))
        kind = kinds[i % len(kinds)]
        src_t, dst_t = pairs[rng.integers(len(pairs))]
        if kind == "median_box_3d" and src_t == np.float32 and dst_t != np.float32:
            dst_t = np.float32
        src = random_volume(shape, src_t, seed=1000 + i, lo=0 if np.dtype(src_t).kind == "u" else None,
                            hi=3000 if np.dtype(src_t).itemsize >= 2 and np.dtype(src_t).kind != "f" else None)
        sigma = tuple(float(rng.choice([0.0, rng.uniform(0.5, 4.0)], p=[0.1, 0.9])) for _ in range(3))
        radius = tuple(int(rng.integers(0 if kind != "median_box_3d" else 1, 6 if kind != "median_box_3d" else 3))
                       for _ in range(3))
        jobs.append((kind, src, dst_t, sigma, radius))
    return jobs


def run_batch(jobs):
    import clkmi
    keep, descs, outs = [], [], []
    for kind, src, dst_t, sigma, radius in jobs:
        ts, si = to_device(src)
        td, di = empty_device(src.shape, dst_t)
        keep += [ts, td]
        outs.append((td, src.shape, dst_t))
        descs.append(clkmi.FilterJob(kind, di, si, sigma, radius))
    clkmi.batch_filter_3d(descs)
    return [from_device(*o) for o in outs]


def run_single(job):
    import clkmi
    kind, src, dst_t, sigma, radius = job
    ts, si = to_device(src)
    td, di = empty_device(src.shape, dst_t)
    if kind == "gaussian_blur_3d":
        clkmi.gaussian_blur_3d(si, di, *sigma)
    elif kind == "median_box_3d":
        clkmi.median_box_3d(di, si, *(2 * r + 1 for r in radius))
    else:
        getattr(clkmi, kind)(si, di, *radius)
    return from_device(td, src.shape, dst_t)


PAIRS = [(np.uint16, np.uint16), (np.uint16, np.float32), (np.float32, np.float32), (np.uint8, np.uint8),
         (np.int16, np.float32), (np.float32, np.uint16), (np.int8, np.int8), (np.uint16, np.uint8)]


def test_batch_packed_medians_slide(gpu):
    """Batched 3x3x1 and 3x3x3 medians (the packed networks, sliding along y
    and z in runs of 16) on shapes spanning several runs and 512-wide tiles,
    equal to single calls and the oracle."""
    jobs = []
    for i, (shape, radius, st, dt) in enumerate([
            ((3, 37, 530), (1, 1, 0), np.uint8, np.uint8), ((2, 16, 20), (1, 1, 0), np.uint16, np.uint16),
            ((37, 5, 600), (1, 1, 1), np.uint8, np.uint8), ((17, 19, 33), (1, 1, 1), np.int16, np.int16),
            ((40, 3, 9), (1, 1, 1), np.float32, np.float32), ((5, 33, 40), (1, 1, 0), np.uint16, np.uint8)]):
        src = random_volume(shape, st, seed=500 + i)
        jobs.append(("median_box_3d", src, dt, (0.0, 0.0, 0.0), radius))
    got = run_batch(jobs)
    for i, (job, g) in enumerate(zip(jobs, got)):
        assert same_bits_or_value(g, run_single(job)), "job %d single" % i
        assert same_bits_or_value(g, oracle_job(*job)), "job %d oracle" % i


def test_batch_median_register_tiers(gpu):
    """Batched medians on every register tier (3x3x1 .. 5x5x5 boxes, 16- and
    32-bit dst) and on a sparse box that keeps the generic search."""
    jobs = []
    radii = [(1, 1, 0), (2, 2, 0), (0, 1, 1), (3, 3, 0), (2, 2, 1), (2, 2, 2), (2, 1, 2), (0, 0, 2)]
    for i, radius in enumerate(radii):
        st, dt = [(np.uint16, np.uint16), (np.float32, np.float32), (np.int16, np.float32)][i % 3]
        src = random_volume((9 + i, 11, 37 + 5 * i), st, seed=600 + i)
        jobs.append(("median_box_3d", src, dt, (0.0, 0.0, 0.0), radius))
    got = run_batch(jobs)
    for i, (job, g) in enumerate(zip(jobs, got)):
        assert same_bits_or_value(g, run_single(job)), "job %d single" % i
        assert same_bits_or_value(g, oracle_job(*job)), "job %d oracle" % i


@pytest.mark.parametrize("seed", [1, 2])
def test_batch_matches_single_calls_and_oracle(gpu, seed):
    rng = np.random.default_rng(seed)
    jobs = make_jobs(rng, 40, ["gaussian_blur_3d", "minimum_box_3d", "maximum_box_3d", "mean_box_3d",
                               "median_box_3d"], PAIRS)
    got = run_batch(jobs)
    for i, (job, g) in enumerate(zip(jobs, got)):
        want = oracle_job(job[0], job[1], job[2], job[3], job[4])
        assert same_bits_or_value(g, want), "job %d %s %s" % (i, job[0], job[1].shape)
        assert same_bits_or_value(g, run_single(job)), "job %d single call" % i


@pytest.mark.parametrize("seg", [8, 32, 256])
def test_batch_tuning_does_not_change_results(gpu, seg):
    import clkmi
    rng = np.random.default_rng(7)
    jobs = make_jobs(rng, 12, ["gaussian_blur_3d", "maximum_box_3d"], PAIRS[:3], lo=20, hi=70)
    old = clkmi.get_tuning("batch_seg")
    try:
        clkmi.set_tuning("batch_seg", seg)
        got = run_batch(jobs)
    finally:
        clkmi.set_tuning("batch_seg", old)
    for job, g in zip(jobs, got):
        assert same_bits_or_value(g, oracle_job(*job))


def test_batch_c4_shaped_workload(gpu):
    """The uneven-batch workload at reduced count: uint16 volumes with axes
    in [32, 256), Gaussian sigma in [0.5, 4] -> float32, min/max r 1..5."""
    rng = np.random.default_rng(4)
    jobs = []
    for i in range(12):
        shape = tuple(int(v) for v in rng.integers(32, 120, 3))
        src = rng.poisson(200, size=shape).astype(np.uint16)
        k = ["gaussian_blur_3d", "minimum_box_3d", "maximum_box_3d"][i % 3]
        sigma = tuple(float(v) for v in rng.uniform(0.5, 4.0, 3))
        r = int(rng.integers(1, 6))
        jobs.append((k, src, np.float32 if k.startswith("g") else np.uint16, sigma, (r, r, r)))
    got = run_batch(jobs)
    for job, g in zip(jobs, got):
        assert same_bits_or_value(g, oracle_job(*job))


def test_batch_errors(gpu):
    import clkmi
    src = random_volume((4, 5, 6), np.uint8, seed=0)
    ts, si = to_device(src)
    td, di = empty_device((4, 5, 6), np.uint8)
    with pytest.raises(clkmi.ClkError) as e:
        clkmi.batch_filter_3d([clkmi.FilterJob("maximum_box_3d", si, si, radius=(1, 1, 1))])
    assert e.value.code == -8
    with pytest.raises(clkmi.ClkError) as e:
        clkmi.batch_filter_3d([clkmi.FilterJob("minimum_box_3d", di, si, radius=(1, -1, 1))])
    assert e.value.code == -30
    clkmi.batch_filter_3d([])


def test_batch_yz_tier_classes(gpu):
    """The y/z tier classes (a run-time half width up to 5, 8 or 16, one
    launch per tier) give the generic strided classes' bits and the
    oracle's: half widths 1..18 (beyond the last tier), odd, even and
    multiple-of-four inner extents (1, 2 or 4 columns per lane), several
    segments along an axis."""
    import clkmi
    rng = np.random.default_rng(21)
    jobs = []
    pairs = [(np.uint16, np.float32), (np.float32, np.float32), (np.uint8, np.uint8), (np.float32, np.uint16),
             (np.int16, np.int16), (np.uint16, np.uint16)]
    for i in range(40):
        shape = (int(rng.integers(3, 90)), int(rng.integers(3, 50)), int(rng.integers(3, 70)))
        kind = ["gaussian_blur_3d", "minimum_box_3d", "maximum_box_3d", "mean_box_3d"][i % 4]
        st, dt = pairs[i % len(pairs)]
        src = random_volume(shape, st, seed=2100 + i, lo=0 if np.dtype(st).kind == "u" else None)
        sigma = tuple(float(v) for v in rng.uniform(0.3, 8.6, 3))
        radius = tuple(int(v) for v in rng.integers(1, 19, 3))
        jobs.append((kind, src, dt, sigma, radius))
    got = run_batch(jobs)
    old = clkmi.get_tuning("strided_variant")
    try:
        clkmi.set_tuning("strided_variant", 1)
        generic = run_batch(jobs)
    finally:
        clkmi.set_tuning("strided_variant", old)
    for i, (job, g, h) in enumerate(zip(jobs, got, generic)):
        assert np.array_equal(g.view(np.uint8), h.view(np.uint8)), "job %d tiers != generic" % i
        assert same_bits_or_value(g, oracle_job(*job)), "job %d %s %s" % (i, job[0], job[1].shape)


def _offset_device(a, offset):
    """Upload a (z, y, x) volume `offset` bytes into a fresh buffer (rows
    that do not start 16-byte aligned); returns (keepalive, Image)."""
    import torch
    import clkmi
    from helpers import CODES
    raw = np.ascontiguousarray(a).view(np.uint8).reshape(-1)
    t = torch.zeros(raw.size + offset + 64, dtype=torch.uint8, device="cuda")
    t[offset:offset + raw.size] = torch.from_numpy(raw.copy()).cuda()
    d, h, w = a.shape
    return t, clkmi.Image(t.data_ptr() + offset, CODES[a.dtype], w, h, d)


def test_batch_x_tier_classes(gpu):
    """The x tier classes (register windows over per-row 8-pixel chunks, a
    run-time half width up to 4, 8 or 16) against the oracle and against the
    LDS row classes they replace: row lengths 1..1000 (partial last chunks,
    rows shorter than a chunk), half widths 0..17 (the last beyond the
    tiers), x-only jobs (the pass writes the job's dst) and full chains,
    aligned and unaligned rows."""
    import clkmi
    rng = np.random.default_rng(31)
    widths = [1, 3, 7, 8, 9, 15, 16, 17, 31, 33, 64, 100, 255, 256, 257, 300, 513, 1000]
    pairs = [(np.uint16, np.float32), (np.uint16, np.uint16), (np.uint8, np.uint8), (np.float32, np.float32),
             (np.int16, np.float32), (np.int8, np.int8)]
    jobs, offsets = [], []
    for i, W in enumerate(widths * 2):
        kind = ["gaussian_blur_3d", "minimum_box_3d", "maximum_box_3d", "mean_box_3d"][i % 4]
        st, dt = pairs[i % len(pairs)]
        shape = (int(rng.integers(1, 6)), int(rng.integers(1, 9)), W)
        src = random_volume(shape, st, seed=3100 + i, lo=0 if np.dtype(st).kind == "u" else None)
        c = i % 18
        if kind == "gaussian_blur_3d":
            s = {0: 0.1}.get(c, (2 * c + 0.5) / 8.0)  # int(8 s) = 2c: N = 2c + 1
            sigma = (s, 0.0, 0.0) if i < len(widths) else (s, 1.0, 0.6)
            radius = (0, 0, 0)
        else:
            sigma = (0.0, 0.0, 0.0)
            radius = (max(c, 1), 0, 0) if i < len(widths) else (max(c, 1), 1, 2)
        jobs.append((kind, src, dt, sigma, radius))
        offsets.append(0 if i % 3 else (np.dtype(st).itemsize * 3))

    def run(no_xtier):
        keep, descs, outs = [], [], []
        for (kind, src, dst_t, sigma, radius), off in zip(jobs, offsets):
            ts, si = _offset_device(src, off)
            td, di = empty_device(src.shape, dst_t)
            keep += [ts, td]
            outs.append((td, src.shape, dst_t))
            descs.append(clkmi.FilterJob(kind, di, si, sigma, radius))
        old = clkmi.get_tuning("no_xtier")
        try:
            clkmi.set_tuning("no_xtier", no_xtier)
            clkmi.batch_filter_3d(descs)
        finally:
            clkmi.set_tuning("no_xtier", old)
        return [from_device(*o) for o in outs]

    print("x tier batch", len(jobs), flush=True)
    got = run(0)
    rows = run(1)
    for i, (job, g, h) in enumerate(zip(jobs, got, rows)):
        ctx = "job %d %s %s %s sigma=%s radius=%s" % (i, job[0], job[1].dtype, job[1].shape, job[3], job[4])
        assert np.array_equal(g.view(np.uint8), h.view(np.uint8)), ctx + ": tiers != LDS rows"
        assert same_bits_or_value(g, oracle_job(*job)), ctx + ": != oracle"
