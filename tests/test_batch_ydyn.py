"""y/z tier classes of the batch engine (a run-time half width up to a
compiled tier, the register ring anchored on the window's newest row).

Two bodies run the Gaussian and the mean: the branch-free one for chains
from integer images (every tap of the tier, weight 0 outside the job's half
width; their float temporaries are finite) and the tap-skipping one for
float images (which may hold NaN and infinities) and for minimum/maximum.
Every job must give the same bits from both bodies, at 1, 2 and 4 columns
per lane, with y-pass rows padded to whole waves, to 16 lanes or packed, over segments
shorter than one step and several per column, and the oracle chain's
values.  Half widths cover 0..16 on y and z (tiers 5, 8
and 16, every c inside each)."""
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


def bits(a):
    return a.view(np.uint32) if a.dtype == np.float32 else a


# (z, y, x): rows of 1..132 pixels (odd widths take 1 column per lane, even
# 2, multiples of 4 up to 4), y and z from 1 row to several 16-row segments
SHAPES = [(5, 7, 1), (3, 40, 2), (37, 3, 4), (2, 1, 9), (23, 19, 6), (9, 70, 12), (70, 5, 8), (4, 33, 132),
          (18, 18, 18), (1, 50, 3)]


def make_jobs(seed, float_src=False, specials=False):
    rng = np.random.default_rng(seed)
    jobs = []
    int_types = [(np.uint16, np.float32), (np.uint8, np.uint8), (np.int16, np.float32), (np.int8, np.int16),
                 (np.uint16, np.uint16)]
    flt_types = [(np.float32, np.float32), (np.float32, np.uint16), (np.float32, np.int8)]
    for i, shape in enumerate(SHAPES * 2):
        kind = ("gaussian_blur_3d", "mean_box_3d", "minimum_box_3d", "maximum_box_3d")[(i + seed) % 4]
        if float_src:
            st, dt = flt_types[int(rng.integers(len(flt_types)))]
        else:
            st, dt = int_types[int(rng.integers(len(int_types)))]
            if kind in ("minimum_box_3d", "maximum_box_3d"):
                dt = st
        src = random_volume(shape, st, seed=seed * 1000 + i)
        if specials:
            flat = src.reshape(-1)
            idx = rng.choice(flat.size, min(flat.size, 6), replace=False)
            vals = [np.nan, np.inf, -np.inf, -0.0, 1e-42, np.nan]
            for j, k in enumerate(idx):
                flat[k] = vals[j]
        cy, cz = int(rng.integers(0, 17)), int(rng.integers(0, 17))
        if kind == "gaussian_blur_3d":
            sigma, radius = (sigma_for_c(int(rng.integers(0, 5))), sigma_for_c(cy), sigma_for_c(cz)), (0, 0, 0)
        else:
            sigma, radius = (0.0, 0.0, 0.0), (int(rng.integers(0, 3)), cy, cz)
        jobs.append((kind, src, dt, sigma, radius))
    return jobs


def check(jobs):
    legs = {}
    for skip in (0, 1):
        for ycols in (0, 1):
            legs[(skip, ycols)] = run_batch(jobs, ydyn_skip=skip, batch_ycols=ycols, batch_seg=16)
    legs["unpadded"] = run_batch(jobs, batch_ypad=0, batch_seg=16)
    legs["pad16"] = run_batch(jobs, batch_ypad=16, batch_seg=16)
    want = legs[(1, 1)]
    for i, job in enumerate(jobs):
        tag = "job %d %s %s %s->%s sigma=%s r=%s" % (i, job[0], job[1].shape, job[1].dtype, np.dtype(job[2]),
                                                  job[3], job[4])
        for leg, got in legs.items():
            assert np.array_equal(bits(got[i]), bits(want[i])), tag + " leg %s differs" % (leg,)
        assert same_bits_or_value(want[i], oracle_job(*job)), tag + " != oracle"


@pytest.mark.parametrize("seed", [21, 22])
def test_ydyn_integer_chains_branch_free_equals_skipping(gpu, seed):
    check(make_jobs(seed))


def test_ydyn_float_images_with_specials(gpu):
    """Float images (NaN, infinities, -0, subnormals) keep the tap-skipping
    body: same values as the oracle, for the Gaussian, the mean and the
    float minimum/maximum (centre first)."""
    check(make_jobs(23, float_src=True, specials=True))


def test_ydyn_weights_every_c(gpu):
    """One Gaussian and one mean per half width 0..16 on y (x and z small),
    from a uint16 image: both bodies, the oracle's values."""
    jobs = []
    for c in range(17):
        src = random_volume((6, 45, 20), np.uint16, seed=300 + c)
        jobs.append(("gaussian_blur_3d", src, np.float32, (0.05, sigma_for_c(c), sigma_for_c(16 - c)), (0, 0, 0)))
        jobs.append(("mean_box_3d", src, np.uint16, (0.0, 0.0, 0.0), (1, c, 16 - c)))
    check(jobs)
