"""Shared test helpers: seeded volumes and device round trips through the
C-ABI's raw Image descriptors."""
import numpy as np

DTYPES = [np.int8, np.uint8, np.int16, np.uint16, np.float32]
DTYPE_NAMES = {np.int8: "i8", np.uint8: "u8", np.int16: "i16", np.uint16: "u16", np.float32: "f32"}
CODES = {np.dtype(np.int8): 0, np.dtype(np.uint8): 1, np.dtype(np.int16): 2,
         np.dtype(np.uint16): 3, np.dtype(np.float32): 4}


def random_volume(shape, dtype, seed, lo=None, hi=None):
    """Values spanning the type's range (ints) or [-1000, 1000) (float), so
    that conversions to narrower types saturate somewhere."""
    rng = np.random.default_rng(seed)
    dt = np.dtype(dtype)
    if dt.kind == "f":
        lo = -1000.0 if lo is None else lo
        hi = 1000.0 if hi is None else hi
        return rng.uniform(lo, hi, size=shape).astype(np.float32)
    info = np.iinfo(dt)
    lo = info.min if lo is None else lo
    hi = info.max if hi is None else hi
    return rng.integers(lo, hi, size=shape, endpoint=True).astype(dt)


def to_device(a):
    """Upload a (z, y, x) numpy volume; returns (keepalive tensor, Image)."""
    import torch
    import clkmi
    a = np.ascontiguousarray(a)
    t = torch.from_numpy(a.view(np.uint8).reshape(-1).copy()).cuda()
    d, h, w = a.shape
    return t, clkmi.Image(t.data_ptr(), CODES[a.dtype], w, h, d)


def empty_device(shape, dtype):
    import torch
    import clkmi
    dt = np.dtype(dtype)
    n = int(np.prod(shape)) * dt.itemsize
    t = torch.empty(max(n, 1), dtype=torch.uint8, device="cuda")
    d, h, w = shape
    return t, clkmi.Image(t.data_ptr(), CODES[dt], w, h, d)


def from_device(t, shape, dtype):
    import torch
    torch.cuda.synchronize()
    dt = np.dtype(dtype)
    n = int(np.prod(shape)) * dt.itemsize
    return t[:n].cpu().numpy().view(dt).reshape(shape)


def same_bits_or_value(a, b):
    """Equality for parity: integers exactly; floats by value with NaN == NaN
    (so +0 == -0)."""
    if a.dtype.kind == "f":
        return np.array_equal(a, b, equal_nan=True)
    return np.array_equal(a, b)


def nanzero_volume(shape, kind, seed):
    """Float volumes for the insertion-sort median rules on NaN and -0:
    'ties' draws from {-0, +0, -1, 1, 2, NaN} (zeros at the median position,
    windows split by NaNs); 'sparse' is uniform with a few -0 and NaN (most
    8x8x8 bricks hold neither); 'slab' has NaN over a z slab (windows whose
    middle member is NaN) and -0 / +0 runs elsewhere."""
    rng = np.random.default_rng(seed)
    if kind == "ties":
        vals = np.array([-0.0, 0.0, -1.0, 1.0, 2.0, np.nan], dtype=np.float32)
        return vals[rng.choice(6, size=shape, p=[0.25, 0.2, 0.15, 0.15, 0.1, 0.15])]
    a = rng.uniform(-1000, 1000, size=shape).astype(np.float32)
    flat = a.reshape(-1)
    if kind == "sparse":
        flat[rng.choice(flat.size, max(1, flat.size // 900), replace=False)] = np.nan
        flat[rng.choice(flat.size, max(1, flat.size // 700), replace=False)] = -0.0
        return a
    z = shape[0]
    a[z // 3: z // 3 + 3] = np.nan
    zeros = rng.random(shape) < 0.4
    a[zeros & (a == a)] = np.where(rng.random(int((zeros & (a == a)).sum())) < 0.5, -0.0, 0.0)
    return a


def same_float_bits(a, b):
    """Floats equal bit for bit (sign of zero included), NaN == NaN."""
    an, bn = np.isnan(a), np.isnan(b)
    return bool(np.array_equal(an, bn) and np.array_equal(a.view(np.uint32)[~an], b.view(np.uint32)[~bn]))
