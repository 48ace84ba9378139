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
