"""numpy front end of the CPU restatement in clk_oracle.c, plus an FP64
reference of the Gaussian pass.

TEST INFRASTRUCTURE ONLY: the oracle.  Only tests/, __graft_entry__.smoke()
and the cpu_baseline leg of bench.py import this module; the product library
never calls it.

Volumes are numpy arrays of shape (depth, height, width) = (z, y, x), C
order, so the flat index is x + y*W + z*W*H as in the reference's buffers.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libclk_oracle.so")
# the same source built -O3 for x86-64-v3 (bit-identical results): what
# bench.py's cpu_baseline leg times
_FAST_PATH = os.path.join(_HERE, "_build", "libclk_oracle_fast.so")

DTYPE_CODES = {
    np.dtype(np.int8): 0,
    np.dtype(np.uint8): 1,
    np.dtype(np.int16): 2,
    np.dtype(np.uint16): 3,
    np.dtype(np.float32): 4,
}
CODE_DTYPES = {v: k for k, v in DTYPE_CODES.items()}

OP_GAUSS, OP_MEAN, OP_MIN, OP_MAX = 0, 1, 2, 3
OP_NAMES = {
    "gaussian_blur_separable_3d": OP_GAUSS,
    "mean_separable_3d": OP_MEAN,
    "minimum_separable_3d": OP_MIN,
    "maximum_separable_3d": OP_MAX,
}

_lib = None
_fast = None


def build():
    """Compile clk_oracle.c (gcc, OpenMP) into _build/."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        _lib = _load(_LIB_PATH)
    return _lib


def fast_lib():
    """The -O3 x86-64-v3 build of the same restatement (for timing)."""
    global _fast
    if _fast is None:
        _fast = _load(_FAST_PATH)
    return _fast


def _load(path):
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    i64, i32, f32, vp = ctypes.c_int64, ctypes.c_int, ctypes.c_float, ctypes.c_void_p
    L.clko_separable.argtypes = [i32, vp, i32, vp, i32, i64, i64, i64, i32, i32, f32, i32]
    L.clko_separable.restype = i32
    L.clko_median_box_3d.argtypes = [vp, i32, vp, i32, i64, i64, i64, i32, i32, i32, i32]
    L.clko_median_box_3d.restype = i32
    L.clko_sphere_3d.argtypes = [i32, vp, i32, vp, i32, i64, i64, i64, i32, i32, i32, i32]
    L.clko_sphere_3d.restype = i32
    L.clko_median_slice_by_slice_box_3d.argtypes = [vp, i32, vp, i32, i64, i64, i64, i32, i32, i32]
    L.clko_median_slice_by_slice_box_3d.restype = i32
    L.clko_copy_3d.argtypes = [vp, i32, vp, i32, i64]
    L.clko_copy_3d.restype = i32
    L.clko_detect_maxima_3d.argtypes = [vp, i32, vp, i32, i64, i64, i64]
    L.clko_detect_maxima_3d.restype = i32
    L.clko_maximum_z_projection.argtypes = [vp, i32, vp, i32, i64, i64, i64]
    L.clko_maximum_z_projection.restype = i32
    L.clko_separable_2d.argtypes = [i32, vp, i32, vp, i32, i64, i64, i32, i32, f32, i32]
    L.clko_separable_2d.restype = i32
    L.clko_median_box_2d.argtypes = [vp, i32, vp, i32, i64, i64, i32, i32, i32]
    L.clko_median_box_2d.restype = i32
    L.clko_expf.argtypes = [f32]
    L.clko_expf.restype = f32
    L.clko_convert.argtypes = [f32, i32]
    L.clko_convert.restype = f32
    L.clko_div_uniform.argtypes = [f32, f32, f32]
    L.clko_div_uniform.restype = f32
    L.clko_div_uniform_binade_mismatches.argtypes = [f32]
    L.clko_div_uniform_binade_mismatches.restype = i64
    L.clko_num_threads.argtypes = []
    L.clko_num_threads.restype = i32
    return L


def default_threads():
    return max(1, lib().clko_num_threads())


def _vol(a):
    a = np.ascontiguousarray(a)
    if a.ndim != 3:
        raise ValueError("volumes are (z, y, x) arrays")
    if a.dtype not in DTYPE_CODES:
        raise ValueError("unsupported pixel type %s" % a.dtype)
    return a


def separable(op, src, dst_dtype, dim, N, s=0.0, nthreads=None, fast=False):
    """One single-axis launch of gaussian/mean/minimum/maximum_separable_3d
    (fast=True: through the -O3 x86-64-v3 build, same results)."""
    if isinstance(op, str):
        op = OP_NAMES[op]
    src = _vol(src)
    dst = np.empty(src.shape, dtype=np.dtype(dst_dtype))
    D, H, W = src.shape
    rc = (fast_lib() if fast else lib()).clko_separable(
        op, dst.ctypes.data, DTYPE_CODES[dst.dtype], src.ctypes.data, DTYPE_CODES[src.dtype],
        W, H, D, int(dim), int(N), float(s), int(nthreads or default_threads()))
    if rc != 0:
        raise ValueError("oracle rejected arguments (%d)" % rc)
    return dst


def gaussian_blur_separable_3d(src, dst_dtype, dim, N, s, nthreads=None):
    return separable(OP_GAUSS, src, dst_dtype, dim, N, s, nthreads)


def mean_separable_3d(src, dst_dtype, dim, N, s=0.0, nthreads=None):
    return separable(OP_MEAN, src, dst_dtype, dim, N, s, nthreads)


def minimum_separable_3d(src, dst_dtype, dim, N, s=0.0, nthreads=None):
    return separable(OP_MIN, src, dst_dtype, dim, N, s, nthreads)


def maximum_separable_3d(src, dst_dtype, dim, N, s=0.0, nthreads=None):
    return separable(OP_MAX, src, dst_dtype, dim, N, s, nthreads)


def median_box_3d(src, dst_dtype, Nx, Ny, Nz, nthreads=None):
    src = _vol(src)
    dst = np.empty(src.shape, dtype=np.dtype(dst_dtype))
    D, H, W = src.shape
    rc = lib().clko_median_box_3d(
        dst.ctypes.data, DTYPE_CODES[dst.dtype], src.ctypes.data, DTYPE_CODES[src.dtype],
        W, H, D, int(Nx), int(Ny), int(Nz), int(nthreads or default_threads()))
    if rc != 0:
        raise ValueError("oracle rejected arguments (%d)" % rc)
    return dst


def separable_2d(op, src, dst_dtype, dim, N, s=0.0, nthreads=None):
    """The 2D separable kernels (…_separable_2d_x.cl) on a (H, W) image."""
    if isinstance(op, str):
        op = OP_NAMES[op.replace("_2d", "_3d")]
    src = np.ascontiguousarray(src)
    assert src.ndim == 2
    dst = np.empty(src.shape, dtype=np.dtype(dst_dtype))
    H, W = src.shape
    rc = lib().clko_separable_2d(op, dst.ctypes.data, DTYPE_CODES[dst.dtype], src.ctypes.data,
                                 DTYPE_CODES[src.dtype], W, H, int(dim), int(N), float(s),
                                 int(nthreads or default_threads()))
    if rc != 0:
        raise ValueError("oracle rejected arguments (%d)" % rc)
    return dst


def median_box_2d(src, dst_dtype, Nx, Ny, nthreads=None):
    """median_box_2d_x.cl on a (H, W) image (C-cast narrowing of integers)."""
    src = np.ascontiguousarray(src)
    assert src.ndim == 2
    dst = np.empty(src.shape, dtype=np.dtype(dst_dtype))
    H, W = src.shape
    rc = lib().clko_median_box_2d(dst.ctypes.data, DTYPE_CODES[dst.dtype], src.ctypes.data,
                                  DTYPE_CODES[src.dtype], W, H, int(Nx), int(Ny), int(nthreads or default_threads()))
    if rc != 0:
        raise ValueError("oracle rejected arguments (%d)" % rc)
    return dst


SPHERE_OPS = {"minimum_sphere_3d": 0, "maximum_sphere_3d": 1, "mean_sphere_3d": 2, "median_sphere_3d": 3}


def sphere_3d(op, src, dst_dtype, Nx, Ny, Nz, nthreads=None):
    """minimum/maximum/mean/median_sphere_3d on a (D, H, W) volume."""
    op = SPHERE_OPS[op] if isinstance(op, str) else int(op)
    src = _vol(src)
    dst = np.empty(src.shape, dtype=np.dtype(dst_dtype))
    D, H, W = src.shape
    rc = lib().clko_sphere_3d(op, dst.ctypes.data, DTYPE_CODES[dst.dtype], src.ctypes.data, DTYPE_CODES[src.dtype],
                              W, H, D, int(Nx), int(Ny), int(Nz), int(nthreads or default_threads()))
    if rc != 0:
        raise ValueError("oracle rejected arguments (%d)" % rc)
    return dst


def median_slice_by_slice_box_3d(src, dst_dtype, Nx, Ny, nthreads=None):
    src = _vol(src)
    dst = np.empty(src.shape, dtype=np.dtype(dst_dtype))
    D, H, W = src.shape
    rc = lib().clko_median_slice_by_slice_box_3d(dst.ctypes.data, DTYPE_CODES[dst.dtype], src.ctypes.data,
                                                 DTYPE_CODES[src.dtype], W, H, D, int(Nx), int(Ny),
                                                 int(nthreads or default_threads()))
    if rc != 0:
        raise ValueError("oracle rejected arguments (%d)" % rc)
    return dst


def copy_3d(src, dst_dtype):
    src = np.ascontiguousarray(src)
    dst = np.empty(src.shape, dtype=np.dtype(dst_dtype))
    rc = lib().clko_copy_3d(dst.ctypes.data, DTYPE_CODES[dst.dtype], src.ctypes.data, DTYPE_CODES[src.dtype], src.size)
    if rc != 0:
        raise ValueError("oracle rejected arguments (%d)" % rc)
    return dst


def detect_maxima_3d(src, dst_dtype):
    src = _vol(src)
    dst = np.empty(src.shape, dtype=np.dtype(dst_dtype))
    D, H, W = src.shape
    rc = lib().clko_detect_maxima_3d(src.ctypes.data, DTYPE_CODES[src.dtype], dst.ctypes.data,
                                     DTYPE_CODES[dst.dtype], W, H, D)
    if rc != 0:
        raise ValueError("oracle rejected arguments (%d)" % rc)
    return dst


def maximum_z_projection(src, dst_dtype):
    """(D, H, W) -> (1, H, W)."""
    src = _vol(src)
    D, H, W = src.shape
    dst = np.empty((1, H, W), dtype=np.dtype(dst_dtype))
    rc = lib().clko_maximum_z_projection(dst.ctypes.data, DTYPE_CODES[dst.dtype], src.ctypes.data,
                                         DTYPE_CODES[src.dtype], W, H, D)
    if rc != 0:
        raise ValueError("oracle rejected arguments (%d)" % rc)
    return dst


def convert(values, dst_dtype):
    """CONVERT_dst_PIXEL_TYPE applied to each float32 value."""
    code = DTYPE_CODES[np.dtype(dst_dtype)]
    L = lib()
    flat = np.asarray(values, dtype=np.float32).ravel()
    out = np.array([L.clko_convert(float(v), code) for v in flat], dtype=np.float32)
    return out.astype(np.dtype(dst_dtype)).reshape(np.shape(values))


def expf(x):
    return np.float32(lib().clko_expf(float(x)))


def sigma_to_n(sigma):
    """CLIJ's host convention for a Gaussian kernel size: int(8*sigma), odd."""
    n = int(np.float32(sigma) * np.float32(8.0))
    return n + 1 if n % 2 == 0 else n


def chain3(op, src, dst_dtype, Ns, ss=(0.0, 0.0, 0.0), active=(True, True, True), nthreads=None):
    """The 3-pass host chain (x, y, z) with float32 intermediates; inactive
    axes are skipped; with none active, a converting copy."""
    acts = [a for a in range(3) if active[a]]
    if not acts:
        return convert(np.asarray(src, dtype=np.float32), dst_dtype).reshape(src.shape)
    cur = src
    for i, a in enumerate(acts):
        out_t = dst_dtype if i == len(acts) - 1 else np.float32
        cur = separable(op, cur, out_t, a, Ns[a], ss[a], nthreads)
    return cur


def gaussian_blur_3d(src, dst_dtype, sx, sy, sz, nthreads=None):
    sig = (sx, sy, sz)
    return chain3(OP_GAUSS, src, dst_dtype, [sigma_to_n(s) for s in sig], sig,
                  [s > 0 for s in sig], nthreads)


def box_3d(op, src, dst_dtype, rx, ry, rz, nthreads=None):
    r = (rx, ry, rz)
    return chain3(op, src, dst_dtype, [2 * v + 1 for v in r], (0.0, 0.0, 0.0),
                  [v > 0 for v in r], nthreads)


# ---- FP64 reference of the Gaussian pass (independent of the restatement) ----

def gaussian_pass_fp64(src, dim, N, s):
    """Single-axis Gaussian in float64: normalised weights exp(-v^2/(2 s^2))
    over v = -c..c, clamp-to-edge border (scipy 'nearest')."""
    from scipy import ndimage
    c = int((N - 1) / 2) if N >= 1 else 0
    v = np.arange(-c, c + 1, dtype=np.float64)
    w = np.exp(-(v * v) / (2.0 * float(s) * float(s)))
    w /= w.sum()
    x = np.asarray(src, dtype=np.float64)
    if dim not in (0, 1, 2):
        return x.copy()
    axis = 2 - dim  # numpy axes are (z, y, x)
    return ndimage.correlate1d(x, w, axis=axis, mode="nearest")
