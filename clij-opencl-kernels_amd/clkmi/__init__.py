"""clkmi: MI355X (gfx950) implementation of the 3D neighbourhood filters of
clEsperanto/clij-opencl-kernels behind the reference's kernel contract.

The functions below mirror how a CLIJ-style host launches the reference
kernels: by kernel name with named parameters (``execute``), or one function
per kernel with the parameters in the kernel's declaration order
(``dst, src, dim, N, s`` for the separable kernels, ``dst, src, Nx, Ny, Nz``
for ``median_box_3d``).  Images are torch tensors on the GPU shaped
(depth, height, width) -- x fastest, as the reference's buffers -- or
``Image`` descriptors holding a raw device pointer.  Every call goes through
libclkmi.so (ctypes); there is no CPU fallback.
"""
import ctypes
from collections import namedtuple

from . import _lib
from ._lib import ClkError, lib, check

__all__ = [
    "ClkError", "Image", "execute", "gaussian_blur_separable_3d", "mean_separable_3d",
    "minimum_separable_3d", "maximum_separable_3d", "median_box_3d", "gaussian_blur_3d",
    "mean_box_3d", "minimum_box_3d", "maximum_box_3d", "sigma_to_kernel_size", "KERNELS",
    "FilterJob", "batch_filter_3d", "set_tuning", "get_tuning", "autotune_save", "autotune_load",
    "gaussian_blur_separable_2d", "mean_separable_2d", "minimum_separable_2d", "maximum_separable_2d",
    "median_box_2d", "gaussian_blur_2d", "mean_box_2d", "minimum_box_2d", "maximum_box_2d",
    "minimum_sphere_3d", "maximum_sphere_3d", "mean_sphere_3d", "median_sphere_3d",
    "median_slice_by_slice_box_3d", "copy_3d", "detect_maxima_3d", "maximum_z_projection",
    "release_temporaries", "temporaries_held",
 "pack_jobs", "PackedJobs",]

# Kernel name -> declared parameters (reference .cl declarations).
KERNELS = {
    "gaussian_blur_separable_3d": ("dst", "src", "dim", "N", "s"),  # gaussian_blur_separable_3d_x.cl:8-11
    "mean_separable_3d": ("dst", "src", "dim", "N", "s"),           # mean_separable_3d_x.cl:5-9
    "minimum_separable_3d": ("dst", "src", "dim", "N", "s"),        # minimum_separable_3d_x.cl:5-9
    "maximum_separable_3d": ("dst", "src", "dim", "N", "s"),        # maximum_separable_3d_x.cl:5-9
    "median_box_3d": ("dst", "src", "Nx", "Ny", "Nz"),              # median_box_3d_x.cl:25-29
    "gaussian_blur_separable_2d": ("dst", "src", "dim", "N", "s"),  # gaussian_blur_separable_2d_x.cl:8-12
    "mean_separable_2d": ("dst", "src", "dim", "N", "s"),           # mean_separable_2d_x.cl:5-9
    "minimum_separable_2d": ("dst", "src", "dim", "N", "s"),        # minimum_separable_2d_x.cl:5-9
    "maximum_separable_2d": ("dst", "src", "dim", "N", "s"),        # maximum_separable_2d_x.cl:5-9
    "median_box_2d": ("dst", "src", "Nx", "Ny"),                    # median_box_2d_x.cl:24-27
    "minimum_sphere_3d": ("dst", "src", "Nx", "Ny", "Nz"),          # minimum_sphere_3d_x.cl:5-9
    "maximum_sphere_3d": ("dst", "src", "Nx", "Ny", "Nz"),          # maximum_sphere_3d_x.cl:5-9
    "mean_sphere_3d": ("dst", "src", "Nx", "Ny", "Nz"),             # mean_sphere_3d_x.cl:5-9
    "median_sphere_3d": ("dst", "src", "Nx", "Ny", "Nz"),           # median_sphere_3d_x.cl:24-28
    "median_slice_by_slice_box_3d": ("dst", "src", "Nx", "Ny"),     # median_slice_by_slice_box_3d_x.cl:24-27
    "copy_3d": ("dst", "src"),                                      # copy_3d_x.cl:2-3
    "detect_maxima_3d": ("src", "dst"),                             # detect_maxima_3d_x.cl:3-4
    "maximum_z_projection": ("dst_max", "src"),                     # maximum_z_projection_x.cl:3-4
}

Image = namedtuple("Image", "ptr dtype width height depth")
Image.__doc__ = "A raw device buffer: pointer, clk dtype code, and size (x, y, z)."


def _torch_codes():
    import torch
    return {
        torch.int8: _lib.CLK_CHAR,
        torch.uint8: _lib.CLK_UCHAR,
        torch.int16: _lib.CLK_SHORT,
        torch.uint16: _lib.CLK_USHORT,
        torch.float32: _lib.CLK_FLOAT,
    }


def to_clk_image(img):
    """Describe a torch tensor (z, y, x) or (y, x), or an Image, as a clk_image."""
    if isinstance(img, Image):
        return _lib.ClkImage(img.ptr, int(img.dtype), 0, int(img.width), int(img.height), int(img.depth))
    codes = _torch_codes()
    if img.dtype not in codes:
        raise TypeError("unsupported pixel type %s (char/uchar/short/ushort/float)" % img.dtype)
    if img.dim() not in (2, 3):
        raise ValueError("images are tensors shaped (depth, height, width) or (height, width)")
    if not img.is_contiguous():
        raise ValueError("images must be contiguous")
    if img.device.type != "cuda":
        raise ValueError("images must live on the GPU")
    d, h, w = img.shape if img.dim() == 3 else (1,) + tuple(img.shape)
    return _lib.ClkImage(img.data_ptr(), codes[img.dtype], 0, w, h, d)


def _stream_ptr(stream):
    if stream is None:
        try:
            import torch
            return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        except Exception:
            return ctypes.c_void_p(0)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


def _sep(name, dst, src, dim, N, s, stream):
    d, sr = to_clk_image(dst), to_clk_image(src)
    fn = getattr(lib(), "clk_" + name)
    check(fn(_stream_ptr(stream), ctypes.byref(d), ctypes.byref(sr), int(dim), int(N), float(s)))
    return dst


def gaussian_blur_separable_3d(dst, src, dim, N, s, stream=None):
    """gaussian_blur_separable_3d(dst, src, dim, N, s): 1D Gaussian along axis dim."""
    return _sep("gaussian_blur_separable_3d", dst, src, dim, N, s, stream)


def mean_separable_3d(dst, src, dim, N, s=0.0, stream=None):
    """mean_separable_3d(dst, src, dim, N, s): 1D box mean along axis dim."""
    return _sep("mean_separable_3d", dst, src, dim, N, s, stream)


def minimum_separable_3d(dst, src, dim, N, s=0.0, stream=None):
    """minimum_separable_3d(dst, src, dim, N, s): 1D box minimum along axis dim."""
    return _sep("minimum_separable_3d", dst, src, dim, N, s, stream)


def maximum_separable_3d(dst, src, dim, N, s=0.0, stream=None):
    """maximum_separable_3d(dst, src, dim, N, s): 1D box maximum along axis dim."""
    return _sep("maximum_separable_3d", dst, src, dim, N, s, stream)


def median_box_3d(dst, src, Nx, Ny, Nz, stream=None):
    """median_box_3d(dst, src, Nx, Ny, Nz): median of the clamped box."""
    d, sr = to_clk_image(dst), to_clk_image(src)
    check(lib().clk_median_box_3d(_stream_ptr(stream), ctypes.byref(d), ctypes.byref(sr),
                                  int(Nx), int(Ny), int(Nz)))
    return dst


def gaussian_blur_separable_2d(dst, src, dim, N, s, stream=None):
    """gaussian_blur_separable_2d(dst, src, dim, N, s) on images of depth 1."""
    return _sep("gaussian_blur_separable_2d", dst, src, dim, N, s, stream)


def mean_separable_2d(dst, src, dim, N, s=0.0, stream=None):
    return _sep("mean_separable_2d", dst, src, dim, N, s, stream)


def minimum_separable_2d(dst, src, dim, N, s=0.0, stream=None):
    return _sep("minimum_separable_2d", dst, src, dim, N, s, stream)


def maximum_separable_2d(dst, src, dim, N, s=0.0, stream=None):
    return _sep("maximum_separable_2d", dst, src, dim, N, s, stream)


def median_box_2d(dst, src, Nx, Ny, stream=None):
    """median_box_2d(dst, src, Nx, Ny): integer values narrowed by a C cast."""
    d, sr = to_clk_image(dst), to_clk_image(src)
    check(lib().clk_median_box_2d(_stream_ptr(stream), ctypes.byref(d), ctypes.byref(sr), int(Nx), int(Ny)))
    return dst


def _nbh(name, dst, src, sizes, stream):
    d, sr = to_clk_image(dst), to_clk_image(src)
    check(getattr(lib(), "clk_" + name)(_stream_ptr(stream), ctypes.byref(d), ctypes.byref(sr),
                                        *[int(v) for v in sizes]))
    return dst


def minimum_sphere_3d(dst, src, Nx, Ny, Nz, stream=None):
    return _nbh("minimum_sphere_3d", dst, src, (Nx, Ny, Nz), stream)


def maximum_sphere_3d(dst, src, Nx, Ny, Nz, stream=None):
    return _nbh("maximum_sphere_3d", dst, src, (Nx, Ny, Nz), stream)


def mean_sphere_3d(dst, src, Nx, Ny, Nz, stream=None):
    return _nbh("mean_sphere_3d", dst, src, (Nx, Ny, Nz), stream)


def median_sphere_3d(dst, src, Nx, Ny, Nz, stream=None):
    return _nbh("median_sphere_3d", dst, src, (Nx, Ny, Nz), stream)


def median_slice_by_slice_box_3d(dst, src, Nx, Ny, stream=None):
    return _nbh("median_slice_by_slice_box_3d", dst, src, (Nx, Ny), stream)


def copy_3d(dst, src, stream=None):
    """dst = CONVERT(src) (copy_3d_x.cl)."""
    return _nbh("copy_3d", dst, src, (), stream)


def detect_maxima_3d(src, dst, stream=None):
    """detect_maxima_3d(src, dst): 1 at local maxima of the 3x3x3 box (note src first)."""
    d, sr = to_clk_image(dst), to_clk_image(src)
    check(lib().clk_detect_maxima_3d(_stream_ptr(stream), ctypes.byref(sr), ctypes.byref(d)))
    return dst


def maximum_z_projection(dst_max, src, stream=None):
    """maximum_z_projection(dst_max, src): maximum over z into a W x H x 1 image."""
    d, sr = to_clk_image(dst_max), to_clk_image(src)
    check(lib().clk_maximum_z_projection(_stream_ptr(stream), ctypes.byref(d), ctypes.byref(sr)))
    return dst_max


def execute(kernel_name, parameters, stream=None):
    """Launch a kernel by name with parameters bound by name, as a CLIJ host's
    execute(kernelName, parameterMap).  The name may also be the program file
    name (e.g. "median_box_3d_x.cl")."""
    images = []
    args = (_lib.ClkArg * len(parameters))()
    for i, (name, value) in enumerate(parameters.items()):
        a = args[i]
        a.name = name.encode()
        if isinstance(value, (int, bool)) and not isinstance(value, float):
            a.kind, a.i = _lib.CLK_ARG_INT, int(value)
        elif isinstance(value, float):
            a.kind, a.f = _lib.CLK_ARG_FLOAT, float(value)
        else:
            im = to_clk_image(value)
            images.append(im)
            a.kind = _lib.CLK_ARG_IMAGE
            a.image = ctypes.pointer(im)
    check(lib().clk_execute(_stream_ptr(stream), kernel_name.encode(), args, len(parameters)))
    return parameters.get("dst", parameters.get("dst_max"))


def sigma_to_kernel_size(sigma):
    """int(8*sigma), made odd: the CLIJ host convention for a Gaussian's N."""
    import numpy as np
    n = int(np.float32(sigma) * np.float32(8.0))
    return n + 1 if n % 2 == 0 else n


def gaussian_blur_3d(src, dst, sigma_x, sigma_y, sigma_z, stream=None):
    """3D Gaussian as three single-axis passes (float32 intermediates)."""
    d, sr = to_clk_image(dst), to_clk_image(src)
    check(lib().clk_gaussian_blur_3d(_stream_ptr(stream), ctypes.byref(d), ctypes.byref(sr),
                                     float(sigma_x), float(sigma_y), float(sigma_z)))
    return dst


def gaussian_blur_2d(src, dst, sigma_x, sigma_y, stream=None):
    """2D Gaussian (depth-1 images) as two single-axis passes."""
    d, sr = to_clk_image(dst), to_clk_image(src)
    check(lib().clk_gaussian_blur_2d(_stream_ptr(stream), ctypes.byref(d), ctypes.byref(sr),
                                     float(sigma_x), float(sigma_y)))
    return dst


def _box2(name, src, dst, rx, ry, stream):
    d, sr = to_clk_image(dst), to_clk_image(src)
    check(getattr(lib(), "clk_" + name)(_stream_ptr(stream), ctypes.byref(d), ctypes.byref(sr), int(rx), int(ry)))
    return dst


def mean_box_2d(src, dst, radius_x, radius_y, stream=None):
    return _box2("mean_box_2d", src, dst, radius_x, radius_y, stream)


def minimum_box_2d(src, dst, radius_x, radius_y, stream=None):
    return _box2("minimum_box_2d", src, dst, radius_x, radius_y, stream)


def maximum_box_2d(src, dst, radius_x, radius_y, stream=None):
    return _box2("maximum_box_2d", src, dst, radius_x, radius_y, stream)


def _box(name, src, dst, rx, ry, rz, stream):
    d, sr = to_clk_image(dst), to_clk_image(src)
    check(getattr(lib(), "clk_" + name)(_stream_ptr(stream), ctypes.byref(d), ctypes.byref(sr),
                                        int(rx), int(ry), int(rz)))
    return dst


def mean_box_3d(src, dst, radius_x, radius_y, radius_z, stream=None):
    return _box("mean_box_3d", src, dst, radius_x, radius_y, radius_z, stream)


def minimum_box_3d(src, dst, radius_x, radius_y, radius_z, stream=None):
    return _box("minimum_box_3d", src, dst, radius_x, radius_y, radius_z, stream)


def maximum_box_3d(src, dst, radius_x, radius_y, radius_z, stream=None):
    return _box("maximum_box_3d", src, dst, radius_x, radius_y, radius_z, stream)


# ---- batches ---------------------------------------------------------------

JOB_KINDS = {
    "gaussian_blur_3d": _lib.CLK_JOB_GAUSSIAN_BLUR_3D,
    "mean_box_3d": _lib.CLK_JOB_MEAN_BOX_3D,
    "minimum_box_3d": _lib.CLK_JOB_MINIMUM_BOX_3D,
    "maximum_box_3d": _lib.CLK_JOB_MAXIMUM_BOX_3D,
    "median_box_3d": _lib.CLK_JOB_MEDIAN_BOX_3D,
}


class FilterJob(namedtuple("FilterJob", "kind dst src sigma radius")):
    """One 3D filter of a batch: kind is a JOB_KINDS name; sigma (Gaussian)
    or radius (box filters, median) per axis x, y, z."""

    def __new__(cls, kind, dst, src, sigma=(0.0, 0.0, 0.0), radius=(0, 0, 0)):
        return super().__new__(cls, kind, dst, src, tuple(sigma), tuple(radius))


def _image_fields(img):
    """(ptr, dtype code, width, height, depth) of an Image or a tensor."""
    if isinstance(img, Image):
        return img.ptr, int(img.dtype), int(img.width), int(img.height), int(img.depth)
    c = to_clk_image(img)
    return c.ptr or 0, c.dtype, c.width, c.height, c.depth


class PackedJobs:
    """A batch's clk_filter_job array, packed once (numpy, in the C layout)
    and reusable for every call on the same images."""

    def __init__(self, jobs):
        import numpy as np
        n = len(jobs)
        a = np.zeros(max(1, n), dtype=np.dtype(_lib.ClkFilterJob))
        if n:
            a["kind"][:n] = [JOB_KINDS[j.kind] if isinstance(j.kind, str) else int(j.kind) for j in jobs]
            for side in ("dst", "src"):
                f = np.array([_image_fields(getattr(j, side)) for j in jobs], dtype=np.int64)
                col = a[side]
                col["ptr"][:n] = f[:, 0].astype(np.uint64)
                col["dtype"][:n] = f[:, 1]
                col["width"][:n] = f[:, 2]
                col["height"][:n] = f[:, 3]
                col["depth"][:n] = f[:, 4]
            a["sigma"][:n] = [j.sigma for j in jobs]
            a["radius"][:n] = [j.radius for j in jobs]
        self.array = a
        self.n = n
        self.ptr = ctypes.c_void_p(a.ctypes.data)


def pack_jobs(jobs):
    """Pack a list of FilterJob once for repeated batch_filter_3d calls."""
    return PackedJobs(jobs)


def batch_filter_3d(jobs, stream=None):
    """Run many independent 3D filters as one batch (a few launches per
    pass stage instead of one per volume and axis); results equal the single
    calls bit for bit.  `jobs` is a list of FilterJob or a PackedJobs."""
    packed = jobs if isinstance(jobs, PackedJobs) else PackedJobs(jobs)
    check(lib().clk_batch_filter_3d(_stream_ptr(stream), ctypes.cast(packed.ptr, ctypes.POINTER(_lib.ClkFilterJob)),
                                    packed.n))


def release_temporaries(stream=None):
    """Free the temporaries kept for `stream` (waits for it) and return the
    library's unused pooled device memory to the driver."""
    check(lib().clk_release_temporaries(_stream_ptr(stream)))


def temporaries_held(stream=None):
    """(bytes kept for `stream`, bytes the library's pool holds on the device)."""
    a, b = ctypes.c_int64(), ctypes.c_int64()
    check(lib().clk_temporaries_held(_stream_ptr(stream), ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


def set_tuning(key, value):
    """Set a launch-shape knob (results never depend on it)."""
    check(lib().clk_set_tuning(key.encode(), int(value)))


def get_tuning(key):
    v = ctypes.c_int32()
    check(lib().clk_get_tuning(key.encode(), ctypes.byref(v)))
    return v.value


def autotune_save(path):
    """Write the launch shapes chosen by the first-launch autotuner to a file."""
    check(lib().clk_autotune_save(str(path).encode()))


def autotune_load(path):
    """Merge launch shapes saved by autotune_save (no timing runs needed)."""
    check(lib().clk_autotune_load(str(path).encode()))
