"""ctypes binding of libclkmi.so (the C-ABI declared in include/clkmi.h).

This is the binding a Python CLIJ/clEsperanto-style host uses in place of an
OpenCL enqueue.  Loading fails loudly when the library has not been built:
there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libclkmi.so")

CLK_CHAR, CLK_UCHAR, CLK_SHORT, CLK_USHORT, CLK_FLOAT = 0, 1, 2, 3, 4
CLK_ARG_IMAGE, CLK_ARG_INT, CLK_ARG_FLOAT = 0, 1, 2

ERROR_NAMES = {
    -4: "CL_MEM_OBJECT_ALLOCATION_FAILURE",
    -5: "CL_OUT_OF_RESOURCES",
    -8: "CL_MEM_COPY_OVERLAP",
    -30: "CL_INVALID_VALUE",
    -38: "CL_INVALID_MEM_OBJECT",
    -46: "CL_INVALID_KERNEL_NAME",
    -52: "CL_INVALID_KERNEL_ARGS",
    -53: "CL_INVALID_WORK_DIMENSION",
    -59: "CL_INVALID_OPERATION",
}

# Every symbol the header declares; tests check the library exports them all.
EXPORTED = (
    "clk_gaussian_blur_separable_3d", "clk_mean_separable_3d", "clk_minimum_separable_3d",
    "clk_maximum_separable_3d", "clk_median_box_3d", "clk_execute", "clk_gaussian_blur_3d",
    "clk_mean_box_3d", "clk_minimum_box_3d", "clk_maximum_box_3d", "clk_init", "clk_malloc",
    "clk_free", "clk_memcpy_h2d", "clk_memcpy_d2h", "clk_stream_create", "clk_stream_destroy",
    "clk_stream_sync", "clk_dtype_size", "clk_last_error_string", "clk_version",
    "clk_batch_filter_3d", "clk_set_tuning", "clk_get_tuning", "clk_autotune_save", "clk_autotune_load",
    "clk_gaussian_blur_separable_2d", "clk_mean_separable_2d", "clk_minimum_separable_2d",
    "clk_maximum_separable_2d", "clk_median_box_2d", "clk_gaussian_blur_2d", "clk_mean_box_2d",
    "clk_minimum_box_2d", "clk_maximum_box_2d", "clk_minimum_sphere_3d", "clk_maximum_sphere_3d",
    "clk_mean_sphere_3d", "clk_median_sphere_3d", "clk_median_slice_by_slice_box_3d", "clk_copy_3d",
    "clk_detect_maxima_3d", "clk_maximum_z_projection", "clk_release_temporaries", "clk_temporaries_held",
    "clk_filter_3d_noalloc",
)

CLK_JOB_GAUSSIAN_BLUR_3D, CLK_JOB_MEAN_BOX_3D, CLK_JOB_MINIMUM_BOX_3D, CLK_JOB_MAXIMUM_BOX_3D, \
    CLK_JOB_MEDIAN_BOX_3D = 0, 1, 2, 3, 4


class ClkImage(ctypes.Structure):
    _fields_ = [
        ("ptr", ctypes.c_void_p),
        ("dtype", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("width", ctypes.c_int64),
        ("height", ctypes.c_int64),
        ("depth", ctypes.c_int64),
    ]


class ClkArg(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char_p),
        ("kind", ctypes.c_int32),
        ("i", ctypes.c_int32),
        ("f", ctypes.c_float),
        ("pad", ctypes.c_int32),
        ("image", ctypes.POINTER(ClkImage)),
    ]


class ClkFilterJob(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("dst", ClkImage),
        ("src", ClkImage),
        ("sigma", ctypes.c_float * 3),
        ("radius", ctypes.c_int32 * 3),
    ]


class ClkError(RuntimeError):
    """A negative OpenCL-style status returned by libclkmi."""

    def __init__(self, code, message):
        self.code = code
        super().__init__("%s (%d): %s" % (ERROR_NAMES.get(code, "error"), code, message))


_lib = None


def lib():
    """Load libclkmi.so once.  Raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libclkmi.so is not built (expected at %s); run __graft_entry__.build() "
                "or `make -C clij-opencl-kernels_amd`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, f32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_float, ctypes.c_int64
        img = ctypes.POINTER(ClkImage)
        for name in ("clk_gaussian_blur_separable_3d", "clk_mean_separable_3d",
                     "clk_minimum_separable_3d", "clk_maximum_separable_3d"):
            fn = getattr(L, name)
            fn.argtypes = [vp, img, img, i32, i32, f32]
            fn.restype = i32
        L.clk_median_box_3d.argtypes = [vp, img, img, i32, i32, i32]
        L.clk_median_box_3d.restype = i32
        for name in ("clk_gaussian_blur_separable_2d", "clk_mean_separable_2d",
                     "clk_minimum_separable_2d", "clk_maximum_separable_2d"):
            fn = getattr(L, name)
            fn.argtypes = [vp, img, img, i32, i32, f32]
            fn.restype = i32
        for name in ("clk_minimum_sphere_3d", "clk_maximum_sphere_3d", "clk_mean_sphere_3d",
                     "clk_median_sphere_3d"):
            fn = getattr(L, name)
            fn.argtypes = [vp, img, img, i32, i32, i32]
            fn.restype = i32
        L.clk_median_slice_by_slice_box_3d.argtypes = [vp, img, img, i32, i32]
        L.clk_median_slice_by_slice_box_3d.restype = i32
        L.clk_copy_3d.argtypes = [vp, img, img]
        L.clk_copy_3d.restype = i32
        L.clk_detect_maxima_3d.argtypes = [vp, img, img]
        L.clk_detect_maxima_3d.restype = i32
        L.clk_maximum_z_projection.argtypes = [vp, img, img]
        L.clk_maximum_z_projection.restype = i32
        L.clk_median_box_2d.argtypes = [vp, img, img, i32, i32]
        L.clk_median_box_2d.restype = i32
        L.clk_gaussian_blur_2d.argtypes = [vp, img, img, f32, f32]
        L.clk_gaussian_blur_2d.restype = i32
        for name in ("clk_mean_box_2d", "clk_minimum_box_2d", "clk_maximum_box_2d"):
            fn = getattr(L, name)
            fn.argtypes = [vp, img, img, i32, i32]
            fn.restype = i32
        L.clk_execute.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ClkArg), i32]
        L.clk_execute.restype = i32
        L.clk_gaussian_blur_3d.argtypes = [vp, img, img, f32, f32, f32]
        L.clk_gaussian_blur_3d.restype = i32
        for name in ("clk_mean_box_3d", "clk_minimum_box_3d", "clk_maximum_box_3d"):
            fn = getattr(L, name)
            fn.argtypes = [vp, img, img, i32, i32, i32]
            fn.restype = i32
        L.clk_batch_filter_3d.argtypes = [vp, ctypes.POINTER(ClkFilterJob), i32]
        L.clk_batch_filter_3d.restype = i32
        L.clk_filter_3d_noalloc.argtypes = [vp, ctypes.POINTER(ClkFilterJob)]
        L.clk_filter_3d_noalloc.restype = i32
        L.clk_release_temporaries.argtypes = [vp]
        L.clk_release_temporaries.restype = i32
        L.clk_temporaries_held.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.clk_temporaries_held.restype = i32
        L.clk_set_tuning.argtypes = [ctypes.c_char_p, i32]
        L.clk_set_tuning.restype = i32
        L.clk_get_tuning.argtypes = [ctypes.c_char_p, ctypes.POINTER(i32)]
        L.clk_get_tuning.restype = i32
        for name in ("clk_autotune_save", "clk_autotune_load"):
            fn = getattr(L, name)
            fn.argtypes = [ctypes.c_char_p]
            fn.restype = i32
        L.clk_init.argtypes = []
        L.clk_init.restype = i32
        L.clk_malloc.argtypes = [ctypes.POINTER(vp), i64]
        L.clk_malloc.restype = i32
        L.clk_free.argtypes = [vp]
        L.clk_free.restype = i32
        L.clk_memcpy_h2d.argtypes = [vp, vp, i64, vp]
        L.clk_memcpy_h2d.restype = i32
        L.clk_memcpy_d2h.argtypes = [vp, vp, i64, vp]
        L.clk_memcpy_d2h.restype = i32
        L.clk_stream_create.argtypes = [ctypes.POINTER(vp)]
        L.clk_stream_create.restype = i32
        L.clk_stream_destroy.argtypes = [vp]
        L.clk_stream_destroy.restype = i32
        L.clk_stream_sync.argtypes = [vp]
        L.clk_stream_sync.restype = i32
        L.clk_dtype_size.argtypes = [i32]
        L.clk_dtype_size.restype = i32
        L.clk_last_error_string.argtypes = []
        L.clk_last_error_string.restype = ctypes.c_char_p
        L.clk_version.argtypes = []
        L.clk_version.restype = i32
        _lib = L
    return _lib


def check(rc):
    """Raise ClkError for a negative status."""
    if rc != 0:
        msg = lib().clk_last_error_string()
        raise ClkError(rc, msg.decode() if msg else "")
    return rc
