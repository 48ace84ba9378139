/*
 * clkmi.h -- C-ABI of the MI355X (gfx950) implementation of the 3D
 * neighbourhood-filter kernels of clEsperanto/clij-opencl-kernels.
 *
 * Each typed entry point replaces one OpenCL enqueue of the kernel of the
 * same name: same named parameters, same meaning, same pixel types
 * (char/uchar/short/ushort/float), clamp-to-edge border and saturating
 * CONVERT_dst_PIXEL_TYPE.  The launch domain is the dst image, exactly as the
 * reference host enqueues global = (W,H,D) of dst.
 *
 * Conventions
 *   - Images are dense device buffers, x fastest:
 *       offset(x,y,z) = x + y*width + z*width*height      (elements)
 *   - Every entry point is asynchronous on `stream` (a hipStream_t, NULL =
 *     the null stream) and returns 0 or a negative OpenCL-style error code.
 *   - No C++ exception crosses this boundary.  clk_last_error_string() gives
 *     the message of the last failing call on the calling thread.
 *   - Calls are thread-safe.  The library borrows the caller's pointers until
 *     the stream reaches the enqueued work; it never frees caller memory.
 *   - src and dst must not overlap (the kernels read neighbours of src);
 *     overlap returns CLK_MEM_COPY_OVERLAP.
 */
#ifndef CLKMI_H
#define CLKMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Pixel types: the values IMAGE_*_PIXEL_TYPE expands to
 * (reference README.md:64-73). */
typedef enum clk_dtype {
  CLK_CHAR   = 0, /* int8   */
  CLK_UCHAR  = 1, /* uint8  */
  CLK_SHORT  = 2, /* int16  */
  CLK_USHORT = 3, /* uint16 */
  CLK_FLOAT  = 4  /* float32 */
} clk_dtype;

/* Error codes: the OpenCL cl_int values a CLIJ host already handles. */
#define CLK_SUCCESS                          0
#define CLK_OUT_OF_RESOURCES                (-5)
#define CLK_MEM_OBJECT_ALLOCATION_FAILURE   (-4)
#define CLK_MEM_COPY_OVERLAP                (-8)
#define CLK_INVALID_VALUE                  (-30)
#define CLK_INVALID_MEM_OBJECT             (-38)
#define CLK_INVALID_KERNEL_NAME            (-46)
#define CLK_INVALID_KERNEL_ARGS            (-52)
#define CLK_INVALID_WORK_DIMENSION         (-53)
#define CLK_INVALID_OPERATION              (-59)

/* Image descriptor: replaces a cl_mem buffer plus the compile-time
 * GET_IMAGE_WIDTH/HEIGHT/DEPTH constants (reference README.md:75-87). */
typedef struct clk_image {
  void*   ptr;     /* device pointer to width*height*depth pixels */
  int32_t dtype;   /* clk_dtype */
  int32_t reserved;/* must be 0 */
  int64_t width;   /* x */
  int64_t height;  /* y */
  int64_t depth;   /* z */
} clk_image;

/* Largest Gaussian tap count N accepted by
 * clk_gaussian_blur_separable_3d (weights travel as kernel arguments). */
#define CLK_GAUSSIAN_MAX_TAPS 511

/* ---- the five kernels of the path ------------------------------------- */

/* gaussian_blur_separable_3d(dst, src, dim, N, s)
 * replaces kernels/gaussian_blur_separable_3d_x.cl:6-12.
 * 1D Gaussian along axis `dim` (0=x,1=y,2=z; any other value makes every tap
 * read the centre voxel, as the reference's dir vector does), c=(N-1)/2,
 * weights exp(v*v/(-2*s*s)), normalised by their sum. */
int32_t clk_gaussian_blur_separable_3d(void* stream, const clk_image* dst,
                                       const clk_image* src, int32_t dim,
                                       int32_t N, float s);

/* mean_separable_3d(dst, src, dim, N, s)
 * replaces kernels/mean_separable_3d_x.cl:3-10; s is accepted and ignored
 * as in the reference. */
int32_t clk_mean_separable_3d(void* stream, const clk_image* dst,
                              const clk_image* src, int32_t dim, int32_t N,
                              float s);

/* minimum_separable_3d(dst, src, dim, N, s)
 * replaces kernels/minimum_separable_3d_x.cl:3-10. */
int32_t clk_minimum_separable_3d(void* stream, const clk_image* dst,
                                 const clk_image* src, int32_t dim, int32_t N,
                                 float s);

/* maximum_separable_3d(dst, src, dim, N, s)
 * replaces kernels/maximum_separable_3d_x.cl:3-10. */
int32_t clk_maximum_separable_3d(void* stream, const clk_image* dst,
                                 const clk_image* src, int32_t dim, int32_t N,
                                 float s);

/* median_box_3d(dst, src, Nx, Ny, Nz)
 * replaces kernels/median_box_3d_x.cl:23-30.  The host define
 * MAX_ARRAY_SIZE (median_box_3d_x.cl:36) is not needed: any box works.
 * Float -> float medians keep the reference's insertion-sort result where
 * windows hold NaN or -0 (a NaN splits the sort, ties keep gather order),
 * as do the 2D, slice-by-slice and ellipsoid medians below. */
int32_t clk_median_box_3d(void* stream, const clk_image* dst,
                          const clk_image* src, int32_t Nx, int32_t Ny,
                          int32_t Nz);

/* ---- the 2D siblings (images of depth 1) -------------------------------- */

/* gaussian_blur_separable_2d / mean_separable_2d / minimum_separable_2d /
 * maximum_separable_2d (dst, src, dim, N, s) replace
 * kernels/gaussian_blur_separable_2d_x.cl:6-13, mean_separable_2d_x.cl:3-10,
 * minimum_separable_2d_x.cl:3-10, maximum_separable_2d_x.cl:3-10: the same
 * per-pixel arithmetic as the 3D kernels at depth 1.  Images must have
 * depth 1 (CLK_INVALID_VALUE otherwise). */
int32_t clk_gaussian_blur_separable_2d(void* stream, const clk_image* dst,
                                       const clk_image* src, int32_t dim,
                                       int32_t N, float s);
int32_t clk_mean_separable_2d(void* stream, const clk_image* dst,
                              const clk_image* src, int32_t dim, int32_t N,
                              float s);
int32_t clk_minimum_separable_2d(void* stream, const clk_image* dst,
                                 const clk_image* src, int32_t dim, int32_t N,
                                 float s);
int32_t clk_maximum_separable_2d(void* stream, const clk_image* dst,
                                 const clk_image* src, int32_t dim, int32_t N,
                                 float s);

/* median_box_2d(dst, src, Nx, Ny) replaces kernels/median_box_2d_x.cl:22-28.
 * Unlike median_box_3d it stores (dst_t)READ(src).x, a C cast of the widened
 * integer: an integer narrowed to a smaller integer dst keeps its low bits
 * (e.g. short -5 -> uchar 251) instead of saturating. */
int32_t clk_median_box_2d(void* stream, const clk_image* dst,
                          const clk_image* src, int32_t Nx, int32_t Ny);

/* ---- ellipsoid neighbourhoods, slice medians, converting copy ---------- */

/* minimum_sphere_3d / maximum_sphere_3d / mean_sphere_3d / median_sphere_3d
 * (dst, src, Nx, Ny, Nz) replace kernels/minimum_sphere_3d_x.cl:3-9,
 * maximum_sphere_3d_x.cl:3-9, mean_sphere_3d_x.cl:3-9 and
 * median_sphere_3d_x.cl:22-28: the box of half widths (N-1)/2 restricted to
 * x^2/a^2 + y^2/b^2 + z^2/c^2 <= 1 (FLT_MIN for a zero half width).  Half
 * widths up to 31 per axis; median_sphere_3d rejects a negative N. */
int32_t clk_minimum_sphere_3d(void* stream, const clk_image* dst,
                              const clk_image* src, int32_t Nx, int32_t Ny,
                              int32_t Nz);
int32_t clk_maximum_sphere_3d(void* stream, const clk_image* dst,
                              const clk_image* src, int32_t Nx, int32_t Ny,
                              int32_t Nz);
int32_t clk_mean_sphere_3d(void* stream, const clk_image* dst,
                           const clk_image* src, int32_t Nx, int32_t Ny,
                           int32_t Nz);
int32_t clk_median_sphere_3d(void* stream, const clk_image* dst,
                             const clk_image* src, int32_t Nx, int32_t Ny,
                             int32_t Nz);

/* median_slice_by_slice_box_3d(dst, src, Nx, Ny) replaces
 * kernels/median_slice_by_slice_box_3d_x.cl:22-28: median_box_2d on every z
 * slice (values stored by the same C cast). */
int32_t clk_median_slice_by_slice_box_3d(void* stream, const clk_image* dst,
                                         const clk_image* src, int32_t Nx,
                                         int32_t Ny);

/* copy_3d(dst, src) replaces kernels/copy_3d_x.cl:1-15: dst = CONVERT(src). */
int32_t clk_copy_3d(void* stream, const clk_image* dst, const clk_image* src);

/* ---- consumers usually chained after smoothing ------------------------ */

/* detect_maxima_3d(src, dst) replaces kernels/detect_maxima_3d_x.cl:2-5
 * (note the reference's parameter order, src first): dst = 1 where the first
 * strict maximum of the clamped 3x3x3 neighbourhood, in the loop order x, y,
 * z and starting from centre - 1, lies at the centre itself, else 0.  The
 * reference's debug printf from work-item (0,0,0) is not reproduced. */
int32_t clk_detect_maxima_3d(void* stream, const clk_image* src,
                             const clk_image* dst);

/* maximum_z_projection(dst_max, src) replaces
 * kernels/maximum_z_projection_x.cl:2-5: dst_max (width x height x 1) =
 * CONVERT(maximum over z), the first value then strictly greater ones. */
int32_t clk_maximum_z_projection(void* stream, const clk_image* dst_max,
                                 const clk_image* src);

/* ---- by-name dispatch (CLIJ execute(kernelName, parameterMap)) --------- */

typedef enum clk_arg_kind {
  CLK_ARG_IMAGE = 0,
  CLK_ARG_INT   = 1,
  CLK_ARG_FLOAT = 2
} clk_arg_kind;

typedef struct clk_arg {
  const char*      name;  /* parameter name as declared in the .cl kernel */
  int32_t          kind;  /* clk_arg_kind */
  int32_t          i;     /* CLK_ARG_INT */
  float            f;     /* CLK_ARG_FLOAT */
  int32_t          pad;
  const clk_image* image; /* CLK_ARG_IMAGE */
} clk_arg;

/* Launch a kernel by its reference name ("gaussian_blur_separable_3d", or the
 * program file name "gaussian_blur_separable_3d_x.cl"; the names are the five
 * 3D kernels, their _2d siblings, the four _sphere_3d filters,
 * median_slice_by_slice_box_3d, copy_3d, detect_maxima_3d and
 * maximum_z_projection) with arguments bound
 * by parameter name, as a CLIJ host binds them (reference README.md:26).
 * An int argument is accepted for a float parameter and vice versa, as the
 * reference hosts convert Java/Python numbers.  Unknown names or missing
 * parameters return CLK_INVALID_KERNEL_ARGS; unknown kernels
 * CLK_INVALID_KERNEL_NAME. */
int32_t clk_execute(void* stream, const char* kernel_name, const clk_arg* args,
                    int32_t nargs);

/* ---- host chain: full 3D filters (three single-axis passes) ------------- */

/* The reference leaves the 3-pass chain to its host.  These run the passes
 * x, y, z with float32 temporaries (the source type for minimum/maximum);
 * an axis with sigma <= 0 (Gaussian) or radius 0 is skipped, and with no
 * axis left the call is a converting copy.  N for a Gaussian axis is
 * int(8*sigma) rounded up to odd; for the box filters N = 2*radius+1.
 *
 * Temporaries are device blocks of the library's, kept per (device, stream)
 * and reused by the next call on that stream (grown when a call needs more)
 * until clk_release_temporaries(stream).  Calls on one stream from several
 * host threads are serialised while they issue their passes; calls on
 * different streams are not.  On a stream under graph capture, and on
 * hipStreamPerThread, each call allocates its temporaries in stream order
 * and frees them after its last pass (under capture the graph owns them).
 * Before destroying a stream of your own on which these calls ran, call
 * clk_release_temporaries for it (clk_stream_destroy does). */
int32_t clk_gaussian_blur_3d(void* stream, const clk_image* dst,
                             const clk_image* src, float sigma_x,
                             float sigma_y, float sigma_z);
int32_t clk_mean_box_3d(void* stream, const clk_image* dst,
                        const clk_image* src, int32_t rx, int32_t ry,
                        int32_t rz);
int32_t clk_minimum_box_3d(void* stream, const clk_image* dst,
                           const clk_image* src, int32_t rx, int32_t ry,
                           int32_t rz);
int32_t clk_maximum_box_3d(void* stream, const clk_image* dst,
                           const clk_image* src, int32_t rx, int32_t ry,
                           int32_t rz);

/* The same chains for 2D images (depth 1): passes x and y. */
int32_t clk_gaussian_blur_2d(void* stream, const clk_image* dst,
                             const clk_image* src, float sigma_x,
                             float sigma_y);
int32_t clk_mean_box_2d(void* stream, const clk_image* dst,
                        const clk_image* src, int32_t rx, int32_t ry);
int32_t clk_minimum_box_2d(void* stream, const clk_image* dst,
                           const clk_image* src, int32_t rx, int32_t ry);
int32_t clk_maximum_box_2d(void* stream, const clk_image* dst,
                           const clk_image* src, int32_t rx, int32_t ry);

/* ---- batches: many 3D filters in a few launches ------------------------ */

/* One 3D filter of a batch, with the semantics of the matching single call:
 *   CLK_JOB_GAUSSIAN_BLUR_3D  clk_gaussian_blur_3d(dst, src, sigma[0..2])
 *   CLK_JOB_MEAN_BOX_3D       clk_mean_box_3d(dst, src, radius[0..2])
 *   CLK_JOB_MINIMUM_BOX_3D    clk_minimum_box_3d(...)
 *   CLK_JOB_MAXIMUM_BOX_3D    clk_maximum_box_3d(...)
 *   CLK_JOB_MEDIAN_BOX_3D     clk_median_box_3d(dst, src, 2r+1 per axis)  */
enum {
  CLK_JOB_GAUSSIAN_BLUR_3D = 0,
  CLK_JOB_MEAN_BOX_3D = 1,
  CLK_JOB_MINIMUM_BOX_3D = 2,
  CLK_JOB_MAXIMUM_BOX_3D = 3,
  CLK_JOB_MEDIAN_BOX_3D = 4
};

typedef struct clk_filter_job {
  int32_t   kind;       /* CLK_JOB_* */
  int32_t   reserved;   /* must be 0 */
  clk_image dst;
  clk_image src;
  float     sigma[3];   /* Gaussian, per axis x,y,z */
  int32_t   radius[3];  /* box filters and median, per axis */
} clk_filter_job;

/* Run njobs independent 3D filters (volumes of any sizes and pixel types) as
 * one batch: the single-axis passes of all jobs run stage by stage, with one
 * kernel launch per kernel class and stage instead of one per volume and
 * axis.  Results equal those of the single calls bit for bit.  The batch's
 * temporaries and launch tables live in device blocks kept per (device,
 * stream), as the host chain's (see above).  Not for a stream under graph
 * capture (CLK_INVALID_OPERATION): the launch tables go up from reused host
 * staging memory.  No reference counterpart: the reference leaves launching
 * to its host. */
int32_t clk_batch_filter_3d(void* stream, const clk_filter_job* jobs, int32_t njobs);

/* Size of the fixed device array that holds the temporaries of
 * clk_filter_3d_noalloc (one per GPU the library runs on, held until the
 * process ends): 128 MiB. */
#define CLK_NOALLOC_SCRATCH_BYTES ((int64_t)128 << 20)

/* clk_filter_3d_noalloc(stream, job): the 3D filter of one clk_filter_job
 * (kinds as for clk_batch_filter_3d), with the results of the matching
 * clk_gaussian_blur_3d / clk_mean_box_3d / clk_minimum_box_3d /
 * clk_maximum_box_3d / clk_median_box_3d call, bit for bit, but calling no
 * allocator: the temporaries (two images of float32, or of the source type
 * for minimum/maximum, when three axes are filtered; one for two axes) live
 * in a fixed device array of CLK_NOALLOC_SCRATCH_BYTES per GPU, and no
 * launch shape is timed (tuned shapes chosen earlier are used).  A job whose
 * temporaries exceed the array is rejected with CLK_INVALID_VALUE (float32
 * temporaries of a 256^3 volume fit exactly).  The array's address on each
 * GPU is resolved on the first call there.  Every call on one GPU shares
 * that GPU's array and leaves its contents behind; nothing in it needs to
 * start at zero (each pass writes all of its output; a float -> float
 * median keeps one flag byte per 8x8x8 brick there, all rewritten first).  Calls on one stream
 * are ordered; calls on different streams of one GPU whose work overlaps in
 * time race on the array and must not be made.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). */
int32_t clk_filter_3d_noalloc(void* stream, const clk_filter_job* job);

/* Release the temporaries kept for `stream` on the current device: waits
 * for the stream, frees its blocks, and returns the library's unused pooled
 * device memory to the driver.  No reference counterpart. */
int32_t clk_release_temporaries(void* stream);

/* Bytes of device memory held for temporaries on the current device: kept
 * for `stream` (*stream_bytes) and by the library's pool as a whole, kept
 * blocks of every stream plus freed memory not yet returned to the driver
 * (*pool_bytes).  Either pointer may be NULL.  No reference counterpart. */
int32_t clk_temporaries_held(void* stream, int64_t* stream_bytes, int64_t* pool_bytes);

/* Launch-shape knobs; results never depend on them.  Keys: "strided_rounds"
 * (y/z blocks per resident slot), "x_sub" (LDS x pass segments per wave),
 * "batch_seg", "no_xreg", "strided_variant", "no_hist8", "xreg_maxc",
 * "ystatic_cols", "strided_order", "xreg_units", "x_dynamic", "x_tasks",
 * "no_xpack", "yring" (float y/z passes: 0 = lockstep ring on y passes,
 * 1 = never, 2 = always), "no_xtier" (1 = batched x passes keep the LDS row
 * classes), "no_xy" (1 = batched 3D filters keep separate x and y passes),
 * "batch_xyseg" (fused x+y batch pass: y outputs per block, 0 = 64),
 * "batch_serial" (1, the default: the classes of a batch stage run one
 * after another on the caller's stream; 0: side by side on side streams),
 * "batch_ycols" (batch y/z passes: at most 1, 2 or 4 columns per lane; 0 =
 * as the half-width tier allows; default 1), "no_xymm" (1 = batched 16-bit
 * minimum/maximum filters keep separate x and y passes), "no_xyg" (1, the
 * default: batched Gaussian and mean filters keep separate x and y passes;
 * 0: their register fused x+y class), "hist8_rows" (1 = 8-bit median
 * histograms kept per output instead of per window column), "no_xyzmm" (1,
 * the default: batched 16-bit minimum/maximum never run as one whole-3D
 * pass; 0: they do where it applies),
 * "batch_zseg" (whole-3D batch pass: planes per block, 0 = 64), "no_zmm"
 * (1, the default: batched 16-bit minimum/maximum z passes keep the y/z
 * ring classes; 0: they run the register min/max kernel over planes),
 * "ydyn_skip" (1 = batched Gaussian and mean y/z passes of chains from
 * integer images skip the taps outside each job's half width by branches,
 * as float-image jobs do; 0, the default: they run every tap of the tier
 * with weight 0 outside it, same results, no branches), "batch_ypad"
 * (batched y passes pad each row's columns to 16, 32 or 64 lanes, the
 * default 64: whole waves, so no wave spans the end of one plane's row and
 * the start of the next; 0: columns packed), "xymm_z" (1, the default: batched 16-bit
 * minimum/maximum fuse x with z, a wave's chunks in consecutive rows of
 * one plane, then run y; 0: x with y, then z; same results), "no_ymm"
 * (1, the default: batched 16-bit minimum/maximum y passes keep the y/z
 * ring classes; 0: they run the register min/max kernel over rows),
 * "yring_deep" (1 = the lockstep ring y/z pass takes steps of 4 rows (one
 * column per lane) or 2 and loads 8 or 4 steps ahead; 0, the default: 16, 8
 * or 4 rows, one step ahead), "sphere_direct" (ellipsoid filters: 0 = auto, 1 = per-voxel
 * gathers, 2 = LDS row tiles without z sliding), and
 * "autotune": 1 = the first single-axis launch of each signature (pixel
 * types, axis, N, volume size; volumes of 2^22 voxels and more) times a set
 * of launch shapes on the caller's stream, synchronising it once, and later
 * launches use the fastest.  Not for streams under graph capture (they are
 * left untuned). */
int32_t clk_set_tuning(const char* key, int32_t value);
int32_t clk_get_tuning(const char* key, int32_t* value);

/* Write the launch shapes the autotuner chose so far to a text file, or
 * merge such a file into this process's choices (for the current device),
 * so a later process launches the tuned shapes without timing them again. */
int32_t clk_autotune_save(const char* path);
int32_t clk_autotune_load(const char* path);

/* ---- device memory and streams (for hosts without their own HIP) ------- */

int32_t clk_init(void);
int32_t clk_malloc(void** ptr, int64_t bytes);
int32_t clk_free(void* ptr);
int32_t clk_memcpy_h2d(void* dst_dev, const void* src_host, int64_t bytes,
                       void* stream);
int32_t clk_memcpy_d2h(void* dst_host, const void* src_dev, int64_t bytes,
                       void* stream);
int32_t clk_stream_create(void** stream);
int32_t clk_stream_destroy(void* stream);
int32_t clk_stream_sync(void* stream);

/* Bytes per pixel of a clk_dtype, or 0 for an invalid value. */
int32_t clk_dtype_size(int32_t dtype);

/* Message of the last failing call on this thread ("" if none). */
const char* clk_last_error_string(void);

/* Library version as major*10000 + minor*100 + patch. */
int32_t clk_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CLKMI_H */
