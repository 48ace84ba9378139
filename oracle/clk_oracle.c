/*
 * clk_oracle.c -- CPU restatement of the five 3D neighbourhood kernels of
 * clEsperanto/clij-opencl-kernels and their five 2D siblings, voxel by
 * voxel, in the reference's own operation order.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
 * The product library (libclkmi) never links, calls or falls back to it.
 *
 * Each function cites the reference lines it restates.  Choices the
 * reference leaves to parts that are not in it (the clij preamble and the
 * OpenCL runtime) are made here explicitly and documented in place:
 *   - READ_src_IMAGE with CLK_ADDRESS_CLAMP_TO_EDGE: each coordinate clamped
 *     to [0, size-1]; the pixel is widened exactly and taken as float.
 *   - CONVERT_dst_PIXEL_TYPE: float identity; integers saturate, truncate
 *     toward zero, NaN -> 0 (OpenCL 1.2 convert_<T>_sat, default rtz).
 *   - exp(): (float)exp((double)x), the correctly rounded value in all but
 *     vanishingly rare cases (OpenCL 1.2 allows 3 ulp).
 *   - res += h * x: contracted to one fused multiply-add, as OpenCL C's
 *     default FP_CONTRACT ON gives on every FMA-capable device.  This file is
 *     compiled with -ffp-contract=off so that fmaf() below is the only fusion.
 *   - OpenCL min(x,y) = y < x ? y : x; max(x,y) = x < y ? y : x.
 *
 * Pixel types: 0 char, 1 uchar, 2 short, 3 ushort, 4 float (clkmi.h).
 * Layout: offset(x,y,z) = x + y*W + z*W*H.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

enum { T_CHAR = 0, T_UCHAR = 1, T_SHORT = 2, T_USHORT = 3, T_FLOAT = 4 };

typedef struct {
  const void* p;
  int t;
  int64_t w, h, d;
} img_t;

static int64_t clampi(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* READ_src_IMAGE(src, sampler, pos).x cast to float, clamp-to-edge sampler
 * (gaussian_blur_separable_3d_x.cl:4,30). */
static float read_px(const img_t* im, int64_t x, int64_t y, int64_t z) {
  x = clampi(x, 0, im->w - 1);
  y = clampi(y, 0, im->h - 1);
  z = clampi(z, 0, im->d - 1);
  const int64_t i = x + y * im->w + z * im->w * im->h;
  switch (im->t) {
    case T_CHAR: return (float)((const int8_t*)im->p)[i];
    case T_UCHAR: return (float)((const uint8_t*)im->p)[i];
    case T_SHORT: return (float)((const int16_t*)im->p)[i];
    case T_USHORT: return (float)((const uint16_t*)im->p)[i];
    default: return ((const float*)im->p)[i];
  }
}

static float sat_rtz(float f, float lo, float hi) {
  if (f != f) return 0.f;
  if (f >= hi) return hi;
  if (f <= lo) return lo;
  return (float)(int32_t)f; /* C conversion truncates toward zero */
}

/* CONVERT_dst_PIXEL_TYPE (README.md:35-47), returned as the float holding
 * the converted value. */
float clko_convert(float f, int t) {
  switch (t) {
    case T_CHAR: return sat_rtz(f, -128.f, 127.f);
    case T_UCHAR: return sat_rtz(f, 0.f, 255.f);
    case T_SHORT: return sat_rtz(f, -32768.f, 32767.f);
    case T_USHORT: return sat_rtz(f, 0.f, 65535.f);
    default: return f;
  }
}

/* WRITE_dst_IMAGE(dst, coord, value) with an already converted value. */
static void write_px(void* p, int t, int64_t i, float v) {
  switch (t) {
    case T_CHAR: ((int8_t*)p)[i] = (int8_t)v; break;
    case T_UCHAR: ((uint8_t*)p)[i] = (uint8_t)v; break;
    case T_SHORT: ((int16_t*)p)[i] = (int16_t)v; break;
    case T_USHORT: ((uint16_t*)p)[i] = (uint16_t)v; break;
    default: ((float*)p)[i] = v; break;
  }
}

float clko_expf(float x) { return (float)exp((double)x); }

/* gaussian_blur_separable_3d_x.cl:14-34 for one work-item (i,j,k). */
static float gaussian_voxel(const img_t* src, int64_t i, int64_t j, int64_t k, int dim, int N, float s) {
  const int64_t di = (dim == 0), dj = (dim == 1), dk = (dim == 2); /* :19 */
  const int c = (N - 1) / 2;                                        /* :22 */
  const float n = -2 * s * s;                                       /* :24 */
  float res = 0;                                                    /* :26 */
  float hsum = 0;                                                   /* :27 */
  for (int v = -c; v <= c; v++) {                                   /* :28 */
    const float h = clko_expf((float)(v * v) / n);                  /* :29 */
    res = fmaf(h, read_px(src, i + v * di, j + v * dj, k + v * dk), res); /* :30 */
    hsum += h;                                                      /* :31 */
  }
  res /= hsum; /* :33 */
  return res;
}

/* mean_separable_3d_x.cl:12-29 for one work-item. */
static float mean_voxel(const img_t* src, int64_t i, int64_t j, int64_t k, int dim, int N) {
  const int64_t di = (dim == 0), dj = (dim == 1), dk = (dim == 2); /* :17 */
  const int c = (N - 1) / 2;                                        /* :20 */
  float res = 0;                                                    /* :22 */
  float count = 0;                                                  /* :23 */
  for (int v = -c; v <= c; v++) {                                   /* :24 */
    res = res + read_px(src, i + v * di, j + v * dj, k + v * dk);   /* :25 */
    count = count + 1;                                              /* :26 */
  }
  res /= count; /* :28 */
  return res;
}

/* minimum_separable_3d_x.cl:12-23 / maximum_separable_3d_x.cl:12-23. */
static float minmax_voxel(const img_t* src, int64_t i, int64_t j, int64_t k, int dim, int N, int is_max) {
  const int64_t di = (dim == 0), dj = (dim == 1), dk = (dim == 2); /* :14 */
  const int c = (N - 1) / 2;                                        /* :17 */
  float res = read_px(src, i, j, k);                                /* :19 */
  for (int v = -c; v <= c; v++) {                                   /* :20 */
    const float x = read_px(src, i + v * di, j + v * dj, k + v * dk);
    if (is_max) res = (res < x) ? x : res; /* :21 max(res, x) */
    else res = (x < res) ? x : res;        /* :21 min(res, x) */
  }
  return res;
}

/* op: 0 Gaussian, 1 mean, 2 minimum, 3 maximum.  NDRange = dst size; src
 * and dst must have the same size.  Returns 0, or -30 for bad arguments. */
int clko_separable(int op, void* dst, int dst_t, const void* src, int src_t, int64_t W, int64_t H,
                   int64_t D, int dim, int N, float s, int nthreads) {
  if (op < 0 || op > 3 || src_t < 0 || src_t > 4 || dst_t < 0 || dst_t > 4) return -30;
  const img_t im = {src, src_t, W, H, D};
  if (nthreads < 1) nthreads = 1;
  int64_t k, j;
#pragma omp parallel for collapse(2) num_threads(nthreads) schedule(static)
  for (k = 0; k < D; k++) {
    for (j = 0; j < H; j++) {
      for (int64_t i = 0; i < W; i++) {
        float r;
        switch (op) {
          case 0: r = gaussian_voxel(&im, i, j, k, dim, N, s); break;
          case 1: r = mean_voxel(&im, i, j, k, dim, N); break;
          case 2: r = minmax_voxel(&im, i, j, k, dim, N, 0); break;
          default: r = minmax_voxel(&im, i, j, k, dim, N, 1); break;
        }
        write_px(dst, dst_t, i + j * W + k * W * H, clko_convert(r, dst_t)); /* WRITE(CONVERT(res)) */
      }
    }
  }
  return 0;
}

/* (dst_t)(float)value: the implicit conversion of median_box_3d_x.cl:55.
 * Out of range it is undefined in the reference; here it saturates. */
static float to_dst_type(float f, int t) { return clko_convert(f, t); }

/* sort() and median() of median_box_3d_x.cl:3-20: stable insertion sort
 * with `<`, then element array_size/2. */
static float median_of(float* a, int n) {
  for (int i = 0; i < n; i++) {
    int jj;
    const float temp = a[i];
    for (jj = i - 1; jj >= 0 && temp < a[jj]; jj--) a[jj + 1] = a[jj];
    a[jj + 1] = temp;
  }
  return a[n / 2];
}

/* median_box_3d_x.cl:32-64.  Returns -30 when a negative N leaves the box
 * empty (the reference then reads an uninitialised array element). */
int clko_median_box_3d(void* dst, int dst_t, const void* src, int src_t, int64_t W, int64_t H, int64_t D,
                       int Nx, int Ny, int Nz, int nthreads) {
  if (src_t < 0 || src_t > 4 || dst_t < 0 || dst_t > 4) return -30;
  const int ex = (Nx - 1) / 2, ey = (Ny - 1) / 2, ez = (Nz - 1) / 2; /* :39 */
  if (ex < 0 || ey < 0 || ez < 0) return -30;
  const int n = (2 * ex + 1) * (2 * ey + 1) * (2 * ez + 1);
  const img_t im = {src, src_t, W, H, D};
  if (nthreads < 1) nthreads = 1;
  int rc = 0;
#pragma omp parallel num_threads(nthreads)
  {
    float* arr = (float*)malloc(sizeof(float) * (size_t)n);
    if (!arr) {
#pragma omp atomic write
      rc = -5;
    }
    int64_t k, j;
#pragma omp for collapse(2) schedule(static)
    for (k = 0; k < D; k++) {
      for (j = 0; j < H; j++) {
        if (!arr) continue;
        for (int64_t i = 0; i < W; i++) {
          int count = 0;
          for (int x = -ex; x <= ex; x++)       /* :47 */
            for (int y = -ey; y <= ey; y++)     /* :48 */
              for (int z = -ez; z <= ez; z++) { /* :49 */
                const float value_res = read_px(&im, i + x, j + y, k + z); /* :54 */
                arr[count++] = to_dst_type(value_res, dst_t);               /* :55 */
              }
          const float res = median_of(arr, count); /* :63 */
          write_px(dst, dst_t, i + j * W + k * W * H, res); /* :64, no CONVERT */
        }
      }
    }
    free(arr);
  }
  return rc;
}

/* ---- the 2D siblings: images of one plane, coordinates (i, j) ---------- */

/* gaussian_blur_separable_2d_x.cl:15-34 for one work-item (i, j). */
static float gaussian_pixel_2d(const img_t* src, int64_t i, int64_t j, int dim, int N, float s) {
  const int64_t di = (dim == 0), dj = (dim == 1); /* :19 */
  const int c = (N - 1) / 2;                       /* :22 */
  const float n = -2 * s * s;                      /* :24 */
  float res = 0, hsum = 0;                         /* :26-27 */
  for (int v = -c; v <= c; v++) {                  /* :28 */
    const float h = clko_expf((float)(v * v) / n); /* :29, (v*v)/n with v*v an int */
    res = fmaf(h, read_px(src, i + v * di, j + v * dj, 0), res); /* :30 */
    hsum += h;                                     /* :31 */
  }
  return res / hsum; /* :33 */
}

/* mean_separable_2d_x.cl:12-27. */
static float mean_pixel_2d(const img_t* src, int64_t i, int64_t j, int dim, int N) {
  const int64_t di = (dim == 0), dj = (dim == 1); /* :15 */
  const int c = (N - 1) / 2;                       /* :18 */
  float res = 0, count = 0;                        /* :20-21 */
  for (int v = -c; v <= c; v++) {                  /* :22 */
    res += read_px(src, i + v * di, j + v * dj, 0); /* :23 */
    count += 1;                                    /* :24 */
  }
  return res / count; /* :26 */
}

/* minimum_separable_2d_x.cl:12-25 / maximum_separable_2d_x.cl:12-25: unlike
 * the 3D kernels the centre is not applied a second time (v != 0, :21). */
static float minmax_pixel_2d(const img_t* src, int64_t i, int64_t j, int dim, int N, int is_max) {
  const int64_t di = (dim == 0), dj = (dim == 1); /* :14 */
  const int c = (N - 1) / 2;                       /* :17 */
  float res = read_px(src, i, j, 0);               /* :19 */
  for (int v = -c; v <= c; v++) {                  /* :20 */
    if (v != 0) {                                  /* :21 */
      const float x = read_px(src, i + v * di, j + v * dj, 0);
      if (is_max) res = (res < x) ? x : res;       /* :22 max */
      else res = (x < res) ? x : res;              /* :22 min */
    }
  }
  return res;
}

/* The four separable 2D kernels; images W x H (one plane). */
int clko_separable_2d(int op, void* dst, int dst_t, const void* src, int src_t, int64_t W, int64_t H, int dim,
                      int N, float s, int nthreads) {
  if (op < 0 || op > 3 || src_t < 0 || src_t > 4 || dst_t < 0 || dst_t > 4) return -30;
  const img_t im = {src, src_t, W, H, 1};
  if (nthreads < 1) nthreads = 1;
  int64_t j;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (j = 0; j < H; j++) {
    for (int64_t i = 0; i < W; i++) {
      float r;
      switch (op) {
        case 0: r = gaussian_pixel_2d(&im, i, j, dim, N, s); break;
        case 1: r = mean_pixel_2d(&im, i, j, dim, N); break;
        case 2: r = minmax_pixel_2d(&im, i, j, dim, N, 0); break;
        default: r = minmax_pixel_2d(&im, i, j, dim, N, 1); break;
      }
      write_px(dst, dst_t, i + j * W, clko_convert(r, dst_t)); /* WRITE(CONVERT(res)) */
    }
  }
  return 0;
}

/* (IMAGE_dst_PIXEL_TYPE)READ_src_IMAGE(...).x of median_box_2d_x.cl:45: a C
 * cast of the widened pixel (int for char/short, uint for uchar/ushort,
 * float).  An integer narrowed to a smaller integer type keeps its low bits
 * (two's complement, as C does on every target OpenCL runs on); a float
 * going to an integer type is undefined out of range in C and saturates here,
 * as in median_box_3d. */
static float cast_2d(float v, int src_t, int dst_t) {
  if (src_t != T_FLOAT && dst_t != T_FLOAT) {
    const int32_t x = (int32_t)v; /* exact: every <=16-bit integer */
    switch (dst_t) {
      case T_CHAR: return (float)(int8_t)(uint8_t)(uint32_t)x;
      case T_UCHAR: return (float)(uint8_t)(uint32_t)x;
      case T_SHORT: return (float)(int16_t)(uint16_t)(uint32_t)x;
      default: return (float)(uint16_t)(uint32_t)x;
    }
  }
  return to_dst_type(v, dst_t);
}

/* median_box_2d_x.cl:29-53 (sort/median :3-20 as in 3D).  Returns -30 when
 * a negative N leaves the box empty. */
int clko_median_box_2d(void* dst, int dst_t, const void* src, int src_t, int64_t W, int64_t H, int Nx, int Ny,
                       int nthreads) {
  if (src_t < 0 || src_t > 4 || dst_t < 0 || dst_t > 4) return -30;
  const int ex = (Nx - 1) / 2, ey = (Ny - 1) / 2; /* :36 */
  if (ex < 0 || ey < 0) return -30;
  const int n = (2 * ex + 1) * (2 * ey + 1);
  const img_t im = {src, src_t, W, H, 1};
  if (nthreads < 1) nthreads = 1;
  int rc = 0;
#pragma omp parallel num_threads(nthreads)
  {
    float* arr = (float*)malloc(sizeof(float) * (size_t)n);
    if (!arr) {
#pragma omp atomic write
      rc = -5;
    }
    int64_t j;
#pragma omp for schedule(static)
    for (j = 0; j < H; j++) {
      if (!arr) continue;
      for (int64_t i = 0; i < W; i++) {
        int count = 0;
        for (int x = -ex; x <= ex; x++)   /* :43 */
          for (int y = -ey; y <= ey; y++) /* :44 */
            arr[count++] = cast_2d(read_px(&im, i + x, j + y, 0), src_t, dst_t); /* :45 */
        write_px(dst, dst_t, i + j * W, median_of(arr, count)); /* :52-53 */
      }
    }
    free(arr);
  }
  return rc;
}

/* ---- ellipsoid neighbourhoods (…_sphere_3d_x.cl) ------------------------ */

/* aSquared etc. of minimum_sphere_3d_x.cl:16-28: e*e as an int converted to
 * float, FLT_MIN when zero. */
static float sq_or_min(int e) {
  float a = (float)(e * e);
  if (a == 0) a = 1.17549435e-38f; /* FLT_MIN */
  return a;
}

/* op: 0 minimum_sphere_3d_x.cl:11-53, 1 maximum_sphere_3d_x.cl (the same with
 * >), 2 mean_sphere_3d_x.cl:11-53, 3 median_sphere_3d_x.cl:29-71.  Returns
 * -30 for bad arguments, including an empty median neighbourhood. */
int clko_sphere_3d(int op, void* dst, int dst_t, const void* src, int src_t, int64_t W, int64_t H, int64_t D,
                   int Nx, int Ny, int Nz, int nthreads) {
  if (op < 0 || op > 3 || src_t < 0 || src_t > 4 || dst_t < 0 || dst_t > 4) return -30;
  const int ex = (Nx - 1) / 2, ey = (Ny - 1) / 2, ez = (Nz - 1) / 2; /* :14 */
  if (op == 3 && (ex < 0 || ey < 0 || ez < 0)) return -30;
  const float a2 = sq_or_min(ex), b2 = sq_or_min(ey), c2 = sq_or_min(ez);
  const int64_t cap = (ex >= 0 && ey >= 0 && ez >= 0) ? (int64_t)(2 * ex + 1) * (2 * ey + 1) * (2 * ez + 1) : 1;
  const img_t im = {src, src_t, W, H, D};
  if (nthreads < 1) nthreads = 1;
  int rc = 0;
#pragma omp parallel num_threads(nthreads)
  {
    float* arr = (float*)malloc(sizeof(float) * (size_t)cap);
    if (!arr) {
#pragma omp atomic write
      rc = -5;
    }
    int64_t k, j;
#pragma omp for collapse(2) schedule(static)
    for (k = 0; k < D; k++) {
      for (j = 0; j < H; j++) {
        if (!arr) continue;
        for (int64_t i = 0; i < W; i++) {
          float ext = read_px(&im, i, j, k); /* min/max: the centre first (:15) */
          float sum = 0;                     /* mean (:17) */
          int count = 0;
          for (int x = -ex; x <= ex; x++) {
            const float xs = (float)(x * x);
            for (int y = -ey; y <= ey; y++) {
              const float ys = (float)(y * y);
              for (int z = -ez; z <= ez; z++) {
                const float zs = (float)(z * z);
                if (xs / a2 + ys / b2 + zs / c2 <= 1.0f) { /* :37 */
                  const float v = read_px(&im, i + x, j + y, k + z);
                  if (op == 0) { if (v < ext) ext = v; }      /* :44-46 */
                  else if (op == 1) { if (v > ext) ext = v; }
                  else if (op == 2) sum = sum + v;            /* mean :44 */
                  else arr[count] = to_dst_type(v, dst_t);    /* median :65 */
                  count++;
                }
              }
            }
          }
          float r;
          if (op <= 1) r = clko_convert(ext, dst_t);
          else if (op == 2) r = clko_convert(sum / (float)count, dst_t); /* :50 */
          else r = median_of(arr, count);                                /* :70 */
          write_px(dst, dst_t, i + j * W + k * W * H, r);
        }
      }
    }
    free(arr);
  }
  return rc;
}

/* median_slice_by_slice_box_3d_x.cl:29-53: median_box_2d on every z slice,
 * with the same C cast of the values (:46). */
int clko_median_slice_by_slice_box_3d(void* dst, int dst_t, const void* src, int src_t, int64_t W, int64_t H,
                                      int64_t D, int Nx, int Ny, int nthreads) {
  const int64_t plane_in = W * H * (src_t == T_FLOAT ? 4 : (src_t >= T_SHORT ? 2 : 1));
  const int64_t plane_out = W * H * (dst_t == T_FLOAT ? 4 : (dst_t >= T_SHORT ? 2 : 1));
  for (int64_t k = 0; k < D; k++) {
    const int rc = clko_median_box_2d((char*)dst + k * plane_out, dst_t, (const char*)src + k * plane_in, src_t, W,
                                      H, Nx, Ny, nthreads);
    if (rc) return rc;
  }
  return 0;
}

/* copy_3d_x.cl:11-14: WRITE(CONVERT((float)READ(src))). */
int clko_copy_3d(void* dst, int dst_t, const void* src, int src_t, int64_t n) {
  if (src_t < 0 || src_t > 4 || dst_t < 0 || dst_t > 4) return -30;
  const img_t im = {src, src_t, n, 1, 1};
  for (int64_t i = 0; i < n; i++) write_px(dst, dst_t, i, clko_convert(read_px(&im, i, 0, 0), dst_t));
  return 0;
}

/* ---- consumers: detect_maxima_3d, maximum_z_projection ------------------ */

/* detect_maxima_3d_x.cl:9-50 (src first in the reference's declaration).
 * The position of the running maximum is the UNclamped neighbour position
 * (:27, :37); reads are clamped by the sampler.  The debug printf (:15-17)
 * is left out. */
int clko_detect_maxima_3d(const void* src, int src_t, void* dst, int dst_t, int64_t W, int64_t H, int64_t D) {
  if (src_t < 0 || src_t > 4 || dst_t < 0 || dst_t > 4) return -30;
  const img_t im = {src, src_t, W, H, D};
  const int radius = 1; /* :9 */
  for (int64_t k = 0; k < D; k++)
    for (int64_t j = 0; j < H; j++)
      for (int64_t i = 0; i < W; i++) {
        float local_max = read_px(&im, i, j, k) - 1;          /* :21 */
        int64_t mx = i, my = j, mz = k;                       /* :22 */
        for (int x = -radius; x < radius + 1; x++)           /* :24 */
          for (int y = -radius; y < radius + 1; y++)         /* :26 */
            for (int z = -radius; z < radius + 1; z++) {     /* :28 */
              const float value = read_px(&im, i + x, j + y, k + z); /* :32 */
              if (value > local_max) {                       /* :35 */
                local_max = value;
                mx = i + x; my = j + y; mz = k + z;          /* :37 */
              }
            }
        write_px(dst, dst_t, i + j * W + k * W * H, (mx == i && my == j && mz == k) ? 1.f : 0.f); /* :45-50 */
      }
  return 0;
}

/* maximum_z_projection_x.cl:7-19: dst_max is W x H. */
int clko_maximum_z_projection(void* dst, int dst_t, const void* src, int src_t, int64_t W, int64_t H, int64_t D) {
  if (src_t < 0 || src_t > 4 || dst_t < 0 || dst_t > 4) return -30;
  const img_t im = {src, src_t, W, H, D};
  for (int64_t y = 0; y < H; y++)
    for (int64_t x = 0; x < W; x++) {
      float m = 0;                                  /* :9 */
      for (int64_t z = 0; z < D; z++) {             /* :10 */
        const float value = read_px(&im, x, y, z); /* :13 */
        if (value > m || z == 0) m = value;         /* :14 */
      }
      write_px(dst, dst_t, x + y * W, clko_convert(m, dst_t)); /* :18 */
    }
  return 0;
}

/* Correctly rounded a/b through RN(1/b) and one FMA residual step, the
 * formula the GPU kernels use for a launch-uniform divisor; exported so the
 * tests can check it against the IEEE divide on many inputs. */
float clko_div_uniform(float a, float b, float r) {
  const float q = a * r;
  const float e = fmaf(-q, b, a);
  const float q1 = fmaf(e, r, q);
  const float aa = fabsf(a);
  if (!(aa >= 0x1p-96f && aa <= 0x1p+120f)) return a / b;
  return q1;
}

/* How many numerators a of the binade [1, 2) (all 2^23 floats) give
 * clko_div_uniform(a, b, RN(1/b)) != a / b.  Scaling a or b by a power of
 * two scales q, the residual and the result exactly (away from underflow
 * and overflow, which the range guard sends to the IEEE divide), so one
 * binade of a checks every guarded numerator for this divisor. */
int64_t clko_div_uniform_binade_mismatches(float b) {
  const float r = 1.0f / b;
  int64_t bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
  for (int32_t m = 0; m < (1 << 23); ++m) {
    union { uint32_t u; float f; } v;
    v.u = 0x3f800000u | (uint32_t)m;
    const float a = v.f;
    const float got = clko_div_uniform(a, b, r);
    const float want = a / b;
    if (memcmp(&got, &want, sizeof(float)) != 0) ++bad;
  }
  return bad;
}

int clko_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
