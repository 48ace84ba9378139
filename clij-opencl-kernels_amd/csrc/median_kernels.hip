// median_box_3d for gfx950.
//
// Reference: kernels/median_box_3d_x.cl:23-65.  Each output is the element
// of rank n/2 (0-based, ascending) of the n = (2ex+1)(2ey+1)(2ez+1)
// clamp-to-edge box values, each value taken as (dst_t)(float)src
// (median_box_3d_x.cl:54-55), written without CONVERT (:64).  The reference
// finds it with an insertion sort of a private array (:3-20); the value of
// that element does not depend on how it is found, so the kernels here use
// selection methods that fit registers and the VALU:
//
//  * 3x3x3 boxes: the 27 values sit in registers and a forgetful-selection
//    network (repeatedly discard the minimum and maximum of a k+2 element
//    buffer, then feed the next value) leaves the median; about 400 min/max
//    per voxel, no memory traffic beyond the gather.
//  * any other box: a bisection over the dst_t value range counts, per probe,
//    how many window values are <= the probe (8, 16 or 32 probes).  No array,
//    so no MAX_ARRAY_SIZE limit.
//
// Integer pixel types are exact.  For float dst the value of the rank-n/2
// element is reproduced; a zero median is returned as +0 and NaN inputs have
// no defined order (the reference's insertion sort is order-dependent there).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clk_device.h"
#include "clk_internal.h"

namespace clk {

// (dst_t)(float)READ(src): C conversion of the float to the dst pixel type.
// In range it truncates toward zero; out of range the reference's C
// conversion is undefined and this one saturates (NaN -> 0).
template <class TS, class TD>
__device__ __forceinline__ float gather_value(TS x) {
  return static_cast<float>(convert_sat<TD>(static_cast<float>(x)));
}

// Order-preserving integer key of a dst_t value (held in a float).
template <class TD>
__device__ __forceinline__ uint32_t value_key(float v) {
  if constexpr (Pix<TD>::kInt) {
    return static_cast<uint32_t>(static_cast<int32_t>(v) - static_cast<int32_t>(Pix<TD>::lo));
  } else {
    if (v == 0.f) v = 0.f;  // -0 and +0 are one value
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  }
}

template <class TD>
__device__ __forceinline__ float key_value(uint32_t k) {
  if constexpr (Pix<TD>::kInt) {
    return static_cast<float>(static_cast<int32_t>(k) + static_cast<int32_t>(Pix<TD>::lo));
  } else {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
  }
}

template <class TD> struct KeyBits { static constexpr int v = sizeof(TD) * 8; };
template <> struct KeyBits<float> { static constexpr int v = 32; };

struct MedParams {
  const void* src;
  void* dst;
  int64_t W, H, D;
  int32_t ex, ey, ez;
  int32_t n;
};

template <class TS, class TD>
__global__ __launch_bounds__(256) void median_generic_kernel(const MedParams p) {
  const int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t y = blockIdx.y;
  const int64_t z = blockIdx.z;
  if (x >= p.W) return;
  const TS* __restrict__ src = static_cast<const TS*>(p.src);
  const int64_t k = p.n / 2;  // rank of the result
  uint32_t lo = 0;
  uint32_t hi = (KeyBits<TD>::v == 32) ? 0xffffffffu : ((1u << KeyBits<TD>::v) - 1u);
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    int64_t cnt = 0;
    for (int dz = -p.ez; dz <= p.ez; ++dz) {
      const int64_t zz = clamp64(z + dz, 0, p.D - 1);
      for (int dy = -p.ey; dy <= p.ey; ++dy) {
        const int64_t yy = clamp64(y + dy, 0, p.H - 1);
        const TS* rowp = src + (zz * p.H + yy) * p.W;
        for (int dx = -p.ex; dx <= p.ex; ++dx) {
          const int64_t xx = clamp64(x + dx, 0, p.W - 1);
          cnt += value_key<TD>(gather_value<TS, TD>(rowp[xx])) <= mid;
        }
      }
    }
    if (cnt >= k + 1) hi = mid; else lo = mid + 1;
  }
  static_cast<TD*>(p.dst)[(z * p.H + y) * p.W + x] = static_cast<TD>(key_value<TD>(lo));
}

// Median (rank N/2) of N = 2k+1 values held in registers by forgetful
// selection: keep a buffer of k+2 values; its minimum and maximum cannot be
// the median of the values seen so far plus the rest, so move them to the
// buffer ends with compare-exchanges, drop both, and put the next unseen
// value into the freed slot.  After k-1 rounds the median of the last 3 is
// the answer.
template <int N>
__device__ __forceinline__ float forgetful_median(float (&a)[N]) {
  constexpr int k = N / 2;
  constexpr int m = k + 2;
  constexpr int hi = m - 1;
#pragma unroll
  for (int step = 0; step < N - m; ++step) {
    const int lo = step;
#pragma unroll
    for (int i = lo + 1; i <= hi; ++i) {
      const float u = a[lo], w = a[i];
      a[lo] = fminf(u, w);
      a[i] = fmaxf(u, w);
    }
#pragma unroll
    for (int i = lo + 1; i < hi; ++i) {
      const float u = a[i], w = a[hi];
      a[i] = fminf(u, w);
      a[hi] = fmaxf(u, w);
    }
    a[hi] = a[m + step];
  }
  const float x0 = a[N - m], x1 = a[N - m + 1], x2 = a[hi];  // the last 3 (N - m + 2 == hi)
  return fmaxf(fminf(x0, x1), fminf(fmaxf(x0, x1), x2));
}

template <class TS, class TD>
__global__ __launch_bounds__(256) void median_27_kernel(const MedParams p) {
  const int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t y = blockIdx.y;
  const int64_t z = blockIdx.z;
  if (x >= p.W) return;
  const TS* __restrict__ src = static_cast<const TS*>(p.src);
  float a[27];
  int i = 0;
#pragma unroll
  for (int dz = -1; dz <= 1; ++dz) {
    const int64_t zz = clamp64(z + dz, 0, p.D - 1);
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
      const int64_t yy = clamp64(y + dy, 0, p.H - 1);
      const TS* rowp = src + (zz * p.H + yy) * p.W;
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        const int64_t xx = clamp64(x + dx, 0, p.W - 1);
        a[i++] = gather_value<TS, TD>(rowp[xx]);
      }
    }
  }
  float m = forgetful_median<27>(a);
  if constexpr (!Pix<TD>::kInt) {
    if (m == 0.f) m = 0.f;
  }
  static_cast<TD*>(p.dst)[(z * p.H + y) * p.W + x] = static_cast<TD>(m);
}

template <class TS, class TD>
static int32_t median_typed(hipStream_t st, const MedParams& p, std::string* err) {
  if (p.H > 65535 || p.D > 65535) {
    *err = "median_box_3d: height and depth are limited to 65535";
    return CLK_INVALID_VALUE;
  }
  const dim3 grid((uint32_t)((p.W + 255) / 256), (uint32_t)p.H, (uint32_t)p.D);
  if (p.ex == 1 && p.ey == 1 && p.ez == 1)
    hipLaunchKernelGGL((median_27_kernel<TS, TD>), grid, dim3(256), 0, st, p);
  else
    hipLaunchKernelGGL((median_generic_kernel<TS, TD>), grid, dim3(256), 0, st, p);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    *err = std::string("kernel launch failed: ") + hipGetErrorString(e);
    return CLK_OUT_OF_RESOURCES;
  }
  return 0;
}

template <class TS>
static int32_t median_dst(hipStream_t st, const MedParams& p, int dst_t, std::string* err) {
  switch (dst_t) {
    case CLK_CHAR:   return median_typed<TS, int8_t>(st, p, err);
    case CLK_UCHAR:  return median_typed<TS, uint8_t>(st, p, err);
    case CLK_SHORT:  return median_typed<TS, int16_t>(st, p, err);
    case CLK_USHORT: return median_typed<TS, uint16_t>(st, p, err);
    case CLK_FLOAT:  return median_typed<TS, float>(st, p, err);
  }
  *err = "invalid dst pixel type";
  return CLK_INVALID_VALUE;
}

int32_t launch_median(hipStream_t st, const clk_image& dst, const clk_image& src, int32_t Nx,
                      int32_t Ny, int32_t Nz, std::string* err) {
  MedParams p;
  p.src = src.ptr;
  p.dst = dst.ptr;
  p.W = dst.width;
  p.H = dst.height;
  p.D = dst.depth;
  // median_box_3d_x.cl:39: e = (N-1)/2 per axis (C division)
  p.ex = (Nx - 1) / 2;
  p.ey = (Ny - 1) / 2;
  p.ez = (Nz - 1) / 2;
  if (p.ex < 0 || p.ey < 0 || p.ez < 0) {
    *err = "median_box_3d: Nx, Ny and Nz must be >= 0 (a negative size leaves the box empty)";
    return CLK_INVALID_VALUE;
  }
  const int64_t n = (int64_t)(2 * p.ex + 1) * (2 * p.ey + 1) * (2 * p.ez + 1);
  if (n > 0x7fffffff) {
    *err = "median_box_3d: box too large";
    return CLK_INVALID_VALUE;
  }
  p.n = (int32_t)n;
  if (p.W * p.H * p.D == 0) return 0;
  switch (src.dtype) {
    case CLK_CHAR:   return median_dst<int8_t>(st, p, dst.dtype, err);
    case CLK_UCHAR:  return median_dst<uint8_t>(st, p, dst.dtype, err);
    case CLK_SHORT:  return median_dst<int16_t>(st, p, dst.dtype, err);
    case CLK_USHORT: return median_dst<uint16_t>(st, p, dst.dtype, err);
    case CLK_FLOAT:  return median_dst<float>(st, p, dst.dtype, err);
  }
  *err = "invalid src pixel type";
  return CLK_INVALID_VALUE;
}

}  // namespace clk
