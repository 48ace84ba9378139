// Two consumers usually chained after smoothing: detect_maxima_3d
// (kernels/detect_maxima_3d_x.cl:2-51) and maximum_z_projection
// (kernels/maximum_z_projection_x.cl:2-20).
#include <string.h>

#include <type_traits>

#include "clk_device.h"
#include "clk_internal.h"

namespace clk {

namespace {

// detect_maxima_3d: the first strict maximum, in the loop order x, y, z over
// the 3x3x3 neighbourhood, of values read clamp-to-edge; it starts from
// centre - 1 and records the UNclamped position, so the result is 1 only
// when that first maximum is the centre itself (detect_maxima_3d_x.cl:20-50).
// A border voxel therefore meets its own value again at an outside position
// that comes first in loop order.  The reference's printf of the global size
// from work-item (0,0,0) (:15-17) is not reproduced.
template <class TS, class TD>
__global__ __launch_bounds__(256) void detect_maxima_kernel(const TS* __restrict__ src, TD* __restrict__ dst,
                                                            int64_t W, int64_t H, int64_t D) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t j = blockIdx.y, k = blockIdx.z;
  if (i >= W) return;
  float local_max = static_cast<float>(src[(k * H + j) * W + i]) - 1;  // :21
  int best = 13;  // offset index of the centre in x, y, z order
  int o = 0;
#pragma unroll
  for (int x = -1; x <= 1; ++x) {
    const int64_t xx = clamp64(i + x, 0, W - 1);
#pragma unroll
    for (int y = -1; y <= 1; ++y) {
      const int64_t yy = clamp64(j + y, 0, H - 1);
#pragma unroll
      for (int z = -1; z <= 1; ++z, ++o) {
        const int64_t zz = clamp64(k + z, 0, D - 1);
        const float v = static_cast<float>(src[(zz * H + yy) * W + xx]);  // :32
        if (v > local_max) {                                               // :35
          local_max = v;
          best = o;
        }
      }
    }
  }
  // unclamped positions differ exactly when the offsets differ; with no
  // update at all (e.g. a NaN centre) the position stays the centre's
  dst[(k * H + j) * W + i] = static_cast<TD>(best == 13 ? 1 : 0);  // :45-50
}

// Integer sources: the values and centre - 1 are exact floats and there is
// no NaN, so "the first strict maximum in loop order is the centre"
// (offset 13) holds exactly when every offset before it in loop order holds
// less than the centre and none after it holds more (an earlier value >= c
// would update first and block the centre's update; a later one > c would
// update after it).  Before the centre come the plane x-1, then (x, y-1, *),
// then (x, y, z-1).  A wave takes 64 consecutive x of one row y and slides
// along a run of z: each lane loads one value per row y-1, y, y+1 of a plane
// and takes its x neighbours from lanes l±1 by DPP (lanes 0 and 63 load the
// clamped halo); the next plane is loaded one step ahead.  Lanes past the
// row end act as its last pixel (the clamp) and store nothing.
constexpr int kDetectZSeg = 32;

template <class TS, class TD>
__global__ __launch_bounds__(64) void detect_maxima_int_kernel(const TS* __restrict__ src, TD* __restrict__ dst,
                                                               int64_t W, int64_t H, int64_t D) {
  const int lane = threadIdx.x;
  const int64_t x = (int64_t)blockIdx.x * 64 + lane;
  const int64_t y = blockIdx.y;
  const int64_t z0 = (int64_t)blockIdx.z * kDetectZSeg;
  const int64_t z1 = (z0 + kDetectZSeg < D) ? z0 + kDetectZSeg : D;
  const bool valid = x < W;
  const int64_t xi = valid ? x : W - 1;
  const bool edge = lane == 0 || lane == 63;
  const int64_t xh = clamp64(xi + (lane == 0 ? -1 : 1), 0, W - 1);  // halo column of lanes 0 and 63
  const int64_t PS = H * W;
  int64_t rows[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) rows[r] = clamp64(y + r - 1, 0, H - 1) * W;
  struct Raw {
    int own[3], halo[3];
  };
  struct Plane {
    int v[3][3];  // [dy + 1][dx + 1]
  };
  auto load = [&](int64_t zz, Raw& R) {
    const TS* pl = src + clamp64(zz, 0, D - 1) * PS;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      R.own[r] = (int)pl[rows[r] + xi];
      R.halo[r] = 0;
    }
    if (edge) {
#pragma unroll
      for (int r = 0; r < 3; ++r) R.halo[r] = (int)pl[rows[r] + xh];
    }
  };
  auto expand = [&](const Raw& R, Plane& P) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      P.v[r][1] = R.own[r];
      P.v[r][0] = (int)dpp_from_left((uint32_t)R.halo[r], (uint32_t)R.own[r]);
      P.v[r][2] = (int)dpp_from_right((uint32_t)R.halo[r], (uint32_t)R.own[r]);
    }
  };
  auto mx = [](int a, int b) { return a > b ? a : b; };
  Raw ra, rb, rc;
  Plane A, B, C;
  load(z0 - 1, ra);
  load(z0, rb);
  load(z0 + 1, rc);
  expand(ra, A);
  expand(rb, B);
  for (int64_t z = z0; z < z1; ++z) {
    Raw rn;
    load(z + 2, rn);  // clamped: in bounds after the last step too
    expand(rc, C);
    const int c = B.v[1][1];
    int a = mx(A.v[1][1], mx(A.v[0][1], mx(B.v[0][1], C.v[0][1])));
    int b = mx(C.v[1][1], mx(A.v[2][1], mx(B.v[2][1], C.v[2][1])));
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      a = mx(a, mx(A.v[r][0], mx(B.v[r][0], C.v[r][0])));
      b = mx(b, mx(A.v[r][2], mx(B.v[r][2], C.v[r][2])));
    }
    if (valid) dst[(z * H + y) * W + x] = static_cast<TD>((a < c && b <= c) ? 1 : 0);
    A = B;
    B = C;
    rc = rn;
  }
}

// maximum_z_projection of integer sources: the reference's running value
// (the first plane's, then each strictly greater one) is the plain maximum,
// which does not depend on order, so the depth is split: the four waves of a
// block take a quarter of the planes each for the same 256 x (4 pixels per
// lane, a vector load when the rows are 4-pixel aligned), and the partial
// maxima meet in LDS.
template <class TS, class TD>
__global__ __launch_bounds__(256) void max_z_projection_int_kernel(const TS* __restrict__ src, TD* __restrict__ dst,
                                                                   int64_t W, int64_t H, int64_t D, int32_t vec) {
  __shared__ int part[4][256];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t x0 = (int64_t)blockIdx.x * 256 + lane * 4;
  const int64_t y = blockIdx.y;
  const int64_t zc = (D + 3) / 4;
  const int64_t za = wv * zc, zb = (za + zc < D) ? za + zc : D;
  int m[4] = {INT32_MIN, INT32_MIN, INT32_MIN, INT32_MIN};
  const bool full = vec && x0 + 4 <= W;
  const int64_t plane = H * W;
  const TS* col = src + y * W + x0;
#pragma unroll 4
  for (int64_t z = za; z < zb; ++z) {
    const TS* p = col + z * plane;
    if (full) {
      const PixVec<TS, 4> v = load_nt<TS, 4>(p);
#pragma unroll
      for (int j = 0; j < 4; ++j) m[j] = m[j] > (int)v.v[j] ? m[j] : (int)v.v[j];
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (x0 + j < W) m[j] = m[j] > (int)p[j] ? m[j] : (int)p[j];
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) part[wv][lane * 4 + j] = m[j];
  __syncthreads();
  const int t = threadIdx.x;
  const int64_t x = (int64_t)blockIdx.x * 256 + t;
  if (x >= W) return;
  int r = part[0][t];  // the first quarter is never empty
#pragma unroll
  for (int w = 1; w < 4; ++w) r = r > part[w][t] ? r : part[w][t];
  dst[y * W + x] = convert_sat<TD>(static_cast<float>(r));  // :18, CONVERT_dst_max
}

// maximum_z_projection of float sources, depth split exactly.  The
// reference's chain (the first plane's value, then every strictly greater
// one) ends at the first occurrence of the largest non-NaN value, unless the
// first value is NaN, which nothing replaces.  Wave 0 runs that chain on the
// first quarter of the planes; waves 1..3 reduce their quarters to the first
// occurrence of the largest non-NaN value (NaN when there is none).  Combined
// in z order with the same strict comparison — a NaN running value stays, a
// NaN part never wins, an equal part leaves the earlier value (-0 / +0) —
// the parts give the chain's result bit for bit.
template <class TD>
__global__ __launch_bounds__(256) void max_z_projection_f32_kernel(const float* __restrict__ src, TD* __restrict__ dst,
                                                                   int64_t W, int64_t H, int64_t D, int32_t vec) {
  __shared__ float part[4][256];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t x0 = (int64_t)blockIdx.x * 256 + lane * 4;
  const int64_t y = blockIdx.y;
  const int64_t zc = (D + 3) / 4;
  const int64_t za = wv * zc, zb = (za + zc < D) ? za + zc : D;
  const float none = __builtin_nanf("");
  float m[4] = {none, none, none, none};
  const bool full = vec && x0 + 4 <= W;
  const int64_t plane = H * W;
  const float* col = src + y * W + x0;
  auto take = [&](int j, float v, bool first) {
    if (first) m[j] = v;  // :14, the first plane
    else if (v > m[j] || (m[j] != m[j] && wv != 0)) m[j] = v;  // wave 0: strict chain; others: skip "none"
  };
#pragma unroll 4
  for (int64_t z = za; z < zb; ++z) {
    const float* p = col + z * plane;
    const bool first = z == 0;
    if (full) {
      const PixVec<float, 4> v = load_nt<float, 4>(p);
#pragma unroll
      for (int j = 0; j < 4; ++j) take(j, v.v[j], first);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (x0 + j < W) take(j, p[j], first);
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) part[wv][lane * 4 + j] = m[j];
  __syncthreads();
  const int t = threadIdx.x;
  const int64_t x = (int64_t)blockIdx.x * 256 + t;
  if (x >= W) return;
  float r = part[0][t];
#pragma unroll
  for (int w = 1; w < 4; ++w)
    if (part[w][t] > r) r = part[w][t];
  dst[y * W + x] = convert_sat<TD>(r);  // :18, CONVERT_dst_max
}

template <class TS, class TD>
void launch_detect(hipStream_t st, const clk_image& dst, const clk_image& src) {
  if constexpr (std::is_integral<TS>::value) {
    const dim3 grid((uint32_t)((src.width + 63) / 64), (uint32_t)src.height,
                    (uint32_t)((src.depth + kDetectZSeg - 1) / kDetectZSeg));
    hipLaunchKernelGGL((detect_maxima_int_kernel<TS, TD>), grid, dim3(64), 0, st, static_cast<const TS*>(src.ptr),
                       static_cast<TD*>(dst.ptr), src.width, src.height, src.depth);
    return;
  }
  const dim3 grid((uint32_t)((src.width + 255) / 256), (uint32_t)src.height, (uint32_t)src.depth);
  hipLaunchKernelGGL((detect_maxima_kernel<TS, TD>), grid, dim3(256), 0, st, static_cast<const TS*>(src.ptr),
                     static_cast<TD*>(dst.ptr), src.width, src.height, src.depth);
}

template <class TS, class TD>
void launch_proj(hipStream_t st, const clk_image& dst, const clk_image& src) {
  const dim3 grid((uint32_t)((src.width + 255) / 256), (uint32_t)src.height, 1);
  if constexpr (std::is_integral<TS>::value) {
    const int32_t vec = src.width % 4 == 0 && (uintptr_t)src.ptr % (4 * sizeof(TS)) == 0;
    hipLaunchKernelGGL((max_z_projection_int_kernel<TS, TD>), grid, dim3(256), 0, st, static_cast<const TS*>(src.ptr),
                       static_cast<TD*>(dst.ptr), src.width, src.height, src.depth, vec);
    return;
  }
  const int32_t vec = src.width % 4 == 0 && (uintptr_t)src.ptr % 16 == 0;
  hipLaunchKernelGGL((max_z_projection_f32_kernel<TD>), grid, dim3(256), 0, st, static_cast<const float*>(src.ptr),
                     static_cast<TD*>(dst.ptr), src.width, src.height, src.depth, vec);
}

template <class TS>
void dispatch_dst(bool detect, hipStream_t st, const clk_image& dst, const clk_image& src) {
  switch (dst.dtype) {
    case CLK_CHAR: detect ? launch_detect<TS, int8_t>(st, dst, src) : launch_proj<TS, int8_t>(st, dst, src); break;
    case CLK_UCHAR: detect ? launch_detect<TS, uint8_t>(st, dst, src) : launch_proj<TS, uint8_t>(st, dst, src); break;
    case CLK_SHORT: detect ? launch_detect<TS, int16_t>(st, dst, src) : launch_proj<TS, int16_t>(st, dst, src); break;
    case CLK_USHORT:
      detect ? launch_detect<TS, uint16_t>(st, dst, src) : launch_proj<TS, uint16_t>(st, dst, src);
      break;
    default: detect ? launch_detect<TS, float>(st, dst, src) : launch_proj<TS, float>(st, dst, src); break;
  }
}

}  // namespace

int32_t launch_consumer(bool detect, hipStream_t st, const clk_image& dst, const clk_image& src, std::string* err) {
  if (src.width * src.height * (detect ? src.depth : 1) == 0) return 0;
  if (src.height > 65535 || (detect && src.depth > 65535)) {
    *err = "height and depth are limited to 65535";
    return CLK_INVALID_VALUE;
  }
  switch (src.dtype) {
    case CLK_CHAR: dispatch_dst<int8_t>(detect, st, dst, src); break;
    case CLK_UCHAR: dispatch_dst<uint8_t>(detect, st, dst, src); break;
    case CLK_SHORT: dispatch_dst<int16_t>(detect, st, dst, src); break;
    case CLK_USHORT: dispatch_dst<uint16_t>(detect, st, dst, src); break;
    default: dispatch_dst<float>(detect, st, dst, src); break;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    *err = std::string("kernel launch failed: ") + hipGetErrorString(e);
    return CLK_OUT_OF_RESOURCES;
  }
  return 0;
}

}  // namespace clk
