// Two consumers usually chained after smoothing: detect_maxima_3d
// (kernels/detect_maxima_3d_x.cl:2-51) and maximum_z_projection
// (kernels/maximum_z_projection_x.cl:2-20).
#include <string.h>

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

template <class TS, class TD>
__global__ __launch_bounds__(256) void max_z_projection_kernel(const TS* __restrict__ src, TD* __restrict__ dst,
                                                               int64_t W, int64_t H, int64_t D) {
  const int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t y = blockIdx.y;
  if (x >= W) return;
  float m = 0;  // :9
  for (int64_t z = 0; z < D; ++z) {
    const float v = static_cast<float>(src[(z * H + y) * W + x]);
    if (v > m || z == 0) m = v;  // :14
  }
  dst[y * W + x] = convert_sat<TD>(m);  // :18, CONVERT_dst_max
}

template <class TS, class TD>
void launch_detect(hipStream_t st, const clk_image& dst, const clk_image& src) {
  const dim3 grid((uint32_t)((src.width + 255) / 256), (uint32_t)src.height, (uint32_t)src.depth);
  hipLaunchKernelGGL((detect_maxima_kernel<TS, TD>), grid, dim3(256), 0, st, static_cast<const TS*>(src.ptr),
                     static_cast<TD*>(dst.ptr), src.width, src.height, src.depth);
}

template <class TS, class TD>
void launch_proj(hipStream_t st, const clk_image& dst, const clk_image& src) {
  const dim3 grid((uint32_t)((src.width + 255) / 256), (uint32_t)src.height, 1);
  hipLaunchKernelGGL((max_z_projection_kernel<TS, TD>), grid, dim3(256), 0, st, static_cast<const TS*>(src.ptr),
                     static_cast<TD*>(dst.ptr), src.width, src.height, src.depth);
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
