// Ellipsoid-neighbourhood filters: minimum/maximum/mean/median_sphere_3d
// (kernels/minimum_sphere_3d_x.cl, maximum_sphere_3d_x.cl,
// mean_sphere_3d_x.cl, median_sphere_3d_x.cl).
//
// The neighbourhood is the box of half widths e = (N-1)/2 restricted to
// x^2/a^2 + y^2/b^2 + z^2/c^2 <= 1 (a^2 = e.x^2, or FLT_MIN when e.x = 0,
// minimum_sphere_3d_x.cl:16-28,37).  The per-axis terms x^2/a^2 are
// computed on the host with the same float divisions and passed in, so the
// kernel only adds three floats per offset in the reference's order.  One
// thread per voxel; offsets in the reference's loop order x, y, z.
#pragma once

#include "clk_device.h"
#include "clk_median.h"
#include "median_impl.h"

namespace clk {

constexpr int kSphereMaxE = 31;  // half width per axis held in the parameter block

enum SphereOp : int { SPH_MIN = 0, SPH_MAX = 1, SPH_MEAN = 2, SPH_MEDIAN = 3 };

struct SphereParams {
  const void* src;
  void* dst;
  int64_t W, H, D;
  int32_t ex, ey, ez;
  int32_t n;  // offsets inside the ellipsoid
  float tx[2 * kSphereMaxE + 1], ty[2 * kSphereMaxE + 1], tz[2 * kSphereMaxE + 1];
};

__device__ __forceinline__ bool sphere_in(const SphereParams& p, int x, int y, int z) {
  return p.tx[x + p.ex] + p.ty[y + p.ey] + p.tz[z + p.ez] <= 1.0f;  // :37, sum in x, y, z order
}

template <class TS, class TD, int OP>
__global__ __launch_bounds__(256) void sphere_kernel(const SphereParams p) {
  const int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t y = blockIdx.y, z = blockIdx.z;
  if (x >= p.W) return;
  const TS* __restrict__ src = static_cast<const TS*>(p.src);
  auto rd = [&](int dx, int dy, int dz) -> float {
    const int64_t xx = clamp64(x + dx, 0, p.W - 1), yy = clamp64(y + dy, 0, p.H - 1),
                  zz = clamp64(z + dz, 0, p.D - 1);
    return static_cast<float>(src[(zz * p.H + yy) * p.W + xx]);
  };
  TD* const out = static_cast<TD*>(p.dst) + (z * p.H + y) * p.W + x;
  if constexpr (OP == SPH_MEDIAN) {
    // rank count/2 of the dst-typed values (median_sphere_3d_x.cl:54-68), by
    // bisection over order-preserving keys as median_box_3d
    const int64_t k = p.n / 2;
    uint32_t lo = 0;
    uint32_t hi = (KeyBits<TD>::v == 32) ? 0xffffffffu : ((1u << KeyBits<TD>::v) - 1u);
    while (lo < hi) {
      const uint32_t mid = lo + ((hi - lo) >> 1);
      int64_t cnt = 0;
      for (int dx = -p.ex; dx <= p.ex; ++dx)
        for (int dy = -p.ey; dy <= p.ey; ++dy)
          for (int dz = -p.ez; dz <= p.ez; ++dz)
            if (sphere_in(p, dx, dy, dz))
              cnt += value_key<TD>(gather_value<TS, TD>(static_cast<TS>(rd(dx, dy, dz)), 0)) <= mid;
      if (cnt >= k + 1) hi = mid; else lo = mid + 1;
    }
    *out = static_cast<TD>(key_value<TD>(lo));
  } else if constexpr (OP == SPH_MEAN) {
    float sum = 0.f;  // mean_sphere_3d_x.cl:17, sequential in loop order
    int count = 0;
    for (int dx = -p.ex; dx <= p.ex; ++dx)
      for (int dy = -p.ey; dy <= p.ey; ++dy)
        for (int dz = -p.ez; dz <= p.ez; ++dz)
          if (sphere_in(p, dx, dy, dz)) {
            sum = sum + rd(dx, dy, dz);  // :44
            ++count;
          }
    *out = convert_sat<TD>(sum / static_cast<float>(count));  // :50, IEEE divide
  } else {
    float m = rd(0, 0, 0);  // minimum_sphere_3d_x.cl:15: the centre first
    for (int dx = -p.ex; dx <= p.ex; ++dx)
      for (int dy = -p.ey; dy <= p.ey; ++dy)
        for (int dz = -p.ez; dz <= p.ez; ++dz)
          if (sphere_in(p, dx, dy, dz)) {
            const float v = rd(dx, dy, dz);
            if (OP == SPH_MIN ? (v < m) : (v > m)) m = v;  // :44-46
          }
    *out = convert_sat<TD>(m);
  }
}

}  // namespace clk
