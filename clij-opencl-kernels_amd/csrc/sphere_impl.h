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

// LDS-tiled form for small neighbourhoods: a block filters 256 consecutive x
// of one (y, z).  It stages every source row (y+dy, z+dz) that holds a
// neighbourhood offset, x0-ex .. x0+255+ex clamped, once in LDS, then each
// thread walks the offset list (host-built, in the reference's x, y, z loop
// order) reading LDS.  Values are staged as floats (min/max/mean) or as
// order-preserving keys of the dst-typed value (median, by bisection).
constexpr int kSphereTileW = 256;
constexpr int kSphereTileMaxRows = 49;
constexpr int kSphereTileMaxOff = 400;

struct SphereTileParams {
  const void* src;
  void* dst;
  int64_t W, H, D;
  int32_t ex, nrows, noff, n;
  int8_t rdy[kSphereTileMaxRows], rdz[kSphereTileMaxRows];
  uint32_t off[kSphereTileMaxOff];  // (dx + 128) | row << 8
};

template <class TS, class TD, int OP>
__global__ __launch_bounds__(256) void sphere_tile_kernel(const SphereTileParams p) {
  extern __shared__ uint32_t tile[];
  const int t = threadIdx.x;
  const int64_t x0 = (int64_t)blockIdx.x * kSphereTileW;
  const int64_t y = blockIdx.y, z = blockIdx.z;
  const int rowlen = kSphereTileW + 2 * p.ex;
  const TS* __restrict__ src = static_cast<const TS*>(p.src);
  for (int r = 0; r < p.nrows; ++r) {
    const int64_t yy = clamp64(y + p.rdy[r], 0, p.H - 1), zz = clamp64(z + p.rdz[r], 0, p.D - 1);
    const TS* row = src + (zz * p.H + yy) * p.W;
    for (int e = t; e < rowlen; e += kSphereTileW) {
      const TS v = row[clamp64(x0 - p.ex + e, 0, p.W - 1)];
      uint32_t w;
      if constexpr (OP == SPH_MEDIAN) w = value_key<TD>(gather_value<TS, TD>(v, 0));
      else w = __float_as_uint(static_cast<float>(v));
      tile[r * rowlen + e] = w;
    }
  }
  __syncthreads();
  const int64_t x = x0 + t;
  if (x >= p.W) return;
  auto at = [&](uint32_t o) -> uint32_t {
    return tile[(int)(o >> 8) * rowlen + t + p.ex + ((int)(o & 0xff) - 128)];
  };
  TD* const out = static_cast<TD*>(p.dst) + (z * p.H + y) * p.W + x;
  if constexpr (OP == SPH_MEDIAN) {
    const int k = p.n / 2;
    uint32_t lo = 0;
    uint32_t hi = (KeyBits<TD>::v == 32) ? 0xffffffffu : ((1u << KeyBits<TD>::v) - 1u);
    while (lo < hi) {
      const uint32_t mid = lo + ((hi - lo) >> 1);
      int cnt = 0;
      for (int i = 0; i < p.noff; ++i) cnt += at(p.off[i]) <= mid;
      if (cnt >= k + 1) hi = mid; else lo = mid + 1;
    }
    *out = static_cast<TD>(key_value<TD>(lo));
  } else if constexpr (OP == SPH_MEAN) {
    float sum = 0.f;
    for (int i = 0; i < p.noff; ++i) sum = sum + __uint_as_float(at(p.off[i]));
    *out = convert_sat<TD>(sum / static_cast<float>(p.noff));
  } else {
    // the centre is an offset of every neighbourhood that is not empty; an
    // empty one (negative N) never takes this kernel
    float m = static_cast<float>(src[(z * p.H + y) * p.W + x]);
    for (int i = 0; i < p.noff; ++i) {
      const float v = __uint_as_float(at(p.off[i]));
      if (OP == SPH_MIN ? (v < m) : (v > m)) m = v;
    }
    *out = convert_sat<TD>(m);
  }
}

}  // namespace clk
