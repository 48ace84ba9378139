// Host side of the ellipsoid-neighbourhood filters (sphere_impl.h).
#include <float.h>
#include <string.h>

#include "clk_internal.h"
#include "clk_tuning.h"
#include "sphere_impl.h"

namespace clk {

namespace {

template <class TS, class TD>
int32_t sphere_typed(int op, hipStream_t st, const SphereParams& p, const SphereTileParams* tp) {
  const dim3 grid((uint32_t)((p.W + 255) / 256), (uint32_t)p.H, (uint32_t)p.D);
  if (tp) {
    const size_t shmem = (size_t)tp->nrows * (kSphereTileW + 2 * tp->ex) * 4;
    switch (op) {
      case SPH_MIN: hipLaunchKernelGGL((sphere_tile_kernel<TS, TD, SPH_MIN>), grid, dim3(256), shmem, st, *tp); break;
      case SPH_MAX: hipLaunchKernelGGL((sphere_tile_kernel<TS, TD, SPH_MAX>), grid, dim3(256), shmem, st, *tp); break;
      case SPH_MEAN: hipLaunchKernelGGL((sphere_tile_kernel<TS, TD, SPH_MEAN>), grid, dim3(256), shmem, st, *tp); break;
      default: hipLaunchKernelGGL((sphere_tile_kernel<TS, TD, SPH_MEDIAN>), grid, dim3(256), shmem, st, *tp); break;
    }
    return 0;
  }
  switch (op) {
    case SPH_MIN: hipLaunchKernelGGL((sphere_kernel<TS, TD, SPH_MIN>), grid, dim3(256), 0, st, p); break;
    case SPH_MAX: hipLaunchKernelGGL((sphere_kernel<TS, TD, SPH_MAX>), grid, dim3(256), 0, st, p); break;
    case SPH_MEAN: hipLaunchKernelGGL((sphere_kernel<TS, TD, SPH_MEAN>), grid, dim3(256), 0, st, p); break;
    default: hipLaunchKernelGGL((sphere_kernel<TS, TD, SPH_MEDIAN>), grid, dim3(256), 0, st, p); break;
  }
  return 0;
}

template <class TS>
int32_t sphere_dst(int op, hipStream_t st, const SphereParams& p, const SphereTileParams* tp, int dst_t) {
  switch (dst_t) {
    case CLK_CHAR: return sphere_typed<TS, int8_t>(op, st, p, tp);
    case CLK_UCHAR: return sphere_typed<TS, uint8_t>(op, st, p, tp);
    case CLK_SHORT: return sphere_typed<TS, int16_t>(op, st, p, tp);
    case CLK_USHORT: return sphere_typed<TS, uint16_t>(op, st, p, tp);
    default: return sphere_typed<TS, float>(op, st, p, tp);
  }
}

// The tiled kernel's rows and offsets, or false when the neighbourhood is
// empty or too large for it.
bool plan_tile(const SphereParams& p, SphereTileParams* tp) {
  if (p.ex < 0 || p.ey < 0 || p.ez < 0 || p.ex > 100 || p.n == 0 || p.n > kSphereTileMaxOff ||
      tuning(TUNE_SPHERE_DIRECT) == 1)
    return false;
  memset(tp, 0, sizeof(*tp));
  tp->src = p.src;
  tp->dst = p.dst;
  tp->W = p.W;
  tp->H = p.H;
  tp->D = p.D;
  tp->ex = p.ex;
  int row_of[2 * kSphereMaxE + 1][2 * kSphereMaxE + 1];
  for (int y = -p.ey; y <= p.ey; ++y)
    for (int z = -p.ez; z <= p.ez; ++z) {
      row_of[y + p.ey][z + p.ez] = -1;
      bool any = false;
      for (int x = -p.ex; x <= p.ex; ++x) any |= p.tx[x + p.ex] + p.ty[y + p.ey] + p.tz[z + p.ez] <= 1.0f;
      if (!any) continue;
      if (tp->nrows == kSphereTileMaxRows) return false;
      tp->rdy[tp->nrows] = (int8_t)y;
      tp->rdz[tp->nrows] = (int8_t)z;
      row_of[y + p.ey][z + p.ez] = tp->nrows++;
    }
  if ((size_t)tp->nrows * (kSphereTileW + 2 * p.ex) * 4 > 64 * 1024) return false;
  for (int x = -p.ex; x <= p.ex; ++x)  // the reference's loop order
    for (int y = -p.ey; y <= p.ey; ++y)
      for (int z = -p.ez; z <= p.ez; ++z)
        if (p.tx[x + p.ex] + p.ty[y + p.ey] + p.tz[z + p.ez] <= 1.0f)
          tp->off[tp->noff++] = (uint32_t)(x + 128) | ((uint32_t)row_of[y + p.ey][z + p.ez] << 8);
  tp->n = tp->noff;
  return true;
}

// x^2 / a^2 per offset of one axis with a^2 = e^2 as a float, FLT_MIN when
// e = 0 (minimum_sphere_3d_x.cl:16-28; x*x is an int converted to float).
void axis_terms(int e, float* t) {
  float a2 = (float)(e * e);
  if (a2 == 0) a2 = FLT_MIN;
  for (int x = -e; x <= e; ++x) t[x + e] = (float)(x * x) / a2;
}

}  // namespace

int32_t launch_sphere(int op, hipStream_t st, const clk_image& dst, const clk_image& src, int32_t Nx, int32_t Ny,
                      int32_t Nz, std::string* err) {
  SphereParams* pp = new (std::nothrow) SphereParams;
  if (!pp) {
    *err = "host allocation failed";
    return CLK_OUT_OF_RESOURCES;
  }
  SphereParams& p = *pp;
  memset(&p, 0, sizeof(p));
  p.src = src.ptr;
  p.dst = dst.ptr;
  p.W = dst.width;
  p.H = dst.height;
  p.D = dst.depth;
  p.ex = (Nx - 1) / 2;  // :14, C division
  p.ey = (Ny - 1) / 2;
  p.ez = (Nz - 1) / 2;
  int32_t rc = 0;
  if (p.ex > kSphereMaxE || p.ey > kSphereMaxE || p.ez > kSphereMaxE) {
    *err = "sphere filters take N <= " + std::to_string(2 * kSphereMaxE + 1) + " per axis";
    rc = CLK_INVALID_VALUE;
  } else if (op == SPH_MEDIAN && (p.ex < 0 || p.ey < 0 || p.ez < 0)) {
    *err = "median_sphere_3d: Nx, Ny and Nz must be >= 0 (a negative size leaves the neighbourhood empty)";
    rc = CLK_INVALID_VALUE;
  } else if (p.H > 65535 || p.D > 65535) {
    *err = "sphere filters: height and depth are limited to 65535";
    rc = CLK_INVALID_VALUE;
  }
  if (rc == 0 && p.W * p.H * p.D > 0) {
    if (p.ex >= 0) axis_terms(p.ex, p.tx);
    if (p.ey >= 0) axis_terms(p.ey, p.ty);
    if (p.ez >= 0) axis_terms(p.ez, p.tz);
    for (int x = -p.ex; x <= p.ex; ++x)
      for (int y = -p.ey; y <= p.ey; ++y)
        for (int z = -p.ez; z <= p.ez; ++z)
          if (p.tx[x + p.ex] + p.ty[y + p.ey] + p.tz[z + p.ez] <= 1.0f) ++p.n;
    // the sliding form where the neighbourhood fits it, else LDS row tiles,
    // else per-voxel gathers
    if (!launch_sphere_slide(op, st, p, src.dtype, dst.dtype)) {
      SphereTileParams* tp = new (std::nothrow) SphereTileParams;
      const SphereTileParams* use = (tp && plan_tile(p, tp)) ? tp : nullptr;
      switch (src.dtype) {
        case CLK_CHAR: rc = sphere_dst<int8_t>(op, st, p, use, dst.dtype); break;
        case CLK_UCHAR: rc = sphere_dst<uint8_t>(op, st, p, use, dst.dtype); break;
        case CLK_SHORT: rc = sphere_dst<int16_t>(op, st, p, use, dst.dtype); break;
        case CLK_USHORT: rc = sphere_dst<uint16_t>(op, st, p, use, dst.dtype); break;
        default: rc = sphere_dst<float>(op, st, p, use, dst.dtype); break;
      }
      delete tp;
    }
    const hipError_t e = hipGetLastError();
    if (rc == 0 && e != hipSuccess) {
      *err = std::string("kernel launch failed: ") + hipGetErrorString(e);
      rc = CLK_OUT_OF_RESOURCES;
    }
  }
  delete pp;
  return rc;
}

}  // namespace clk
