// Host side of the sliding-window ellipsoid filters (sphere_slide_kernel in
// sphere_impl.h): planning and the typed launches.
#include <string.h>

#include <type_traits>

#include "clk_internal.h"
#include "clk_tuning.h"
#include "sphere_impl.h"

namespace clk {

namespace {

// the register network for medians of 7 or 33 values (the balls of N = 3
// and 5): 16-bit key pairs for an 8- or 16-bit dst, 32-bit keys for float
inline bool slide_net(int op, int n, int /*dst_t*/) { return op == SPH_MEDIAN && (n == 7 || n == 33); }

template <class TS, class TD, int TY>
void slide_op(int op, hipStream_t st, const SphereSlideParams& sp, dim3 grid, size_t shmem) {
#define CLK_SLIDE(OP, NET, PK) \
  hipLaunchKernelGGL((sphere_slide_kernel<TS, TD, OP, TY, NET, PK>), grid, dim3(256), shmem, st, sp)
  if constexpr (sizeof(TD) <= 2) {
    if (op == SPH_MEDIAN && sp.noff == 7) {
      CLK_SLIDE(SPH_MEDIAN, 7, 1);
      return;
    }
    if (op == SPH_MEDIAN && sp.noff == 33) {
      CLK_SLIDE(SPH_MEDIAN, 33, 1);
      return;
    }
  } else {
    if (op == SPH_MEDIAN && sp.noff == 7) {
      CLK_SLIDE(SPH_MEDIAN, 7, 0);
      return;
    }
    if (op == SPH_MEDIAN && sp.noff == 33) {
      CLK_SLIDE(SPH_MEDIAN, 33, 0);
      return;
    }
  }
  if constexpr (std::is_integral<TS>::value) {
    // 8/16-bit integer sources: packed row pairs for min/max, and for the
    // mean of bytes while the key sum fits 16 bits
    if (op == SPH_MIN) {
      CLK_SLIDE(SPH_MIN, 0, 1);
      return;
    }
    if (op == SPH_MAX) {
      CLK_SLIDE(SPH_MAX, 0, 1);
      return;
    }
    if constexpr (sizeof(TS) == 1) {
      if (op == SPH_MEAN && sp.noff <= 257) {
        CLK_SLIDE(SPH_MEAN, 0, 1);
        return;
      }
    }
  } else {
    if (op == SPH_MIN) {
      CLK_SLIDE(SPH_MIN, 0, 0);
      return;
    }
    if (op == SPH_MAX) {
      CLK_SLIDE(SPH_MAX, 0, 0);
      return;
    }
  }
  if (op == SPH_MEAN) {
    CLK_SLIDE(SPH_MEAN, 0, 0);
    return;
  }
  CLK_SLIDE(SPH_MEDIAN, 0, 0);
#undef CLK_SLIDE
}

template <class TS, class TD>
void slide_ty(int op, hipStream_t st, const SphereSlideParams& sp, int ty, dim3 grid, size_t shmem) {
  if (ty == 8) slide_op<TS, TD, 8>(op, st, sp, grid, shmem);
  else slide_op<TS, TD, 4>(op, st, sp, grid, shmem);
}

template <class TS>
void slide_dst(int op, hipStream_t st, const SphereSlideParams& sp, int dst_t, int ty, dim3 grid, size_t shmem) {
  switch (dst_t) {
    case CLK_CHAR: slide_ty<TS, int8_t>(op, st, sp, ty, grid, shmem); break;
    case CLK_UCHAR: slide_ty<TS, uint8_t>(op, st, sp, ty, grid, shmem); break;
    case CLK_SHORT: slide_ty<TS, int16_t>(op, st, sp, ty, grid, shmem); break;
    case CLK_USHORT: slide_ty<TS, uint16_t>(op, st, sp, ty, grid, shmem); break;
    default: slide_ty<TS, float>(op, st, sp, ty, grid, shmem); break;
  }
}

template <class TS, class TD>
void hist_nc(hipStream_t st, const SphereHistParams& hp, dim3 grid) {
  if (hp.ncol <= 16) hipLaunchKernelGGL((sphere_median_hist8_kernel<TS, TD, 16>), grid, dim3(64), 0, st, hp);
  else hipLaunchKernelGGL((sphere_median_hist8_kernel<TS, TD, 32>), grid, dim3(64), 0, st, hp);
}

template <class TD>
void hist_src(hipStream_t st, const SphereHistParams& hp, int src_t, dim3 grid) {
  switch (src_t) {
    case CLK_CHAR: hist_nc<int8_t, TD>(st, hp, grid); break;
    case CLK_UCHAR: hist_nc<uint8_t, TD>(st, hp, grid); break;
    case CLK_SHORT: hist_nc<int16_t, TD>(st, hp, grid); break;
    case CLK_USHORT: hist_nc<uint16_t, TD>(st, hp, grid); break;
    default: hist_nc<float, TD>(st, hp, grid); break;
  }
}

// median_sphere_3d with an 8-bit dst by the sliding byte histogram; false
// when it does not apply (other dst, fewer than 32 or more than 255 values,
// more than 32 columns, planes of 2^31 pixels or more, knobs).  Below 32
// values the median moves many bins per z step and the walk's serial LDS
// reads lose to the LDS ring's bisection (u8 512x512x256: 3x3x3 ball 4.97
// vs 0.68 ms, 5x5x5 3.10 vs 3.53 ms, 7x7x7 3.97 vs 15.96 ms).
bool launch_sphere_hist8(hipStream_t st, const SphereParams& p, int src_t, int dst_t) {
  if ((dst_t != CLK_CHAR && dst_t != CLK_UCHAR) || tuning(TUNE_SPHERE_DIRECT) != 0 || tuning(TUNE_NO_HIST8))
    return false;
  if (p.ex < 0 || p.ey < 0 || p.ez < 0 || p.n < 32 || p.n > 255 || p.H * p.W >= (int64_t(1) << 31)) return false;
  SphereHistParams hp;
  memset(&hp, 0, sizeof(hp));
  auto in = [&](int x, int y, int z) { return p.tx[x + p.ex] + p.ty[y + p.ey] + p.tz[z + p.ez] <= 1.0f; };
  int total = 0;
  for (int x = -p.ex; x <= p.ex; ++x)
    for (int y = -p.ey; y <= p.ey; ++y) {
      if (!in(x, y, 0)) {
        for (int z = -p.ez; z <= p.ez; ++z)
          if (in(x, y, z)) return false;  // not a column symmetric about z = 0
        continue;
      }
      int h = 0;
      while (h < p.ez && in(x, y, h + 1)) ++h;
      for (int z = -p.ez; z <= p.ez; ++z)
        if (in(x, y, z) != (z >= -h && z <= h)) return false;
      if (hp.ncol == kSphereHistMaxCols) return false;
      hp.col[hp.ncol++] = (uint32_t)(x + 128) | ((uint32_t)(y + 128) << 8) | ((uint32_t)h << 16);
      total += 2 * h + 1;
    }
  if (total != p.n) return false;
  hp.src = p.src;
  hp.dst = p.dst;
  hp.W = p.W;
  hp.H = p.H;
  hp.D = p.D;
  hp.n = p.n;
  const dim3 grid((uint32_t)((p.W + 63) / 64), (uint32_t)p.H, (uint32_t)((p.D + kHistZSeg - 1) / kHistZSeg));
  if (dst_t == CLK_CHAR) hist_src<int8_t>(st, hp, src_t, grid);
  else hist_src<uint8_t>(st, hp, src_t, grid);
  return true;
}

}  // namespace

bool launch_sphere_slide(int op, hipStream_t st, const SphereParams& p, int src_t, int dst_t) {
  if (tuning(TUNE_SPHERE_DIRECT) != 0) return false;
  if (op == SPH_MEDIAN && !slide_net(op, p.n, dst_t) && launch_sphere_hist8(st, p, src_t, dst_t)) return true;
  if (p.ex < 0 || p.ey < 0 || p.ez < 0 || 2 * p.ex > kSlideRL - 256 || p.ey > kSlideMaxEY || p.n == 0 ||
      p.n > kSphereTileMaxOff)
    return false;
  int ty = 0;
  size_t shmem = 0;
  const size_t cap = slide_net(op, p.n, dst_t) ? (size_t)kSlideMaxLDSNet : (size_t)kSlideMaxLDS;
  for (int c : {8, 4}) {
    const size_t b = (size_t)(2 * p.ez + 2) * (size_t)(c + 2 * p.ey) * kSlideRL * 4;
    if (b <= cap) {
      ty = c;
      shmem = b;
      break;
    }
  }
  if (ty == 0) return false;
  SphereSlideParams* spp = new (std::nothrow) SphereSlideParams;
  if (!spp) return false;
  SphereSlideParams& sp = *spp;
  memset(&sp, 0, sizeof(sp));
  sp.src = p.src;
  sp.dst = p.dst;
  sp.W = p.W;
  sp.H = p.H;
  sp.D = p.D;
  sp.ex = p.ex;
  sp.ey = p.ey;
  sp.ez = p.ez;
  for (int x = -p.ex; x <= p.ex; ++x)  // the reference's loop order
    for (int y = -p.ey; y <= p.ey; ++y)
      for (int z = -p.ez; z <= p.ez; ++z)
        if (p.tx[x + p.ex] + p.ty[y + p.ey] + p.tz[z + p.ez] <= 1.0f)
          sp.off[sp.noff++] = (uint32_t)((y + p.ey) * kSlideRL + x + p.ex) | ((uint32_t)(z + p.ez) << 16);
  for (int i = sp.noff; i % 8 != 0; ++i)  // the centre
    sp.off[i] = (uint32_t)(p.ey * kSlideRL + p.ex) | ((uint32_t)p.ez << 16);
  // about 2048 blocks, each sliding over at least 8 planes (or the depth)
  const int64_t gx = (p.W + 255) / 256, gy = (p.H + ty - 1) / ty;
  int64_t gz = (2048 + gx * gy - 1) / (gx * gy);
  if (gz > p.D) gz = p.D;
  int64_t zs = (p.D + gz - 1) / gz;
  const int64_t zmin = p.D < 8 ? p.D : 8;
  if (zs < zmin) zs = zmin;
  gz = (p.D + zs - 1) / zs;
  sp.zs = (int32_t)zs;
  const dim3 grid((uint32_t)gx, (uint32_t)gy, (uint32_t)gz);
  switch (src_t) {
    case CLK_CHAR: slide_dst<int8_t>(op, st, sp, dst_t, ty, grid, shmem); break;
    case CLK_UCHAR: slide_dst<uint8_t>(op, st, sp, dst_t, ty, grid, shmem); break;
    case CLK_SHORT: slide_dst<int16_t>(op, st, sp, dst_t, ty, grid, shmem); break;
    case CLK_USHORT: slide_dst<uint16_t>(op, st, sp, dst_t, ty, grid, shmem); break;
    default: slide_dst<float>(op, st, sp, dst_t, ty, grid, shmem); break;
  }
  delete spp;
  return true;
}

}  // namespace clk
