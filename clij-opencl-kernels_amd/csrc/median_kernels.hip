// median_box_3d / median_box_2d single launches; the kernels are in
// median_impl.h.
#include <string.h>

#include "median_impl.h"

namespace clk {

template <class TS, class TD>
static int32_t median_typed(hipStream_t st, MedParams& p, std::string* err) {
  if (p.H > 65535 || p.D > 65535) {
    *err = "median_box_3d: height and depth are limited to 65535";
    return CLK_INVALID_VALUE;
  }
  dim3 grid;
  int threads;
  switch ((int)plan_median(p, (int)sizeof(TD), &grid, &threads)) {
    case MED_PK8: {
      // single launches slide along z (runs of kPk8ZRun planes)
      const dim3 g(grid.x, grid.y, (uint32_t)((p.D + kPk8ZRun - 1) / kPk8ZRun));
      hipLaunchKernelGGL((median_27_pk8_slide_kernel<TS, TD>), g, dim3(threads), 0, st, p);
      break;
    }
    case MED_PK9: {
      // single launches slide along y (runs of kPk9YRun rows)
      const dim3 g(grid.x, (uint32_t)((p.H + kPk9YRun - 1) / kPk9YRun), grid.z);
      hipLaunchKernelGGL((median_9_pk8_slide_kernel<TS, TD>), g, dim3(threads), 0, st, p);
      break;
    }
    case MED_HIST8:
      // planned for 8-bit dst only
      if constexpr (sizeof(TD) == 1) {
#define CLK_HIST8_FIXED(EX, EY)                                                                    \
  if (p.ex == EX && p.ey == EY) {                                                                  \
    if (tuning(TUNE_HIST8_ROWS) == 0 && p.H * p.W * (int64_t)sizeof(TS) < (int64_t(1) << 31) &&    \
        p.D < (int64_t(1) << 31)) {                                                                \
      const dim3 g((uint32_t)((p.W + (64 - 2 * EX) - 1) / (64 - 2 * EX)), grid.y, grid.z);          \
      hipLaunchKernelGGL((median_hist8_cols_kernel<TS, TD, EX, EY>), g, dim3(threads), 0, st, p);   \
    } else {                                                                                       \
      hipLaunchKernelGGL((median_hist8_fixed_kernel<TS, TD, EX, EY>), grid, dim3(threads), 0, st, p); \
    }                                                                                              \
    break;                                                                                         \
  }
      CLK_HIST8_FIXED(1, 1) CLK_HIST8_FIXED(1, 2) CLK_HIST8_FIXED(1, 3)
      CLK_HIST8_FIXED(2, 1) CLK_HIST8_FIXED(2, 2) CLK_HIST8_FIXED(2, 3)
      CLK_HIST8_FIXED(3, 1) CLK_HIST8_FIXED(3, 2) CLK_HIST8_FIXED(3, 3)
#undef CLK_HIST8_FIXED
        hipLaunchKernelGGL((median_hist8_kernel<TS, TD>), grid, dim3(threads), 0, st, p);
      }
      break;
    case MED_27:
      // planned for float dst only (8/16-bit dst take the packed network);
      // single launches slide along z (runs of kF27ZRun planes)
      if constexpr (sizeof(TD) == 4) {
        const dim3 g(grid.x, grid.y, (uint32_t)((p.D + kF27ZRun - 1) / kF27ZRun));
        hipLaunchKernelGGL((median_27_key_slide_kernel<TS, TD>), g, dim3(threads), 0, st, p);
      } else {
        hipLaunchKernelGGL((median_27_kernel<TS, TD>), grid, dim3(threads), 0, st, p);
      }
      break;
#define CLK_MED_TIER(I)                                                                     \
  case MED_REGS + I:                                                                        \
    if constexpr (sizeof(TD) >= 2)                                                          \
      hipLaunchKernelGGL((median_regs_kernel<TS, TD, I>), grid, dim3(threads), 0, st, p); \
    break;
      CLK_MED_TIER(0) CLK_MED_TIER(1) CLK_MED_TIER(2) CLK_MED_TIER(3) CLK_MED_TIER(4) CLK_MED_TIER(5)
#undef CLK_MED_TIER
    default: hipLaunchKernelGGL((median_generic_kernel<TS, TD>), grid, dim3(threads), 0, st, p); break;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    *err = std::string("kernel launch failed: ") + hipGetErrorString(e);
    return CLK_OUT_OF_RESOURCES;
  }
  return 0;
}

template <class TS>
static int32_t median_dst(hipStream_t st, MedParams& p, int dst_t, std::string* err) {
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
                      int32_t Ny, int32_t Nz, std::string* err, bool wrap) {
  MedParams p;
  memset(&p, 0, sizeof(p));
  p.wrap = wrap ? 1 : 0;
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
