// Batch classes of the y/z pass with a run-time half width (sep_ydyn.h):
// one launch per (op, pixel types, tier, columns) and stage, whatever each
// job's c, so a batch of jobs with many different radii and sigmas needs a
// few launches per stage.  One compilation unit per op instantiates these
// (batch_ydyn_<op>.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <type_traits>

#include "../../include/clkmi.h"
#include "batch.h"
#include "clk_device.h"
#include "clk_internal.h"
#include "sep_impl.h"

namespace clk {

template <int OP, class TS, class TD, int CMAX, int COLS, int K, bool FIN>
__global__ __launch_bounds__(256) void batch_ydyn_kernel(const SepParams* __restrict__ ps,
                                                         const BatchGrid grid) {
  const BatchTile t = batch_tile(grid);
  sep_ydyn_body<OP, TS, TD, CMAX, COLS, K, FIN>(batch_param(ps, grid, t.job), t.vblk);
}


// fin: the branch-free body (Gaussian and mean of finite data only)
template <int OP, class TS, class TD, int CMAX, int COLS>
inline void go(hipStream_t st, bool fin, const SepParams* ps, const BatchGrid grid, uint32_t n) {
  constexpr int K = (COLS == 4 || (CMAX == 16 && COLS == 2)) ? 4 : 8;
  if constexpr (OP == OP_GAUSS || OP == OP_MEAN) {
    if (fin) {
      hipLaunchKernelGGL((batch_ydyn_kernel<OP, TS, TD, CMAX, COLS, K, true>), dim3(n), dim3(256), 0, st, ps, grid);
      return;
    }
  }
  hipLaunchKernelGGL((batch_ydyn_kernel<OP, TS, TD, CMAX, COLS, K, false>), dim3(n), dim3(256), 0, st, ps, grid);
}

template <int OP, class TS, class TD>
inline bool tier_cols(hipStream_t st, bool fin, int tier, int cols, const SepParams* ps, const BatchGrid grid, uint32_t n) {
  if (tier == 5) {
    if (cols == 4) go<OP, TS, TD, 5, 4>(st, fin, ps, grid, n);
    else if (cols == 2) go<OP, TS, TD, 5, 2>(st, fin, ps, grid, n);
    else go<OP, TS, TD, 5, 1>(st, fin, ps, grid, n);
  } else if (tier == 8) {
    if (cols == 2) go<OP, TS, TD, 8, 2>(st, fin, ps, grid, n);
    else go<OP, TS, TD, 8, 1>(st, fin, ps, grid, n);
  } else if (tier == 16) {
    if (cols == 2) go<OP, TS, TD, 16, 2>(st, fin, ps, grid, n);
    else go<OP, TS, TD, 16, 1>(st, fin, ps, grid, n);
  } else {
    return false;
  }
  return true;
}

template <int OP, class TS>
inline bool dst_type(hipStream_t st, bool fin, int td, int tier, int cols, const SepParams* ps, const BatchGrid grid, uint32_t n) {
  switch (td) {
    case CLK_CHAR:
      if constexpr (std::is_same<TS, int8_t>::value || std::is_same<TS, float>::value)
        return tier_cols<OP, TS, int8_t>(st, fin, tier, cols, ps, grid, n);
      break;
    case CLK_UCHAR:
      if constexpr (std::is_same<TS, uint8_t>::value || std::is_same<TS, float>::value)
        return tier_cols<OP, TS, uint8_t>(st, fin, tier, cols, ps, grid, n);
      break;
    case CLK_SHORT:
      if constexpr (std::is_same<TS, int16_t>::value || std::is_same<TS, float>::value)
        return tier_cols<OP, TS, int16_t>(st, fin, tier, cols, ps, grid, n);
      break;
    case CLK_USHORT:
      if constexpr (std::is_same<TS, uint16_t>::value || std::is_same<TS, float>::value)
        return tier_cols<OP, TS, uint16_t>(st, fin, tier, cols, ps, grid, n);
      break;
    case CLK_FLOAT:
      if constexpr (std::is_same<TS, float>::value) return tier_cols<OP, TS, float>(st, fin, tier, cols, ps, grid, n);
      break;
  }
  return false;
}

template <int OP>
inline bool src_type(hipStream_t st, bool fin, int ts, int td, int tier, int cols, const SepParams* ps, const BatchGrid grid,
              uint32_t n) {
  switch (ts) {
    case CLK_CHAR: return dst_type<OP, int8_t>(st, fin, td, tier, cols, ps, grid, n);
    case CLK_UCHAR: return dst_type<OP, uint8_t>(st, fin, td, tier, cols, ps, grid, n);
    case CLK_SHORT: return dst_type<OP, int16_t>(st, fin, td, tier, cols, ps, grid, n);
    case CLK_USHORT: return dst_type<OP, uint16_t>(st, fin, td, tier, cols, ps, grid, n);
    case CLK_FLOAT: return dst_type<OP, float>(st, fin, td, tier, cols, ps, grid, n);
  }
  return false;
}

// Launch one tier class of op OP; false when no kernel exists for it.
template <int OP>
bool launch_ydyn_op(hipStream_t st, bool fin, int ts, int td, int tier, int cols, const SepParams* ps, const BatchGrid grid,
                    uint32_t n) {
  // Gaussian and mean passes inside chains read float temporaries (only
  // those instantiations exist)
  if constexpr (OP == OP_GAUSS || OP == OP_MEAN) {
    if (ts != CLK_FLOAT) return false;
    return dst_type<OP, float>(st, fin, td, tier, cols, ps, grid, n);
  } else {
    return src_type<OP>(st, fin, ts, td, tier, cols, ps, grid, n);
  }
}

}  // namespace clk
