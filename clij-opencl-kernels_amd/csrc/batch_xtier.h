// Batch classes of the x pass with a run-time half width (sep_xtier.h): one
// launch per (op, pixel types, tier) and stage, whatever each job's row
// length and c.  One compilation unit per op instantiates these
// (batch_xtier_<op>.hip).
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
#include "sep_xtier.h"

namespace clk {

template <class T>
constexpr int dtype_code() {
  if constexpr (std::is_same<T, int8_t>::value) return CLK_CHAR;
  else if constexpr (std::is_same<T, uint8_t>::value) return CLK_UCHAR;
  else if constexpr (std::is_same<T, int16_t>::value) return CLK_SHORT;
  else if constexpr (std::is_same<T, uint16_t>::value) return CLK_USHORT;
  else return CLK_FLOAT;
}

template <int OP, class TS, class TD, int CMAX>
__global__ __launch_bounds__(256) void batch_xtier_kernel(const SepParams* __restrict__ ps,
                                                          const BatchGrid grid) {
  const BatchTile t = batch_tile(grid);
  sep_xtier_body<OP, TS, TD, CMAX, kXTierUnits>(batch_param(ps, grid, t.job), t.vblk);
}

template <int OP, class TS, class TD>
inline bool xtier_go(hipStream_t st, int tier, const SepParams* ps, const BatchGrid grid, uint32_t n) {
  switch (tier) {
    case 4: hipLaunchKernelGGL((batch_xtier_kernel<OP, TS, TD, 4>), dim3(n), dim3(256), 0, st, ps, grid); return true;
    case 8: hipLaunchKernelGGL((batch_xtier_kernel<OP, TS, TD, 8>), dim3(n), dim3(256), 0, st, ps, grid); return true;
    case 16: hipLaunchKernelGGL((batch_xtier_kernel<OP, TS, TD, 16>), dim3(n), dim3(256), 0, st, ps, grid); return true;
  }
  return false;
}

// (TS, float) for every TS and (TS, TS): xtier_pair()
template <int OP, class TS>
inline bool xtier_dst(hipStream_t st, int td, int tier, const SepParams* ps, const BatchGrid grid, uint32_t n) {
  if (td == CLK_FLOAT) return xtier_go<OP, TS, float>(st, tier, ps, grid, n);
  if constexpr (!std::is_same<TS, float>::value) {
    if (td == dtype_code<TS>()) return xtier_go<OP, TS, TS>(st, tier, ps, grid, n);
  }
  return false;
}

// Launch one x tier class of op OP; false when no kernel exists for it.
template <int OP>
bool launch_xtier_op(hipStream_t st, int ts, int td, int tier, const SepParams* ps, const BatchGrid grid,
                     uint32_t n) {
  switch (ts) {
    case CLK_CHAR: return xtier_dst<OP, int8_t>(st, td, tier, ps, grid, n);
    case CLK_UCHAR: return xtier_dst<OP, uint8_t>(st, td, tier, ps, grid, n);
    case CLK_SHORT: return xtier_dst<OP, int16_t>(st, td, tier, ps, grid, n);
    case CLK_USHORT: return xtier_dst<OP, uint16_t>(st, td, tier, ps, grid, n);
    case CLK_FLOAT: return xtier_dst<OP, float>(st, td, tier, ps, grid, n);
  }
  return false;
}

}  // namespace clk
