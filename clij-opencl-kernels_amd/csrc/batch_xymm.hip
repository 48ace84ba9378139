// Batch fused x+y minimum/maximum classes of 16-bit integer pixels.
#include <string>

#include "batch_xymm.h"

namespace clk {

namespace {

template <int OP, class TS>
bool xymm_tier(hipStream_t st, int tier, const XYParams* ps, const BatchGrid grid, uint32_t n) {
  if (tier == 5) hipLaunchKernelGGL((batch_xymm_kernel<OP, TS, 5>), dim3(n), dim3(256), 0, st, ps, grid);
  else if (tier == 8) hipLaunchKernelGGL((batch_xymm_kernel<OP, TS, 8>), dim3(n), dim3(256), 0, st, ps, grid);
  else return false;
  return true;
}

template <int OP>
bool xymm_src(hipStream_t st, int ts, int tier, const XYParams* ps, const BatchGrid grid, uint32_t n) {
  if (ts == CLK_USHORT) return xymm_tier<OP, uint16_t>(st, tier, ps, grid, n);
  if (ts == CLK_SHORT) return xymm_tier<OP, int16_t>(st, tier, ps, grid, n);
  return false;
}

}  // namespace

int32_t launch_xymm_class(hipStream_t st, int op, int ts, int tier, const XYParams* ps, const BatchGrid grid,
                          uint32_t nblocks, std::string* err) {
  bool ok = false;
  if (op == OP_MIN) ok = xymm_src<OP_MIN>(st, ts, tier, ps, grid, nblocks);
  else if (op == OP_MAX) ok = xymm_src<OP_MAX>(st, ts, tier, ps, grid, nblocks);
  if (!ok) {
    *err = "internal: no fused x+y minimum/maximum class for this job";
    return CLK_INVALID_VALUE;
  }
  return 0;
}

}  // namespace clk
