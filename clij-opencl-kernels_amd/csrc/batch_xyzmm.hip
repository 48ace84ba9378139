// Batch whole-3D minimum/maximum classes of 16-bit integer pixels.
#include <string>

#include "batch_xyzmm.h"

namespace clk {

namespace {

template <int OP, class TS>
bool xyzmm_tier(hipStream_t st, int tier, const MMParams* ps, const BatchGrid grid, uint32_t n) {
  if (tier == 5) hipLaunchKernelGGL((batch_xyzmm_kernel<OP, TS, 5>), dim3(n), dim3(kXYZThreads), 0, st, ps, grid);
  else return false;
  return true;
}

template <int OP>
bool xyzmm_src(hipStream_t st, int ts, int tier, const MMParams* ps, const BatchGrid grid, uint32_t n) {
  if (ts == CLK_USHORT) return xyzmm_tier<OP, uint16_t>(st, tier, ps, grid, n);
  if (ts == CLK_SHORT) return xyzmm_tier<OP, int16_t>(st, tier, ps, grid, n);
  return false;
}

}  // namespace

int32_t launch_xyzmm_class(hipStream_t st, int op, int ts, int tier, const MMParams* ps, const BatchGrid grid,
                           uint32_t nblocks, std::string* err) {
  bool ok = false;
  if (op == OP_MIN) ok = xyzmm_src<OP_MIN>(st, ts, tier, ps, grid, nblocks);
  else if (op == OP_MAX) ok = xyzmm_src<OP_MAX>(st, ts, tier, ps, grid, nblocks);
  if (!ok) {
    *err = "internal: no 3D minimum/maximum class for this job";
    return CLK_INVALID_VALUE;
  }
  return 0;
}

}  // namespace clk
