// Batch fused x+y classes of one op (split per op so the units compile in
// parallel).
#include "batch_xy.h"

namespace clk {

bool launch_xy_max(hipStream_t st, int ts, int tier, int cols, const XYParams* ps, const BatchGrid grid, uint32_t n,
                   size_t lds, int nslot) {
  return launch_xy_op<OP_MAX>(st, ts, tier, cols, ps, grid, n, lds, nslot);
}

}  // namespace clk
