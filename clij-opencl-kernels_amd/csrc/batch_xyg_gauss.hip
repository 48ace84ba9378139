// Batch fused x+y classes of one op with float temporaries (split per op
// so the units compile in parallel).
#include "batch_xyg.h"

namespace clk {

bool launch_xyg_gauss(hipStream_t st, int ts, int xtier, int ytier, const XYParams* ps, const BatchGrid grid,
                   uint32_t n) {
  return launch_xyg_op<OP_GAUSS>(st, ts, xtier, ytier, ps, grid, n);
}

}  // namespace clk
