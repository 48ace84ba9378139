// Batch x tier classes of one op (split per op so the units compile in
// parallel).
#include "batch_xtier.h"

namespace clk {

bool launch_xtier_mean(hipStream_t st, int ts, int td, int tier, const SepParams* ps, const BatchGrid grid,
                      uint32_t n) {
  return launch_xtier_op<OP_MEAN>(st, ts, td, tier, ps, grid, n);
}

}  // namespace clk
