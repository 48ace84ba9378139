// Batch y/z tier classes of one op (split per op so the units compile in
// parallel).
#include "batch_ydyn.h"

namespace clk {

bool launch_ydyn_max(hipStream_t st, bool fin, int ts, int td, int tier, int cols, const SepParams* ps, const BatchGrid grid,
                    uint32_t n) {
  return launch_ydyn_op<OP_MAX>(st, fin, ts, td, tier, cols, ps, grid, n);
}

}  // namespace clk
