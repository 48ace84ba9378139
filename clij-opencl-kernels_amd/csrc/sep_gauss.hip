// Instantiations of the single-axis separable pass for one op (all 25
// src/dst pixel-type pairs); see sep_impl.h.
#include "sep_impl.h"

namespace clk {
int32_t launch_sep_gauss(hipStream_t st, SepParams& p, int axis_is_x, int src_t, int dst_t,
                      std::string* err) {
  return launch_sep_op<OP_GAUSS>(st, p, axis_is_x, src_t, dst_t, err);
}
}  // namespace clk
