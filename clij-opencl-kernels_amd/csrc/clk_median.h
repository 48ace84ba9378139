// Parameters of one median_box_3d launch (shared by the single launch and
// the batched launcher).
#pragma once
#include <stdint.h>

namespace clk {

struct MedParams {
  const void* src;
  void* dst;
  int64_t W, H, D;
  int32_t ex, ey, ez;
  int32_t n;
  uint32_t gx, gy;  // grid of the chosen kernel in x and y (batched launches)
  int32_t wrap;     // 1: median_box_2d's C cast of integer values (see gather_value)
  int32_t reserved;
};

}  // namespace clk
