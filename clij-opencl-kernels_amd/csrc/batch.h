// Shared between the batch planner (batch.hip) and the batch class kernels
// of other compilation units (batch_ydyn.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clk_internal.h"

namespace clk {

// One block of a batched class launch: the job (index into the class's
// parameter array) and the block's index within that job's own launch.
struct BatchTile {
  uint32_t job;
  uint32_t vblk;
};

// y/z passes with a half width known at run time (sep_ydyn.h): the tier
// (largest half width compiled: 5, 8 or 16; 0 = none fits), the columns
// per lane the tier allows, and the outputs per step.
inline int ydyn_tier(int c) { return c <= 5 ? 5 : (c <= 8 ? 8 : (c <= 16 ? 16 : 0)); }
inline int ydyn_cols(int tier, bool v4, bool v2) { return (v4 && tier == 5) ? 4 : (v2 ? 2 : 1); }
inline int ydyn_k(int tier, int cols) { return (cols == 4 || (tier == 16 && cols == 2)) ? 4 : 8; }
// Pixel-type pairs with y/z tier classes: same type, and float to an
// integer type (the last pass of a chain with float temporaries).
inline bool ydyn_pair(int op, int ts, int td) {
  if (op == 0 || op == 1) return ts == CLK_FLOAT;  // Gaussian, mean: float temporaries
  return ts == td || ts == CLK_FLOAT;
}

// Launch one y/z tier class (a launch of ntiles 256-thread blocks);
// CLK_INVALID_VALUE when no kernel exists for the pair (the planner only
// forms classes that have one: ydyn_pair, and float sources for the
// Gaussian and the mean).
int32_t launch_ydyn_class(hipStream_t st, int op, int ts, int td, int tier, int cols, const SepParams* ps,
                          const BatchTile* tiles, uint32_t ntiles, std::string* err);

}  // namespace clk
