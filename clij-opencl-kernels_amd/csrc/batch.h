// Shared between the batch planner (batch.hip) and the batch class kernels
// of other compilation units (batch_ydyn.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "clk_internal.h"

namespace clk {

// One block of a batched class launch: the job (index into the class's
// parameter array) and the block's index within that job's own launch.
struct BatchTile {
  uint32_t job;
  uint32_t vblk;
};

// The blocks of a class launch: job j owns blocks off[j] .. off[j+1] - 1
// (off[0] = 0, off[njobs] = the grid), so the table the host uploads grows
// with the jobs, not with the blocks.
struct BatchGrid {
  const uint32_t* off;
  uint32_t njobs;
};

// Job and job-local block of this block: a binary search over off (the
// same for the whole block, so it stays on the scalar unit).
__device__ __forceinline__ BatchTile batch_tile(const BatchGrid g) {
  const uint32_t b = blockIdx.x;
  uint32_t lo = 0, hi = g.njobs;  // off[lo] <= b < off[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (g.off[mid] <= b) lo = mid;
    else hi = mid;
  }
  return {lo, b - g.off[lo]};
}

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

// x passes with a half width known at run time (sep_xtier.h): the tier
// (largest half width compiled: 4, 8 or 16; 0 = none fits), the pixel-type
// pairs compiled (to float, or to the same type), and the 64-chunk units
// each wave filters.
inline int xtier_tier(int c) { return c <= 4 ? 4 : (c <= 8 ? 8 : (c <= 16 ? 16 : 0)); }
inline bool xtier_pair(int ts, int td) { return td == CLK_FLOAT || ts == td; }
constexpr int kXTierUnits = 2;

// Launch one x tier class (nblocks 256-thread blocks); CLK_INVALID_VALUE
// when no kernel exists for it.
int32_t launch_xtier_class(hipStream_t st, int op, int ts, int td, int tier, const SepParams* ps,
                           const BatchGrid g, uint32_t nblocks, std::string* err);

// Launch one y/z tier class (a launch of nblocks 256-thread blocks);
// CLK_INVALID_VALUE when no kernel exists for the pair (the planner only
// forms classes that have one: ydyn_pair, and float sources for the
// Gaussian and the mean).
int32_t launch_ydyn_class(hipStream_t st, int op, int ts, int td, int tier, int cols, const SepParams* ps,
                          const BatchGrid g, uint32_t nblocks, std::string* err);

}  // namespace clk
