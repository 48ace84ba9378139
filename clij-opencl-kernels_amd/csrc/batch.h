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
  uint32_t pstride;  // bytes between the jobs' parameter entries
};

// Job j's entry of a class's parameter table (entries pstride bytes apart;
// an entry may end early where the class never reads further).
template <class P>
__device__ __forceinline__ const P& batch_param(const P* ps, const BatchGrid& g, uint32_t job) {
  return *reinterpret_cast<const P*>(reinterpret_cast<const char*>(ps) + (size_t)job * g.pstride);
}

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

// Fused x+y classes (batch_xy.h): the x pass computed in LDS on the way
// into the y pass, so the x temporary never goes to memory.
constexpr int kXYMaxC = 16;     // largest half width along x and y
constexpr int kXYK = 4;         // y outputs per step
constexpr int kXYSlots = 4;     // staged source pixels per thread and row (at most)

// One fused job (device table).  Plain data, filled by the batch planner.
struct XYParams {
  const void* src;
  void* dst;
  int32_t W, H, D;
  int32_t cx, cy;
  int32_t ng;       // column groups per row: ceil(W / COLS)
  int32_t ngt;      // D * ng
  int32_t rstr;     // staged words per plane row (xy_rstr)
  int32_t seg;      // y outputs per segment (multiple of kXYK)
  int32_t nseg;     // segments along y
  int32_t vec_st;   // W % COLS == 0 and dst aligned: COLS-wide stores
  float divx, rdivx, divy, rdivy;
  int32_t fastx, fasty;
  float wx[2 * kXYMaxC + 1];  // x weights in tap order t = 0..2cx
  float wy[2 * kXYMaxC + 1];  // y weights at CMAX - cy .. CMAX + cy (tier order)
  int64_t sy, sz;  // K_XYMM: elements between rows of the walk, and between
                   // its outer slices (x+y: W and H*W; a z pass: H*W and W)
};

// Staged words per plane row: the padded row rounded up to whole COLS-word
// vectors, so every thread's window starts COLS-aligned.
inline int32_t xy_rstr(int64_t ng, int cols, int cx) {
  return (int32_t)((ng * cols + 2 * cx + cols - 1) / cols * cols);
}

// Staged words (of one row) a block of a job touches at most: the planes a
// run of 256 groups can reach times the padded row, plus COLS - 1 words
// that the last window's final vector may read past it (never used).
inline int64_t xy_nslot(int64_t ng, int64_t D, int cols, int cx) {
  int64_t np = 255 / ng + 2;
  if (np > D) np = D;
  return np * xy_rstr(ng, cols, cx) + cols;
}

// LDS bytes of a fused class launch: K staged rows of nslot words, K x
// results of COLS words per thread.
inline size_t xy_lds_bytes(int64_t nslot, int cols) { return (size_t)4 * (kXYK * nslot + kXYK * cols * 256); }

// Launch one fused class (nblocks 256-thread blocks, lds bytes of dynamic
// LDS, nslot staged words per row); CLK_INVALID_VALUE when no kernel exists
// for it.
int32_t launch_xy_class(hipStream_t st, int op, int ts, int tier, int cols, const XYParams* ps,
                        const BatchGrid g, uint32_t nblocks, size_t lds, int nslot, std::string* err);

// Fused x+y minimum/maximum of 16-bit integer pixels in registers
// (batch_xymm.h): half widths up to kXYMMMaxC along x and y; y tiers 5, 8.
// ng = chunks of 8 pixels per row, ngt = D * ng.
constexpr int kXYMMMaxC = 8;
int32_t launch_xymm_class(hipStream_t st, int op, int ts, int tier, const XYParams* ps, const BatchGrid g,
                          uint32_t nblocks, std::string* err);

// Fused x+y Gaussian/mean with float temporaries, x in registers and y in
// registers, rows handed over through LDS (batch_xyg.h): half widths up to
// kXYMaxC; x tiers 8, 16, y tiers 5, 8, 16; ng = chunks of 8 pixels per row.
int32_t launch_xyg_class(hipStream_t st, int op, int ts, int xtier, int ytier, const XYParams* ps,
                         const BatchGrid g, uint32_t nblocks, std::string* err);

// A whole 3D minimum/maximum of 16-bit integer pixels in one launch
// (batch_xyzmm.h): half widths up to kXYMMMaxC along x and y and kXYZMaxCZ
// along z (one z tier).  A block (kXYZThreads threads) owns
// R = kXYZThreads / ng whole rows (ng = chunks of 8 pixels per row),
// TY = R - 2 cy of them outputs, and a segment of zseg planes; blocks of a
// job = nyt tiles x nzseg segments.
constexpr int kXYZThreads = 768;
constexpr int kXYZMaxCZ = 5;
struct MMParams {
  const void* src;
  void* dst;
  int32_t W, H, D;
  int32_t cx, cy, cz;
  int32_t ng, R, TY, nyt;
  int32_t zseg, nzseg;
};
int32_t launch_xyzmm_class(hipStream_t st, int op, int ts, int tier, const MMParams* ps, const BatchGrid g,
                           uint32_t nblocks, std::string* err);

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
int32_t launch_ydyn_class(hipStream_t st, int op, int ts, int td, int tier, int cols, bool fin, const SepParams* ps,
                          const BatchGrid g, uint32_t nblocks, std::string* err);

}  // namespace clk
