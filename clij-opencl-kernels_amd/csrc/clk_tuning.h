// Launch-shape knobs read by the launchers.  Defaults come from the
// environment (CLK_<KEY> in upper case) at first use and can be changed at
// run time through clk_set_tuning(), which is what the autotuner does.
#pragma once
#include <stdint.h>

namespace clk {

enum TuneKey : int {
  TUNE_STRIDED_ROUNDS = 0,   // y/z single launch: blocks per resident slot (0 = 16)
  TUNE_X_SUB = 1,            // x LDS pass: 512-output segments per wave (0 = auto)
  TUNE_BATCH_SEG = 2,        // batched y/z: outputs per thread along the axis
  TUNE_NO_XREG = 3,          // 1 = never use the register-window x pass
  TUNE_STRIDED_VARIANT = 4,  // 0 = auto; >0 forces a generic y/z variant
  TUNE_NO_HIST8 = 5,         // 1 = never use the histogram median
  TUNE_XREG_MAXC = 6,        // 0 = auto; else the largest half width for the register-window x pass
  TUNE_YSTATIC_COLS = 7,     // 0 = auto; 1, 2 or 4 columns per lane in the static y/z pass
  TUNE_STRIDED_ORDER = 8,    // static y/z pass block order (bit 0: segment-minor, bit 1: XCD ranges)
  TUNE_AUTOTUNE = 9,         // 1 = time launch shapes at the first launch of each signature
  TUNE_XREG_UNITS = 10,      // 0 = auto; 1, 2 or 4 units of 512 outputs per wave in the register x pass
  TUNE_X_DYNAMIC = 11,       // 1 = LDS x pass always takes the half width at run time
  TUNE_X_TASKS = 12,         // 0/1 = one task per wave in the static LDS x pass, 2 = two (prefetched)
  TUNE_SPHERE_DIRECT = 13,   // ellipsoid filters: 0 = auto, 1 = per-voxel gathers, 2 = LDS row tiles (no z sliding)
  TUNE_NO_XPACK = 14,        // 1 = short rows take one row per wave in the LDS x pass
  TUNE_YRING = 15,           // float y/z passes: 0 = lockstep ring on y passes, 1 = never, 2 = always
  TUNE_NO_XTIER = 16,        // 1 = batched x passes never take the register-window tier classes
  TUNE_NO_XY = 17,           // 1 = batched 3D filters never fuse their x and y passes
  TUNE_BATCH_XYSEG = 18,     // fused x+y batch pass: y outputs per block (0 = 64)
  TUNE_BATCH_SERIAL = 19,    // 1 (default) = batch classes of a stage run one after another on the
                             // caller's stream; 0 = side by side on side streams
  TUNE_BATCH_YCOLS = 20,     // batch y/z tier classes: at most this many columns per lane (0 = tier's own; default 1)
  TUNE_NO_XYMM = 21,         // 1 = batched 16-bit minimum/maximum keep separate x and y passes
  TUNE_NO_XYG = 22,          // 1 (default) = batched Gaussian/mean keep separate x and y passes
  TUNE_HIST8_ROWS = 23,      // 1 = 8-bit median histograms per output lane (not per column)
  TUNE_NO_XYZMM = 24,        // 1 (default) = batched 16-bit minimum/maximum never run as one 3D pass
  TUNE_BATCH_ZSEG = 25,      // whole-3D batch pass: planes per block (0 = 64)
  TUNE_NO_ZMM = 26,          // 1 (default) = batched 16-bit minimum/maximum z passes keep the y/z ring classes
  TUNE_YDYN_SKIP = 27,       // 1 = batched Gaussian/mean y/z passes of finite data keep the tap-skipping body
  TUNE_BATCH_YPAD = 28,      // batched y passes pad each row's columns to 16, 32 or 64 lanes (0 = packed)
  TUNE_XYMM_Z = 29,          // 1 (default) = batched 16-bit min/max fuse x with z (then y) instead of x with y
  TUNE_NO_YMM = 30,          // 1 (default) = batched 16-bit min/max y passes keep the y/z ring classes
  TUNE_YRING_DEEP = 31,      // 1 = the lockstep ring y/z pass loads 4 or 8 steps ahead (smaller steps)
  TUNE_COUNT = 32
};

int32_t tuning(TuneKey k);
const char* tuning_name(int k);

// Overrides for the calling thread while in scope (the autotuner's
// candidates); other threads keep the process-wide values.
struct TuneSetting {
  TuneKey key;
  int32_t value;
};
class ScopedTuning {
 public:
  ScopedTuning(const TuneSetting* s, int n);
  ~ScopedTuning();
  ScopedTuning(const ScopedTuning&) = delete;
  ScopedTuning& operator=(const ScopedTuning&) = delete;

 private:
  int32_t saved_[TUNE_COUNT];
  bool saved_set_[TUNE_COUNT];
};

}  // namespace clk
