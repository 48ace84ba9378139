// Launch-shape knobs read by the launchers.  Defaults come from the
// environment (CLK_<KEY> in upper case) at first use and can be changed at
// run time through clk_set_tuning(), which is what the autotuner does.
#pragma once
#include <stdint.h>

namespace clk {

enum TuneKey : int {
  TUNE_STRIDED_ROUNDS = 0,   // y/z single launch: blocks per resident slot
  TUNE_X_SUB = 1,            // x LDS pass: 512-output segments per wave (0 = auto)
  TUNE_BATCH_SEG = 2,        // batched y/z: outputs per thread along the axis
  TUNE_NO_XREG = 3,          // 1 = never use the register-window x pass
  TUNE_STRIDED_VARIANT = 4,  // 0 = auto; >0 forces a generic y/z variant
  TUNE_NO_HIST8 = 5,         // 1 = never use the histogram median
  TUNE_COUNT = 6
};

int32_t tuning(TuneKey k);
const char* tuning_name(int k);

}  // namespace clk
