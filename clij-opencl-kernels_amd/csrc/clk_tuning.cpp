// Process-wide launch-shape knobs (see clk_tuning.h).
#include "clk_tuning.h"

#include <atomic>
#include <ctype.h>
#include <mutex>
#include <stdlib.h>
#include <string.h>

#include "../../include/clkmi.h"

namespace clk {
namespace {

const char* kNames[TUNE_COUNT] = {"strided_rounds", "x_sub", "batch_seg",
                                  "no_xreg", "strided_variant", "no_hist8",
                                  "xreg_maxc", "ystatic_cols", "strided_order",
                                  "autotune", "xreg_units", "x_dynamic", "x_tasks", "sphere_direct",
                                  "no_xpack", "yring", "no_xtier", "no_xy", "batch_xyseg",
                                  "batch_serial", "batch_ycols", "no_xymm", "no_xyg", "hist8_rows",
                                  "no_xyzmm", "batch_zseg", "no_zmm", "ydyn_skip", "batch_ypad", "xymm_z", "no_ymm",
                                  "yring_deep"};
const int32_t kDefaults[TUNE_COUNT] = {16, 0, 32, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 1, 1, 0, 1, 0, 1, 0, 1, 0, 64, 1, 1, 0};
std::atomic<int32_t> g_val[TUNE_COUNT];
std::once_flag g_once;

void init() {
  for (int k = 0; k < TUNE_COUNT; ++k) {
    char env[64] = "CLK_";
    for (int i = 0; kNames[k][i] && i < 58; ++i) env[4 + i] = (char)toupper((unsigned char)kNames[k][i]);
    const char* v = getenv(env);
    g_val[k].store((v && *v) ? (int32_t)atoi(v) : kDefaults[k]);
  }
}

int find(const char* key) {
  if (!key) return -1;
  for (int k = 0; k < TUNE_COUNT; ++k)
    if (strcmp(key, kNames[k]) == 0) return k;
  return -1;
}

}  // namespace

namespace {
thread_local int32_t t_over[TUNE_COUNT];
thread_local bool t_over_set[TUNE_COUNT];
}  // namespace

int32_t tuning(TuneKey k) {
  if (t_over_set[k]) return t_over[k];
  std::call_once(g_once, init);
  return g_val[k].load(std::memory_order_relaxed);
}

ScopedTuning::ScopedTuning(const TuneSetting* s, int n) {
  for (int k = 0; k < TUNE_COUNT; ++k) {
    saved_[k] = t_over[k];
    saved_set_[k] = t_over_set[k];
  }
  for (int i = 0; i < n; ++i) {
    t_over[s[i].key] = s[i].value;
    t_over_set[s[i].key] = true;
  }
}

ScopedTuning::~ScopedTuning() {
  for (int k = 0; k < TUNE_COUNT; ++k) {
    t_over[k] = saved_[k];
    t_over_set[k] = saved_set_[k];
  }
}

const char* tuning_name(int k) { return (k >= 0 && k < TUNE_COUNT) ? kNames[k] : nullptr; }

}  // namespace clk

extern "C" int32_t clk_set_tuning(const char* key, int32_t value) {
  std::call_once(clk::g_once, clk::init);
  const int k = clk::find(key);
  if (k < 0) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_BATCH_SEG && (value < 8 || value > (1 << 20))) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_STRIDED_ROUNDS && (value < 1 || value > 64)) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_XREG_MAXC && (value < 0 || value > 16)) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_YSTATIC_COLS && value != 0 && value != 1 && value != 2 && value != 4)
    return CLK_INVALID_VALUE;
  if (k == clk::TUNE_STRIDED_ORDER && (value < 0 || value > 3)) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_AUTOTUNE && value != 0 && value != 1) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_X_TASKS && (value < 0 || value > 2)) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_SPHERE_DIRECT && (value < 0 || value > 2)) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_YRING && (value < 0 || value > 2)) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_NO_XTIER && value != 0 && value != 1) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_NO_XY && value != 0 && value != 1) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_BATCH_SERIAL && value != 0 && value != 1) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_BATCH_YCOLS && value != 0 && value != 1 && value != 2 && value != 4) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_NO_XYMM && value != 0 && value != 1) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_NO_XYG && value != 0 && value != 1) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_HIST8_ROWS && value != 0 && value != 1) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_NO_XYZMM && value != 0 && value != 1) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_BATCH_ZSEG && (value < 0 || value > (1 << 20))) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_YDYN_SKIP && value != 0 && value != 1) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_BATCH_YPAD && value != 0 && value != 16 && value != 32 && value != 64) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_XYMM_Z && value != 0 && value != 1) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_NO_YMM && value != 0 && value != 1) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_YRING_DEEP && value != 0 && value != 1) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_NO_ZMM && value != 0 && value != 1) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_BATCH_XYSEG && (value < 0 || value > (1 << 20))) return CLK_INVALID_VALUE;
  if (k == clk::TUNE_XREG_UNITS && value != 0 && value != 1 && value != 2 && value != 4)
    return CLK_INVALID_VALUE;
  clk::g_val[k].store(value);
  return CLK_SUCCESS;
}

extern "C" int32_t clk_get_tuning(const char* key, int32_t* value) {
  std::call_once(clk::g_once, clk::init);
  const int k = clk::find(key);
  if (k < 0 || !value) return CLK_INVALID_VALUE;
  *value = clk::g_val[k].load();
  return CLK_SUCCESS;
}
