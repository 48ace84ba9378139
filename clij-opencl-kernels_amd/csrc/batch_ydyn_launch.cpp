// Dispatch of the batch tier classes (x, y/z and fused x+y) to the per-op compilation
// units.
#include <string>

#include "batch.h"
#include "clk_device.h"

namespace clk {

bool launch_ydyn_gauss(hipStream_t, bool, int, int, int, int, const SepParams*, const BatchGrid, uint32_t);
bool launch_ydyn_mean(hipStream_t, bool, int, int, int, int, const SepParams*, const BatchGrid, uint32_t);
bool launch_ydyn_min(hipStream_t, bool, int, int, int, int, const SepParams*, const BatchGrid, uint32_t);
bool launch_ydyn_max(hipStream_t, bool, int, int, int, int, const SepParams*, const BatchGrid, uint32_t);
bool launch_xtier_gauss(hipStream_t, int, int, int, const SepParams*, const BatchGrid, uint32_t);
bool launch_xtier_mean(hipStream_t, int, int, int, const SepParams*, const BatchGrid, uint32_t);
bool launch_xtier_min(hipStream_t, int, int, int, const SepParams*, const BatchGrid, uint32_t);
bool launch_xtier_max(hipStream_t, int, int, int, const SepParams*, const BatchGrid, uint32_t);
bool launch_xy_gauss(hipStream_t, int, int, int, const XYParams*, const BatchGrid, uint32_t, size_t, int);
bool launch_xyg_gauss(hipStream_t, int, int, int, const XYParams*, const BatchGrid, uint32_t);
bool launch_xyg_mean(hipStream_t, int, int, int, const XYParams*, const BatchGrid, uint32_t);
bool launch_xy_mean(hipStream_t, int, int, int, const XYParams*, const BatchGrid, uint32_t, size_t, int);
bool launch_xy_min(hipStream_t, int, int, int, const XYParams*, const BatchGrid, uint32_t, size_t, int);
bool launch_xy_max(hipStream_t, int, int, int, const XYParams*, const BatchGrid, uint32_t, size_t, int);

int32_t launch_xtier_class(hipStream_t st, int op, int ts, int td, int tier, const SepParams* ps,
                           const BatchGrid grid, uint32_t nblocks, std::string* err) {
  bool ok = false;
  switch (op) {
    case OP_GAUSS: ok = launch_xtier_gauss(st, ts, td, tier, ps, grid, nblocks); break;
    case OP_MEAN: ok = launch_xtier_mean(st, ts, td, tier, ps, grid, nblocks); break;
    case OP_MIN: ok = launch_xtier_min(st, ts, td, tier, ps, grid, nblocks); break;
    case OP_MAX: ok = launch_xtier_max(st, ts, td, tier, ps, grid, nblocks); break;
  }
  if (!ok) {
    *err = "internal: no x tier class for this pass";
    return CLK_INVALID_VALUE;
  }
  return 0;
}

int32_t launch_ydyn_class(hipStream_t st, int op, int ts, int td, int tier, int cols, bool fin, const SepParams* ps,
                          const BatchGrid grid, uint32_t nblocks, std::string* err) {
  bool ok = false;
  switch (op) {
    case OP_GAUSS: ok = launch_ydyn_gauss(st, fin, ts, td, tier, cols, ps, grid, nblocks); break;
    case OP_MEAN: ok = launch_ydyn_mean(st, fin, ts, td, tier, cols, ps, grid, nblocks); break;
    case OP_MIN: ok = launch_ydyn_min(st, fin, ts, td, tier, cols, ps, grid, nblocks); break;
    case OP_MAX: ok = launch_ydyn_max(st, fin, ts, td, tier, cols, ps, grid, nblocks); break;
  }
  if (!ok) {
    *err = "internal: no y/z tier class for this pass";
    return CLK_INVALID_VALUE;
  }
  return 0;
}

int32_t launch_xy_class(hipStream_t st, int op, int ts, int tier, int cols, const XYParams* ps,
                        const BatchGrid grid, uint32_t nblocks, size_t lds, int nslot, std::string* err) {
  bool ok = false;
  switch (op) {
    case OP_GAUSS: ok = launch_xy_gauss(st, ts, tier, cols, ps, grid, nblocks, lds, nslot); break;
    case OP_MEAN: ok = launch_xy_mean(st, ts, tier, cols, ps, grid, nblocks, lds, nslot); break;
    case OP_MIN: ok = launch_xy_min(st, ts, tier, cols, ps, grid, nblocks, lds, nslot); break;
    case OP_MAX: ok = launch_xy_max(st, ts, tier, cols, ps, grid, nblocks, lds, nslot); break;
  }
  if (!ok) {
    *err = "internal: no fused x+y class for this job";
    return CLK_INVALID_VALUE;
  }
  return 0;
}

int32_t launch_xyg_class(hipStream_t st, int op, int ts, int xtier, int ytier, const XYParams* ps,
                         const BatchGrid grid, uint32_t nblocks, std::string* err) {
  bool ok = false;
  if (op == OP_GAUSS) ok = launch_xyg_gauss(st, ts, xtier, ytier, ps, grid, nblocks);
  else if (op == OP_MEAN) ok = launch_xyg_mean(st, ts, xtier, ytier, ps, grid, nblocks);
  if (!ok) {
    *err = "internal: no fused x+y Gaussian/mean class for this job";
    return CLK_INVALID_VALUE;
  }
  return 0;
}

}  // namespace clk
