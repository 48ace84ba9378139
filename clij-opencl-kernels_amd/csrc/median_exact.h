// Float medians over windows that hold NaN or -0 (median_exact.hip).
//
// The reference sorts a window with a stable insertion sort on `<`
// (median_box_3d_x.cl:3-14): a NaN never moves and stops every later
// element, and -0 / +0 compare equal so they keep their gather order.  The
// sorted array is therefore the window split at its NaNs, each run sorted
// stably, the NaNs in place; the median is its element n/2.  The fast median
// kernels rank order-preserving keys (NaN above +inf, -0 as +0), which
// gives the same element whenever the window holds neither a NaN nor a -0.
// For float -> float medians (the only pairs where either can reach the
// sorted array) a scan flags the 8x8x8 source bricks holding a NaN or a -0,
// and a second kernel recomputes, with the rule above, every output brick
// whose window reaches a flagged brick; other bricks return at once.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/clkmi.h"

namespace clk {

// Flag bytes (one per 8x8x8 brick) a volume needs.
int64_t nanzero_flag_bytes(const clk_image& img);

// Scan + fix-up after a float -> float median over the box of half widths
// (ex, ey, ez) (sphere = false; ez = 0 for the 2D and slice medians) or the
// ellipsoid of median_sphere_3d (sphere = true); `flags` holds
// nanzero_flag_bytes(src) bytes.  Gather order x, y, z as the reference.
int32_t launch_median_nanzero(hipStream_t st, const clk_image& dst, const clk_image& src, int ex, int ey, int ez,
                              bool sphere, void* flags, std::string* err);

}  // namespace clk
