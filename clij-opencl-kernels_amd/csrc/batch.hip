// Batched execution of many 3D filters (clk_batch_filter_3d).
//
// A 3D separable filter is up to three single-axis passes; a median is one
// launch.  The planner turns every job into its passes (same temporaries
// and pass order as the single calls), then runs "stage" 0, 1, 2 -- the
// first, second and third active pass of every job.  Within a stage the
// passes are grouped by kernel class (kernel kind, op, pixel types, column
// width); each class is ONE launch whose blocks are numbered job after job
// (a device table holds each job's first block; a block finds its job by
// binary search), and each block runs the same kernel body as the single
// launch with that job's parameters.  So a batch of 256
// uneven volumes costs a few launches per stage instead of ~768, and the
// results equal the single calls bit for bit.
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/clkmi.h"
#include "batch.h"
#include "clk_device.h"
#include "clk_internal.h"
#include "clk_tuning.h"
#include "median_exact.h"
#include "median_impl.h"
#include "sep_impl.h"
#include "batch_x.h"

namespace clk {


template <int OP, class TS, class TD>
__global__ __launch_bounds__(256) void batch_x_kernel(const SepParams* __restrict__ ps,
                                                      const BatchGrid grid) {
  extern __shared__ uint32_t lds[];
  const BatchTile t = batch_tile(grid);
  sep_x_body<OP, TS, TD, 1>(batch_param(ps, grid, t.job), t.vblk, lds);
}

template <int OP, class TS, class TD>
__global__ __launch_bounds__(256) void batch_xpack_kernel(const SepParams* __restrict__ ps,
                                                          const BatchGrid grid) {
  extern __shared__ uint32_t lds[];
  const BatchTile t = batch_tile(grid);
  sep_xrows_body<OP, TS, TD>(batch_param(ps, grid, t.job), t.vblk, lds);
}

template <int OP, class TS, class TD, int COLS>
__global__ __launch_bounds__(256) void batch_strided_kernel(const SepParams* __restrict__ ps,
                                                            const BatchGrid grid) {
  const BatchTile t = batch_tile(grid);
  sep_strided_body<OP, TS, TD, 8, COLS>(batch_param(ps, grid, t.job), t.vblk);
}

template <class TS, class TD, int KIND>
__global__ __launch_bounds__(256) void batch_median_kernel(const MedParams* __restrict__ ms,
                                                           const BatchGrid grid) {
  const BatchTile t = batch_tile(grid);
  const MedParams& p = batch_param(ms, grid, t.job);
  const uint32_t bx = t.vblk % p.gx;
  const uint32_t r = t.vblk / p.gx;
  const uint32_t by = r % p.gy, bz = r / p.gy;
  if constexpr (KIND == MED_PK8) median_27_pk8_slide_body<TS, TD>(p, bx, by, bz);
  else if constexpr (KIND == MED_PK9) median_9_pk8_slide_body<TS, TD>(p, bx, by, bz);
  else if constexpr (KIND == MED_27 && sizeof(TD) == 4) median_27_key_slide_body<TS, TD>(p, bx, by, bz);
  else if constexpr (KIND == MED_27) median_27_body<TS, TD>(p, bx, by, bz);
  else if constexpr (KIND >= MED_REGS) median_regs_body<TS, TD, KIND - MED_REGS>(p, bx, by, bz);
  else median_generic_body<TS, TD>(p, bx, by, bz);
}

template <class TS, class TD>
__global__ __launch_bounds__(64) void batch_hist8_kernel(const MedParams* __restrict__ ms,
                                                         const BatchGrid grid) {
  const BatchTile t = batch_tile(grid);
  const MedParams& p = batch_param(ms, grid, t.job);
  const uint32_t bx = t.vblk % p.gx;
  const uint32_t r = t.vblk / p.gx;
  median_hist8_body<TS, TD>(p, bx, r % p.gy, r / p.gy);
}

namespace {

enum ClassKind : int { K_X = 0, K_STRIDED4 = 1, K_STRIDED1 = 2, K_MEDIAN = 3, K_XPACK = 4, K_YDYN = 5, K_XTIER = 6, K_XY = 7,
                       K_XYMM = 8, K_XYG = 9, K_XYZMM = 10 };

// Pixel-type pairs with batched kernels: same type, to float, from float.
bool batch_pair(int ts, int td) { return ts == td || td == CLK_FLOAT || ts == CLK_FLOAT; }

struct ClassKey {
  int kind, op, ts, td;  // op: separable op, or MedKind for medians
  int tier = 0, cols = 0;  // K_YDYN, K_XTIER, K_XY: half-width tier (K_XY: of y); K_YDYN, K_XY: columns per lane
  int fin = 0;             // K_YDYN: the branch-free body (Gaussian and mean of finite data)
  bool operator<(const ClassKey& o) const {
    return std::tie(kind, op, ts, td, tier, cols, fin) < std::tie(o.kind, o.op, o.ts, o.td, o.tier, o.cols, o.fin);
  }
};

struct ClassWork {
  std::vector<char> blob;        // parameter entries, `stride` bytes each
  uint32_t stride = 0;
  std::vector<uint32_t> off{0};  // first block of each job, then the grid size
  uint64_t nblocks = 0;
  size_t lds = 0;
  int64_t nslot = 0;  // K_XY: staged words per row (largest job)
  // device offsets inside the table area
  size_t params_off = 0, off_off = 0;
  // the first `bytes` of *p (every entry of a class takes the same)
  void add_params(const void* p, size_t bytes) {
    if (stride == 0) stride = (uint32_t)((bytes + 15) & ~size_t(15));
    const size_t at = blob.size();
    blob.resize(at + stride);
    memcpy(blob.data() + at, p, bytes < stride ? bytes : stride);
  }
  void add_job(uint64_t blocks) {
    nblocks += blocks;
    off.push_back((uint32_t)(nblocks < 0xffffffffULL ? nblocks : 0xffffffffULL));
  }
  BatchGrid grid(char* dev) const {
    return {reinterpret_cast<const uint32_t*>(dev + off_off), (uint32_t)(off.size() - 1), stride};
  }
};

// Bytes of a separable pass's table entry: the tier classes read weights
// w[0 .. 2*tier] only, so their entries end there; other classes keep the
// whole structure.
size_t sep_entry_bytes(int kind, int tier) {
  if (kind == K_XTIER || kind == K_YDYN) return offsetof(SepParams, w) + (size_t)(2 * tier + 1) * sizeof(float);
  return sizeof(SepParams);
}

// Launch tables go from pinned host memory to the device by a copy kernel
// (tables are whole 256-byte blocks).
__global__ __launch_bounds__(256) void table_upload_kernel(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                                           const uint32_t n16) {
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n16; i += gridDim.x * 256u) dst[i] = src[i];
}

// Pinned staging buffers for the launch tables, reused round robin; a slot
// is reused only after its previous upload completed.
struct Staging {
  void* host = nullptr;
  void* dev = nullptr;  // the same memory as the device reads it
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  bool pending = false;
};
constexpr int kStagingSlots = 8;  // up to three uploads per batch call
std::mutex g_stage_mu;
Staging g_stage[kStagingSlots];
int g_stage_next = 0;

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

int32_t sigma_n(float sigma) {
  int32_t n = (int32_t)(sigma * 8.0f);
  return (n % 2 == 0) ? n + 1 : n;
}

template <int OP, class TS, class TD>
void launch_class_typed(hipStream_t st, const ClassKey& k, const ClassWork& w, char* dev) {
  const uint32_t nt = (uint32_t)w.nblocks;
  const BatchGrid grid = w.grid(dev);
  const SepParams* ps = reinterpret_cast<const SepParams*>(dev + w.params_off);
  if (k.kind == K_X)
    hipLaunchKernelGGL((batch_x_kernel<OP, TS, TD>), dim3(nt), dim3(256), w.lds, st, ps, grid);
  else if (k.kind == K_XPACK)
    hipLaunchKernelGGL((batch_xpack_kernel<OP, TS, TD>), dim3(nt), dim3(256), w.lds, st, ps, grid);
  else if (k.kind == K_STRIDED4)
    hipLaunchKernelGGL((batch_strided_kernel<OP, TS, TD, 4>), dim3(nt), dim3(256), 0, st, ps, grid);
  else
    hipLaunchKernelGGL((batch_strided_kernel<OP, TS, TD, 1>), dim3(nt), dim3(256), 0, st, ps, grid);
}

template <class TS, class TD>
void launch_median_class(hipStream_t st, const ClassKey& k, const ClassWork& w, char* dev) {
  const uint32_t nt = (uint32_t)w.nblocks;
  const BatchGrid grid = w.grid(dev);
  const MedParams* ms = reinterpret_cast<const MedParams*>(dev + w.params_off);
  switch (k.op) {
    case MED_PK8: hipLaunchKernelGGL((batch_median_kernel<TS, TD, MED_PK8>), dim3(nt), dim3(256), 0, st, ms, grid); break;
    case MED_PK9: hipLaunchKernelGGL((batch_median_kernel<TS, TD, MED_PK9>), dim3(nt), dim3(256), 0, st, ms, grid); break;
    case MED_HIST8: hipLaunchKernelGGL((batch_hist8_kernel<TS, TD>), dim3(nt), dim3(64), 0, st, ms, grid); break;
    case MED_27: hipLaunchKernelGGL((batch_median_kernel<TS, TD, MED_27>), dim3(nt), dim3(256), 0, st, ms, grid); break;
#define CLK_MED_TIER(I)                                                                                     \
  case MED_REGS + I:                                                                                        \
    if constexpr (sizeof(TD) >= 2)                                                                          \
      hipLaunchKernelGGL((batch_median_kernel<TS, TD, MED_REGS + I>), dim3(nt), dim3(256), 0, st, ms, grid); \
    break;
    CLK_MED_TIER(0) CLK_MED_TIER(1) CLK_MED_TIER(2) CLK_MED_TIER(3) CLK_MED_TIER(4) CLK_MED_TIER(5)
#undef CLK_MED_TIER
    default: hipLaunchKernelGGL((batch_median_kernel<TS, TD, MED_GENERIC>), dim3(nt), dim3(256), 0, st, ms, grid); break;
  }
}

// Dispatch (op, ts, td) to a template instantiation; only batch_pair()
// combinations reach here.
template <int OP, class TS>
void sep_dst(hipStream_t st, const ClassKey& k, const ClassWork& w, char* dev) {
  switch (k.td) {
    case CLK_CHAR: if constexpr (std::is_same<TS, int8_t>::value || std::is_same<TS, float>::value) launch_class_typed<OP, TS, int8_t>(st, k, w, dev); break;
    case CLK_UCHAR: if constexpr (std::is_same<TS, uint8_t>::value || std::is_same<TS, float>::value) launch_class_typed<OP, TS, uint8_t>(st, k, w, dev); break;
    case CLK_SHORT: if constexpr (std::is_same<TS, int16_t>::value || std::is_same<TS, float>::value) launch_class_typed<OP, TS, int16_t>(st, k, w, dev); break;
    case CLK_USHORT: if constexpr (std::is_same<TS, uint16_t>::value || std::is_same<TS, float>::value) launch_class_typed<OP, TS, uint16_t>(st, k, w, dev); break;
    case CLK_FLOAT: launch_class_typed<OP, TS, float>(st, k, w, dev); break;
  }
}

template <int OP>
void sep_src(hipStream_t st, const ClassKey& k, const ClassWork& w, char* dev) {
  switch (k.ts) {
    case CLK_CHAR: sep_dst<OP, int8_t>(st, k, w, dev); break;
    case CLK_UCHAR: sep_dst<OP, uint8_t>(st, k, w, dev); break;
    case CLK_SHORT: sep_dst<OP, int16_t>(st, k, w, dev); break;
    case CLK_USHORT: sep_dst<OP, uint16_t>(st, k, w, dev); break;
    case CLK_FLOAT: sep_dst<OP, float>(st, k, w, dev); break;
  }
}

template <class TS>
void med_dst(hipStream_t st, const ClassKey& k, const ClassWork& w, char* dev) {
  switch (k.td) {
    case CLK_CHAR: if constexpr (std::is_same<TS, int8_t>::value || std::is_same<TS, float>::value) launch_median_class<TS, int8_t>(st, k, w, dev); break;
    case CLK_UCHAR: if constexpr (std::is_same<TS, uint8_t>::value || std::is_same<TS, float>::value) launch_median_class<TS, uint8_t>(st, k, w, dev); break;
    case CLK_SHORT: if constexpr (std::is_same<TS, int16_t>::value || std::is_same<TS, float>::value) launch_median_class<TS, int16_t>(st, k, w, dev); break;
    case CLK_USHORT: if constexpr (std::is_same<TS, uint16_t>::value || std::is_same<TS, float>::value) launch_median_class<TS, uint16_t>(st, k, w, dev); break;
    case CLK_FLOAT: launch_median_class<TS, float>(st, k, w, dev); break;
  }
}

int32_t launch_class(hipStream_t st, const ClassKey& k, const ClassWork& w, char* dev, std::string* err) {
  if (k.kind == K_XYZMM)
    return launch_xyzmm_class(st, k.op, k.ts, k.tier, reinterpret_cast<const MMParams*>(dev + w.params_off),
                              w.grid(dev), (uint32_t)w.nblocks, err);
  if (k.kind == K_XYG)
    return launch_xyg_class(st, k.op, k.ts, k.cols, k.tier, reinterpret_cast<const XYParams*>(dev + w.params_off),
                            w.grid(dev), (uint32_t)w.nblocks, err);
  if (k.kind == K_XYMM)
    return launch_xymm_class(st, k.op, k.ts, k.tier, reinterpret_cast<const XYParams*>(dev + w.params_off),
                             w.grid(dev), (uint32_t)w.nblocks, err);
  if (k.kind == K_XY)
    return launch_xy_class(st, k.op, k.ts, k.tier, k.cols, reinterpret_cast<const XYParams*>(dev + w.params_off),
                           w.grid(dev), (uint32_t)w.nblocks, w.lds, (int)w.nslot, err);
  if (k.kind == K_XTIER)
    return launch_xtier_class(st, k.op, k.ts, k.td, k.tier, reinterpret_cast<const SepParams*>(dev + w.params_off),
                              w.grid(dev), (uint32_t)w.nblocks, err);
  if (k.kind == K_YDYN)
    return launch_ydyn_class(st, k.op, k.ts, k.td, k.tier, k.cols, k.fin != 0,
                             reinterpret_cast<const SepParams*>(dev + w.params_off),
                             w.grid(dev), (uint32_t)w.nblocks, err);
  if (k.kind == K_MEDIAN) {
    switch (k.ts) {
      case CLK_CHAR: med_dst<int8_t>(st, k, w, dev); break;
      case CLK_UCHAR: med_dst<uint8_t>(st, k, w, dev); break;
      case CLK_SHORT: med_dst<int16_t>(st, k, w, dev); break;
      case CLK_USHORT: med_dst<uint16_t>(st, k, w, dev); break;
      case CLK_FLOAT: med_dst<float>(st, k, w, dev); break;
    }
    return 0;
  }
  switch (k.op) {
    case OP_GAUSS: sep_src<OP_GAUSS>(st, k, w, dev); break;
    case OP_MEAN: sep_src<OP_MEAN>(st, k, w, dev); break;
    case OP_MIN: sep_src<OP_MIN>(st, k, w, dev); break;
    case OP_MAX: sep_src<OP_MAX>(st, k, w, dev); break;
  }
  return 0;
}

// One single-axis pass of a job.  src_tmp/dst_tmp >= 0 are byte offsets
// of a temporary in the batch's temp area (the pointer is set once the area
// is allocated).
struct Pass {
  int op;
  clk_image dst, src;
  int dim, N;
  float s;
  int64_t src_tmp, dst_tmp;
  bool fin;  // the job's source is an integer image: every value finite
};

// The fused x+y pass of a job (batch_xy.h): src -> temporary of type tt at
// byte offset dst_tmp, the y pass's output; the job's z pass reads it.
struct XYPass {
  int op;
  clk_image dst, src;  // dst: the job's image (sizes)
  int Nx, Ny;
  float sx, sy;
  int tt;
  int64_t dst_tmp;
  int cols;  // K_XY: columns per thread; 0: K_XYMM; -1: K_XYG (registers and LDS, 8-pixel chunks)
  bool zaxis = false;  // K_XYMM: the fused axis is z (Ny is z's; the job's y pass follows)
};

// Minimum/maximum of 16-bit integer pixels with half widths up to
// kXYMMMaxC along x and y and kXYZMaxCZ along z, and rows short enough for
// a tile of at least 4 output rows (the threads of a block hold
// R = kXYZThreads / ng rows of ng chunks), can run whole in one pass
// (batch_xyzmm.h).  Off by default: one block per CU and two barriers a
// plane left it at about 1.2 TB/s on the uneven batch, 1.5 ms against 1.2 ms
// for the fused x+y pass and the z pass of the same jobs.
bool xyzmm_ok(int op, const clk_filter_job& jb, const int (&Ns)[3]) {
  if (tuning(TUNE_NO_XYZMM) != 0) return false;
  if (op != OP_MIN && op != OP_MAX) return false;
  if (jb.src.dtype != CLK_USHORT && jb.src.dtype != CLK_SHORT) return false;
  if (jb.dst.dtype != jb.src.dtype) return false;  // the chain's last pass converts; this class writes the source type
  for (int a = 0; a < 3; ++a)
    if (Ns[a] < 1 || (Ns[a] - 1) / 2 > (a == 2 ? kXYZMaxCZ : kXYMMMaxC)) return false;
  const int64_t W = jb.dst.width, H = jb.dst.height, D = jb.dst.depth;
  // rows of at least one 8-pixel chunk (chunk16_start)
  if (W < 8 || H <= 0 || D <= 0 || W * H * D >= (int64_t(1) << 31) - 1) return false;
  const int64_t ng = (W + 7) / 8;
  const int cy = (Ns[1] - 1) / 2;
  return ng <= kXYZThreads && kXYZThreads / ng - 2 * cy >= 4;
}

// Gaussian/mean jobs whose x and y half widths are at most kXYMaxC can take
// the register fused class (batch_xyg.h) for 8-bit unsigned, 16-bit and
// float sources.  Off by default: on the 256-volume uneven batch its classes
// (64 chunks x a y segment per block, 112-177 VGPRs) ran at 110-220 Gvoxel/s
// against the separate x and y classes' combined rate, 1.94 ms against
// 1.29 ms for the same jobs.
bool xyg_ok(int op, const clk_filter_job& jb, const int (&Ns)[3]) {
  if (tuning(TUNE_NO_XYG) != 0) return false;
  if (op != OP_GAUSS && op != OP_MEAN) return false;
  const int ts = jb.src.dtype;
  if (ts != CLK_UCHAR && ts != CLK_SHORT && ts != CLK_USHORT && ts != CLK_FLOAT) return false;
  const int cx = (Ns[0] - 1) / 2, cy = (Ns[1] - 1) / 2;
  if (Ns[0] < 1 || Ns[1] < 1 || cx > kXYMaxC || cy > kXYMaxC) return false;
  const int64_t W = jb.dst.width, H = jb.dst.height, D = jb.dst.depth;
  return W > 0 && H > 0 && D > 0 && W * H * D < (int64_t(1) << 31) - 1;
}

// Minimum/maximum of 16-bit integer pixels with the x half width and that
// of the fused axis (y, or z with xymm_z) up to kXYMMMaxC take the
// register-only fused class (batch_xymm.h).  Minimum and maximum of
// integers are exact in any order, so x with z, then y, gives the bits of
// x, y, z.
bool xymm_ok(int op, const clk_filter_job& jb, const int (&Ns)[3], int axis) {
  if (tuning(TUNE_NO_XYMM) != 0) return false;
  if (op != OP_MIN && op != OP_MAX) return false;
  if (jb.src.dtype != CLK_USHORT && jb.src.dtype != CLK_SHORT) return false;
  const int cx = (Ns[0] - 1) / 2, cy = (Ns[axis] - 1) / 2;
  if (Ns[0] < 1 || Ns[axis] < 1 || cx > kXYMMMaxC || cy > kXYMMMaxC) return false;
  const int64_t W = jb.dst.width, H = jb.dst.height, D = jb.dst.depth;
  // rows of at least one 8-pixel chunk (chunk16_start); images under 2^30
  // bytes (the class's dropped stores use offsets from 2^31)
  return W >= 8 && H > 0 && D > 0 && W * H * D < (int64_t(1) << 29);
}

// Columns per thread of a fused x+y pass for a job, or 0 when the job takes
// the separate passes: half widths up to 16 (c = (N-1)/2 as the kernels
// truncate), integer sources for min/max, 32-bit offsets, and a padded
// staging row that fits the block's slots.  Wider columns cut LDS reads;
// the y tier bounds them by registers as in the y/z tier classes.
int xy_cols(int op, const clk_filter_job& jb, const int (&Ns)[3]) {
  if (tuning(TUNE_NO_XY) != 0) return 0;
  const int cx = (Ns[0] - 1) / 2, cy = (Ns[1] - 1) / 2;
  if (Ns[0] < 1 || Ns[1] < 1 || cx > kXYMaxC || cy > kXYMaxC) return 0;
  if ((op == OP_MIN || op == OP_MAX) && jb.src.dtype == CLK_FLOAT) return 0;
  const int64_t W = jb.dst.width, H = jb.dst.height, D = jb.dst.depth;
  if (W <= 0 || H <= 0 || D <= 0 || W * H * D >= (int64_t(1) << 31) - 1) return 0;
  const int tier = ydyn_tier(cy);
  const int cands[3] = {tier == 5 ? 4 : 2, tier == 5 ? 2 : 1, tier == 5 ? 1 : 0};
  for (int cols : cands) {
    if (cols == 0) break;
    const int64_t ng = (W + cols - 1) / cols;
    if (xy_nslot(ng, D, cols, cx) - cols <= (int64_t)kXYSlots * 256) return cols;
  }
  return 0;
}

// A pass that does not go through a batched class (rare pixel-type pairs,
// degenerate parameters) runs as a single launch, in stage order.
struct Single {
  int stage;
  Pass pass;
};

}  // namespace

int32_t run_batch(hipStream_t st, const clk_filter_job* jobs, int32_t njobs, std::string* err) {
  {
    // The launch tables go up from reused host staging memory: a captured
    // copy would read whatever that memory holds at each replay.
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess) (void)hipGetLastError();
    if (cs != hipStreamCaptureStatusNone) {
      *err = "clk_batch_filter_3d: a stream under graph capture is not supported (launch tables are "
             "uploaded from reused host staging memory)";
      return CLK_INVALID_OPERATION;
    }
  }
  // ---- plan: passes per stage, temporaries --------------------------------
  std::vector<std::vector<Pass>> stage_passes(3);
  std::vector<XYPass> xy_passes;  // stage 0
  std::vector<std::pair<clk_filter_job, std::array<int, 3>>> mm3;  // whole 3D min/max, stage 0
  std::vector<std::pair<clk_filter_job, int>> medians;  // job, unused
  std::vector<std::pair<clk_image, clk_image>> copies;  // no active axis: converting copy

  int64_t temp_bytes = 0;
  for (int j = 0; j < njobs; ++j) {
    const clk_filter_job& jb = jobs[j];
    if (jb.kind == CLK_JOB_MEDIAN_BOX_3D) {
      medians.push_back({jb, 0});
      continue;
    }
    const int op = jb.kind;  // CLK_JOB_* of separable filters equal the Op codes
    bool act[3];
    int Ns[3];
    for (int a = 0; a < 3; ++a) {
      if (op == OP_GAUSS) {
        act[a] = jb.sigma[a] > 0.f;
        Ns[a] = sigma_n(jb.sigma[a]);
      } else {
        act[a] = jb.radius[a] > 0;
        Ns[a] = 2 * jb.radius[a] + 1;
      }
    }
    int nact = (int)act[0] + (int)act[1] + (int)act[2];
    if (nact == 0) {
      copies.push_back({jb.dst, jb.src});
      continue;
    }
    // temporaries: float for Gaussian and mean; the source type for min/max,
    // whose values are exact in any type that holds the source
    const int tdt = (op == OP_MIN || op == OP_MAX) ? jb.src.dtype : CLK_FLOAT;
    const int64_t V = jb.dst.width * jb.dst.height * jb.dst.depth;
    if (nact == 3 && xyzmm_ok(op, jb, Ns)) {
      mm3.push_back({jb, {Ns[0], Ns[1], Ns[2]}});
      continue;
    }
    if (nact == 3) {
      // 16-bit min/max: x with z (a wave's chunks in consecutive rows of one
      // plane, contiguous) when xymm_z asks for it, else x with y
      const bool mmz = tuning(TUNE_XYMM_Z) != 0 && xymm_ok(op, jb, Ns, 2);
      const bool mm = mmz || xymm_ok(op, jb, Ns, 1);
      const bool xg = !mm && xyg_ok(op, jb, Ns);
      const int cols = mm ? 0 : (xg ? -1 : xy_cols(op, jb, Ns));
      if (mm || cols != 0) {
        // x and y (or z) in one pass into one temporary, then the last pass
        const int64_t toff = temp_bytes;
        temp_bytes += (int64_t)align256((size_t)(V * clk_dtype_size(tdt)));
        const bool g = op == OP_GAUSS;
        const int fa = mmz ? 2 : 1, la = mmz ? 1 : 2;  // fused axis, last axis
        XYPass xp{op, jb.dst, jb.src, Ns[0], Ns[fa], g ? jb.sigma[0] : 0.f, g ? jb.sigma[fa] : 0.f, tdt, toff, cols};
        xp.zaxis = mmz;
        xy_passes.push_back(xp);
        clk_image tmp = jb.dst;
        tmp.dtype = tdt;
        tmp.ptr = nullptr;
        stage_passes[1].push_back({op, jb.dst, tmp, la, Ns[la], g ? jb.sigma[la] : 0.f, toff, -1, jb.src.dtype != CLK_FLOAT});
        continue;
      }
    }
    const int ntmp = nact - 1 < 2 ? nact - 1 : 2;
    std::vector<int64_t> toff(ntmp);
    for (int i = 0; i < ntmp; ++i) {
      toff[i] = temp_bytes;
      temp_bytes += (int64_t)align256((size_t)(V * clk_dtype_size(tdt)));
    }
    int done = 0, ti = 0, stage = 0;
    clk_image cur = jb.src;
    int64_t cur_tmp = -1;
    for (int a = 0; a < 3; ++a) {
      if (!act[a]) continue;
      ++done;
      clk_image out = jb.dst;
      int64_t out_tmp = -1;
      if (done < nact) {
        out.dtype = tdt;
        out.ptr = nullptr;
        out_tmp = toff[ti];
        ti ^= 1;
      }
      stage_passes[stage++].push_back({op, out, cur, a, Ns[a], op == OP_GAUSS ? jb.sigma[a] : 0.f,
                                       cur_tmp, out_tmp, jb.src.dtype != CLK_FLOAT});
      cur = out;
      cur_tmp = out_tmp;
    }
  }

  // ---- temporaries, table area ---------------------------------------------
  const int bseg = tuning(TUNE_BATCH_SEG) > 0 ? tuning(TUNE_BATCH_SEG) : 32;
  const int xyseg = tuning(TUNE_BATCH_XYSEG) > 0 ? tuning(TUNE_BATCH_XYSEG) : 64;
  std::vector<Single> singles;
  void* temp_base = nullptr;
  CallScratch scratch(st);  // held until the last launch is issued
  if (temp_bytes > 0 && scratch.get(1, (size_t)temp_bytes, &temp_base) != hipSuccess) {
    *err = "batch temporaries: allocation failed";
    return CLK_MEM_OBJECT_ALLOCATION_FAILURE;
  }
  for (int s = 0; s < 3; ++s)
    for (Pass& ps : stage_passes[s]) {
      if (ps.dst_tmp >= 0) ps.dst.ptr = static_cast<char*>(temp_base) + ps.dst_tmp;
      if (ps.src_tmp >= 0) ps.src.ptr = static_cast<char*>(temp_base) + ps.src_tmp;
    }
  // Each stage's tables go up in their own region of one device block, so
  // the host plans stage s + 1 while stage s runs; the block's size bounds
  // every stage's classes (an entry and a block offset per job, two 256-byte
  // alignments per class, at most one class per job).
  size_t tcap = 0;
  for (int s = 0; s < 3; ++s) tcap += stage_passes[s].size() * (sizeof(SepParams) + 4 + 2 * 256 + 4);
  tcap += xy_passes.size() * (sizeof(XYParams) + 4 + 2 * 256 + 4);
  tcap += mm3.size() * (sizeof(MMParams) + 4 + 2 * 256 + 4);
  tcap += medians.size() * (sizeof(MedParams) + 4 + 2 * 256 + 4);
  char* dev_tables = nullptr;
  if (tcap > 0 && scratch.get(2, tcap, (void**)&dev_tables) != hipSuccess) {
    *err = "batch tables: allocation failed";
    return CLK_MEM_OBJECT_ALLOCATION_FAILURE;
  }
  size_t tused = 0;

  // ---- classes of one stage ---------------------------------------------------
  auto classify_stage = [&](const int s, std::map<ClassKey, ClassWork>& classes) -> int32_t {
    for (const Pass& ps : stage_passes[s]) {
      SepParams p;
      int axis_is_x = 0;
      bool nan_fill = false;
      int32_t rc = prepare_separable(ps.op, ps.dst, ps.src, ps.dim, ps.N, ps.s, p, &axis_is_x, &nan_fill, err);
      if (rc) return rc;
      const int64_t V = ps.dst.width * ps.dst.height * ps.dst.depth;
      if (V == 0) continue;
      if (nan_fill || !batch_pair(ps.src.dtype, ps.dst.dtype) || p.L >= (int64_t(1) << 30)) {
        singles.push_back({s, ps});
        continue;
      }
      if (((ps.dim == 2 && tuning(TUNE_NO_ZMM) == 0) || (ps.dim == 1 && tuning(TUNE_NO_YMM) == 0)) &&
          (ps.op == OP_MIN || ps.op == OP_MAX) && ps.src.dtype == ps.dst.dtype &&
          (ps.src.dtype == CLK_USHORT || ps.src.dtype == CLK_SHORT) && p.c >= 1 && p.c <= kXYMMMaxC &&
          V < (int64_t(1) << 29) && ps.dst.width >= 8) {
        // 16-bit minimum/maximum along z (or y): the register min/max class
        // walking planes (rows), 8-pixel chunks, packed pairs, no x filter.
        // Off by default: on the uneven batch it made the call 3.27 -> 3.48
        // ms against the y/z ring classes as a z pass.
        const bool zw = ps.dim == 2;
        XYParams q;
        memset(&q, 0, sizeof(q));
        q.src = ps.src.ptr;
        q.dst = ps.dst.ptr;
        q.W = (int32_t)ps.dst.width;
        q.H = (int32_t)(zw ? ps.dst.depth : ps.dst.height);  // the walk
        q.D = (int32_t)(zw ? ps.dst.height : ps.dst.depth);  // its outer slices
        q.cx = 0;
        q.cy = p.c;
        q.ng = (q.W + 7) / 8;
        q.ngt = q.D * q.ng;
        int zseg = (xyseg + kXYK - 1) / kXYK * kXYK;
        const int hk = (q.H + kXYK - 1) / kXYK * kXYK;
        if (zseg > hk) zseg = hk;
        q.seg = zseg;
        q.nseg = (q.H + zseg - 1) / zseg;
        q.sy = zw ? ps.dst.width * ps.dst.height : ps.dst.width;
        q.sz = zw ? ps.dst.width : ps.dst.width * ps.dst.height;
        ClassWork& w = classes[ClassKey{K_XYMM, ps.op, ps.src.dtype, ps.dst.dtype, p.c <= 5 ? 5 : 8, 0}];
        w.add_params(&q, sizeof(XYParams));
        const uint64_t waves = (uint64_t)((q.ngt + 60) / 61) * (uint64_t)q.nseg;  // 61 chunks per wave
        w.add_job((waves + 3) / 4);
        continue;
      }
      ClassKey key{0, ps.op, ps.src.dtype, ps.dst.dtype};
      uint32_t blocks = 0;
      size_t lds = 0;
      const int64_t per_row = (p.L + 7) / 8;
      if (axis_is_x && xtier_tier(p.c) > 0 && xtier_pair(ps.src.dtype, ps.dst.dtype) && p.L >= 8 &&
          p.outer * per_row < (int64_t(1) << 31) - 64 && tuning(TUNE_NO_XTIER) == 0) {
        // register windows over per-row 8-pixel chunks, one class per tier
        // whatever each job's row length and c
        key.kind = K_XTIER;
        key.tier = xtier_tier(p.c);
        set_nseg(p, per_row);
        p.ntasks = p.outer * per_row;
        const int ss = clk_dtype_size(ps.src.dtype), ds = clk_dtype_size(ps.dst.dtype);
        p.vec_ld = ((p.L * ss) % 16 == 0) && ((uintptr_t)p.src % 16 == 0);
        p.vec_st = ((p.L * ds) % 16 == 0) && ((uintptr_t)p.dst % 16 == 0);
        if (ps.op == OP_GAUSS) set_tier_weights(p, key.tier);
        const int owned = 64 - 2 * ((key.tier + 7) / 8);  // chunks a wave owns (xt_owned)
        const int64_t units = (p.ntasks + owned - 1) / owned;
        const int64_t waves = (units + kXTierUnits - 1) / kXTierUnits;
        blocks = (uint32_t)((waves + 3) / 4);
      } else if (axis_is_x && p.L <= 256 && p.c <= 256 && tuning(TUNE_NO_XPACK) == 0) {
        key.kind = K_XPACK;  // short rows: several per wave, vector loads and stores of their span
        plan_xpack(p);
        blocks = (uint32_t)((p.ntasks + 3) / 4);
        lds = (size_t)4 * p.lds_stride * 4 + 4 * (512 * (size_t)clk_dtype_size(ps.dst.dtype) + 16);
      } else if (axis_is_x && p.c <= kMaxCX) {
        key.kind = K_X;
        const int64_t W = p.L;
        set_nseg(p, (W + kXSeg - 1) / kXSeg);
        p.ntasks = p.outer * p.nseg;
        p.lds_stride = lds_pad(kXSeg + 2 * p.c) + 1;
        const int ss = clk_dtype_size(ps.src.dtype), ds = clk_dtype_size(ps.dst.dtype);
        p.vec_ld = (W % kXK == 0) && ((uintptr_t)p.src % (kXK * ss) == 0);
        p.vec_st = (W % kXK == 0) && ((uintptr_t)p.dst % (kXK * ds) == 0);
        blocks = (uint32_t)((p.ntasks + 3) / 4);
        lds = (size_t)4 * p.lds_stride * 4;
      } else if (ydyn_tier(p.c) > 0 && ydyn_pair(ps.op, ps.src.dtype, ps.dst.dtype) && tuning(TUNE_STRIDED_VARIANT) == 0) {
        // y/z: the ring body for a run-time half width (one class per tier
        // and column width, whatever each job's c)
        const int ss = clk_dtype_size(ps.src.dtype), ds = clk_dtype_size(ps.dst.dtype);
        const bool v4 = (p.inner % 4 == 0) && ((uintptr_t)p.src % (4 * ss) == 0) &&
                        ((uintptr_t)p.dst % (4 * ds) == 0);
        const bool v2 = (p.inner % 2 == 0) && ((uintptr_t)p.src % (2 * ss) == 0) &&
                        ((uintptr_t)p.dst % (2 * ds) == 0);
        key.kind = K_YDYN;
        key.tier = ydyn_tier(p.c);
        const int ycap = tuning(TUNE_BATCH_YCOLS);
        key.cols = ydyn_cols(key.tier, v4 && (ycap == 0 || ycap >= 4), v2 && (ycap == 0 || ycap >= 2));
        const int K = ydyn_k(key.tier, key.cols);
        // finite data (chains from integer images): the branch-free body
        key.fin = ps.fin && (ps.op == OP_GAUSS || ps.op == OP_MEAN) && tuning(TUNE_YDYN_SKIP) == 0;
        if (ps.op == OP_GAUSS || key.fin) set_ydyn_weights(p, ps.op, key.tier, key.fin != 0);
        p.ncolgroups = p.inner / key.cols;
        p.ntasks = 0;
        const int ypad = tuning(TUNE_BATCH_YPAD);  // lanes: 0 (packed), 16, 32 or 64
        if (ypad > 0 && p.outer > 1 && p.ncolgroups % ypad != 0) {
          // y passes: each row's column groups padded to a whole number of
          // ypad lanes, so a wave's loads hold at most 64 / ypad row pieces,
          // each starting its row (packed, a wave's columns ran from the end
          // of one plane's row into the next plane's at any lane)
          p.ntasks = p.ncolgroups;
          p.ncolgroups = (p.ncolgroups + ypad - 1) / ypad * ypad;
        }
        p.ncols_total = p.outer * p.ncolgroups;
        const int64_t nblk = (p.ncols_total + 255) / 256;
        int64_t seg = ((int64_t)bseg + K - 1) / K * K;
        if (seg > (p.L + K - 1) / K * K) seg = (p.L + K - 1) / K * K;
        const int64_t nseg = (p.L + seg - 1) / seg;
        p.seg = (int32_t)seg;
        set_nseg(p, nseg);
        p.nblk = (uint32_t)nblk;
        blocks = (uint32_t)(nblk * nseg);
      } else {
        const int ss = clk_dtype_size(ps.src.dtype), ds = clk_dtype_size(ps.dst.dtype);
        const bool v4 = (p.inner % 4 == 0) && ((uintptr_t)p.src % (4 * ss) == 0) &&
                        ((uintptr_t)p.dst % (4 * ds) == 0);
        const int cols = v4 ? 4 : 1;
        key.kind = v4 ? K_STRIDED4 : K_STRIDED1;
        p.ncolgroups = p.inner / cols;
        p.ncols_total = p.outer * p.ncolgroups;
        const int64_t nblk = (p.ncols_total + 255) / 256;
        int64_t seg = ((int64_t)bseg + 7) / 8 * 8;
        if (seg > (p.L + 7) / 8 * 8) seg = (p.L + 7) / 8 * 8;
        const int64_t nseg = (p.L + seg - 1) / seg;
        p.seg = (int32_t)seg;
        set_nseg(p, nseg);
        p.nblk = (uint32_t)nblk;
        blocks = (uint32_t)(nblk * nseg);
      }
      ClassWork& w = classes[key];
      w.add_params(&p, sep_entry_bytes(key.kind, key.tier));
      if (lds > w.lds) w.lds = lds;
      w.add_job(blocks);
    }
    if (s != 0) return 0;
    const int zseg = tuning(TUNE_BATCH_ZSEG) > 0 ? tuning(TUNE_BATCH_ZSEG) : 64;
    for (const auto& m : mm3) {
      const clk_filter_job& jb = m.first;
      MMParams q;
      memset(&q, 0, sizeof(q));
      q.src = jb.src.ptr;
      q.dst = jb.dst.ptr;
      q.W = (int32_t)jb.dst.width;
      q.H = (int32_t)jb.dst.height;
      q.D = (int32_t)jb.dst.depth;
      q.cx = (m.second[0] - 1) / 2;
      q.cy = (m.second[1] - 1) / 2;
      q.cz = (m.second[2] - 1) / 2;
      q.ng = (q.W + 7) / 8;
      q.R = kXYZThreads / q.ng;
      q.TY = q.R - 2 * q.cy;
      q.nyt = (q.H + q.TY - 1) / q.TY;
      q.zseg = zseg;
      q.nzseg = (q.D + zseg - 1) / zseg;
      ClassWork& w = classes[ClassKey{K_XYZMM, jb.kind, jb.src.dtype, jb.dst.dtype, kXYZMaxCZ, 0}];
      w.add_params(&q, sizeof(MMParams));
      w.add_job((uint64_t)q.nyt * (uint64_t)q.nzseg);
    }
    for (const XYPass& xp : xy_passes) {
      clk_image tmp = xp.dst;
      tmp.dtype = xp.tt;
      tmp.ptr = static_cast<char*>(temp_base) + xp.dst_tmp;
      SepParams px, py;
      int ax = 0;
      bool nf = false;
      int32_t rc = prepare_separable(xp.op, tmp, xp.src, 0, xp.Nx, xp.sx, px, &ax, &nf, err);
      if (rc == 0 && !nf) rc = prepare_separable(xp.op, tmp, tmp, xp.zaxis ? 2 : 1, xp.Ny, xp.sy, py, &ax, &nf, err);
      if (rc) return rc;
      if (nf) {
        *err = "internal: fused x+y pass planned for an empty tap window";
        return CLK_INVALID_VALUE;
      }
      const int cols = xp.cols;
      const int tier = cols == 0 ? (py.c <= 5 ? 5 : 8) : ydyn_tier(py.c);
      const int xtier = px.c <= 8 ? 8 : 16;  // K_XYG
      XYParams q;
      memset(&q, 0, sizeof(q));
      q.src = xp.src.ptr;
      q.dst = tmp.ptr;
      q.W = (int32_t)xp.dst.width;
      q.H = (int32_t)(xp.zaxis ? xp.dst.depth : xp.dst.height);  // the walk
      q.D = (int32_t)(xp.zaxis ? xp.dst.height : xp.dst.depth);  // its outer slices
      q.cx = px.c;
      q.cy = py.c;
      q.ng = (int32_t)((q.W + (cols > 0 ? cols : 8) - 1) / (cols > 0 ? cols : 8));
      q.ngt = q.D * q.ng;
      q.rstr = cols > 0 ? xy_rstr(q.ng, cols, q.cx) : 0;
      int seg = (xyseg + kXYK - 1) / kXYK * kXYK;
      const int hk = (q.H + kXYK - 1) / kXYK * kXYK;
      if (seg > hk) seg = hk;
      q.seg = seg;
      q.nseg = (q.H + seg - 1) / seg;
      q.sy = xp.zaxis ? (int64_t)q.W * q.D : q.W;  // elements between rows of the walk
      q.sz = xp.zaxis ? q.W : (int64_t)q.W * q.H;  // and between its outer slices
      const int ds = clk_dtype_size(xp.tt);
      q.vec_st = cols > 0 && (q.W % cols == 0) && ((uintptr_t)tmp.ptr % (size_t)(cols * ds) == 0);
      q.divx = px.div;
      q.rdivx = px.rdiv;
      q.fastx = px.fastdiv;
      q.divy = py.div;
      q.rdivy = py.rdiv;
      q.fasty = py.fastdiv;
      if (xp.op == OP_GAUSS) {
        // x weights in tap order (K_XY) or at the x tier's window (K_XYG)
        const int xo = cols < 0 ? xtier - px.c : 0;
        for (int t = 0; t <= 2 * px.c; ++t) q.wx[t + xo] = px.w[t];
        for (int t = 0; t <= 2 * py.c; ++t) q.wy[t + tier - py.c] = py.w[t];
      }
      if (cols < 0) {
        ClassWork& w = classes[ClassKey{K_XYG, xp.op, xp.src.dtype, xp.tt, tier, xtier}];
        w.add_params(&q, sizeof(XYParams));
        w.add_job((uint64_t)((q.ngt + 63) / 64) * (uint64_t)q.nseg);
        continue;
      }
      if (cols == 0) {
        ClassWork& w = classes[ClassKey{K_XYMM, xp.op, xp.src.dtype, xp.tt, tier, 0}];
        w.add_params(&q, sizeof(XYParams));
        const uint64_t waves = (uint64_t)((q.ngt + 60) / 61) * (uint64_t)q.nseg;  // 61 chunks per wave
        w.add_job((waves + 3) / 4);
        continue;
      }
      ClassWork& w = classes[ClassKey{K_XY, xp.op, xp.src.dtype, xp.tt, tier, cols}];
      w.add_params(&q, sizeof(XYParams));
      const int64_t nsl = (xy_nslot(q.ng, q.D, cols, q.cx) + 3) / 4 * 4;
      if (nsl > w.nslot) w.nslot = nsl;
      w.lds = xy_lds_bytes(w.nslot, cols);
      w.add_job((uint64_t)((q.ngt + 255) / 256) * (uint64_t)q.nseg);
    }
    for (auto& mj : medians) {
      const clk_filter_job& jb = mj.first;
      MedParams p;
      memset(&p, 0, sizeof(p));
      p.src = jb.src.ptr;
      p.dst = jb.dst.ptr;
      p.W = jb.dst.width;
      p.H = jb.dst.height;
      p.D = jb.dst.depth;
      p.ex = jb.radius[0];
      p.ey = jb.radius[1];
      p.ez = jb.radius[2];
      p.n = (2 * p.ex + 1) * (2 * p.ey + 1) * (2 * p.ez + 1);
      if (p.W * p.H * p.D == 0) continue;
      if (!batch_pair(jb.src.dtype, jb.dst.dtype)) {
        std::string e2;
        const int32_t rc = launch_median(st, jb.dst, jb.src, 2 * p.ex + 1, 2 * p.ey + 1, 2 * p.ez + 1, &e2);
        if (rc) { *err = e2; return rc; }
        continue;
      }
      dim3 grid;
      int threads;
      const MedKind mk = plan_median(p, clk_dtype_size(jb.dst.dtype), &grid, &threads);
      // the 3x3x3 bodies slide along z and the packed 3x3 one along y, as in
      // single launches
      if (mk == MED_PK8) {
        grid.z = (uint32_t)((p.D + kPk8ZRun - 1) / kPk8ZRun);
      } else if (mk == MED_27 && clk_dtype_size(jb.dst.dtype) == 4) {
        grid.z = (uint32_t)((p.D + kF27ZRun - 1) / kF27ZRun);
      } else if (mk == MED_PK9) {
        grid.y = (uint32_t)((p.H + kPk9YRun - 1) / kPk9YRun);
        p.gy = grid.y;
      }
      ClassKey key{K_MEDIAN, (int)mk, jb.src.dtype, jb.dst.dtype};
      ClassWork& w = classes[key];
      w.add_params(&p, sizeof(MedParams));
      w.add_job((uint64_t)grid.x * grid.y * grid.z);
    }

    return 0;
  };

  // ---- upload one stage's tables (pinned staging ring, stream order) --------
  auto upload_stage = [&](std::map<ClassKey, ClassWork>& classes) -> int32_t {
    size_t tb = 0;
    for (auto& kv : classes) {
      ClassWork& w = kv.second;
      w.params_off = tused + tb;
      tb += align256(w.blob.size());
      w.off_off = tused + tb;
      tb += align256(w.off.size() * sizeof(uint32_t));
    }
    if (tb == 0) return 0;
    if (tused + tb > tcap) {
      *err = "internal: batch tables exceed their bound";
      return CLK_OUT_OF_RESOURCES;
    }
    std::lock_guard<std::mutex> lk(g_stage_mu);
    Staging& sg = g_stage[g_stage_next];
    g_stage_next = (g_stage_next + 1) % kStagingSlots;
    if (sg.pending) {
      (void)hipEventSynchronize(sg.ev);
      sg.pending = false;
    }
    if (!sg.ev && hipEventCreateWithFlags(&sg.ev, hipEventDisableTiming) != hipSuccess) sg.ev = nullptr;
    if (sg.cap < tb) {
      // grows by powers of two from 1 MB: freeing pinned memory waits for
      // the whole device, so a slot is replaced rarely
      size_t cap = size_t(1) << 20;
      while (cap < tb) cap <<= 1;
      if (sg.host) (void)hipHostFree(sg.host);
      sg.host = sg.dev = nullptr;
      sg.cap = 0;
      if (hipHostMalloc(&sg.host, cap, hipHostMallocMapped) != hipSuccess ||
          hipHostGetDevicePointer(&sg.dev, sg.host, 0) != hipSuccess) {
        (void)hipGetLastError();
        if (sg.host) (void)hipHostFree(sg.host);
        sg.host = sg.dev = nullptr;
        *err = "batch tables: pinned staging allocation failed";
        return CLK_OUT_OF_RESOURCES;
      }
      sg.cap = cap;
    }
    char* const h = static_cast<char*>(sg.host);  // device offset tused + i lands at h + i
    for (auto& kv : classes) {
      const ClassWork& w = kv.second;
      if (!w.blob.empty()) memcpy(h + (w.params_off - tused), w.blob.data(), w.blob.size());
      memcpy(h + (w.off_off - tused), w.off.data(), w.off.size() * sizeof(uint32_t));
    }
    // a copy kernel pulls the tables over the bus: hipMemcpyAsync of a few
    // hundred KB from pinned memory blocked the host (it copied them itself,
    // after waiting for the work queued before it, 0.2 to 1.3 ms a stage)
    const uint32_t n16 = (uint32_t)(tb / 16);
    const uint32_t nb = (n16 + 255) / 256 < 256 ? (n16 + 255) / 256 : 256;
    hipLaunchKernelGGL(table_upload_kernel, dim3(nb), dim3(256), 0, st, static_cast<const u32x4*>(sg.dev),
                       reinterpret_cast<u32x4*>(dev_tables + tused), n16);
    if (hipGetLastError() != hipSuccess) {
      *err = "batch tables: upload failed";
      return CLK_OUT_OF_RESOURCES;
    }
    if (sg.ev && hipEventRecord(sg.ev, st) == hipSuccess) sg.pending = true;
    tused += tb;
    return 0;
  };

  // ---- launch stage by stage --------------------------------------------------
  // Stage s + 1 is planned and its tables uploaded while stage s runs.  The
  // classes of one stage touch disjoint jobs; with batch_serial 0 they run
  // side by side on the stream's side streams (forked from st after the
  // stage's table upload, joined back into st; the largest classes first,
  // one per side stream in turn), which overlaps the ramp and tail of
  // launches of a few thousand blocks but costs about 0.5 ms of host time a
  // call in event calls; on the C4 batch the two measured alike, so the
  // default runs them in order on st.
  hipStream_t* aux = nullptr;
  hipEvent_t* joins = nullptr;
  hipEvent_t fork = nullptr;
  const int naux = tuning(TUNE_BATCH_SERIAL) ? 0 : scratch.aux(&aux, &joins, &fork);
  int32_t rc = 0;
  for (auto& c : copies) {
    rc = launch_convert_copy(st, c.first, c.second, err);
    if (rc) return rc;
  }
  int used = 0;  // side streams of the previous stage, not yet joined
  auto join = [&]() {
    for (int a = 0; a < used; ++a)
      if (hipEventRecord(joins[a], aux[a]) != hipSuccess || hipStreamWaitEvent(st, joins[a], 0) != hipSuccess) {
        (void)hipStreamSynchronize(aux[a]);
        if (rc == 0) {
          *err = "batch: side stream join failed";
          rc = CLK_OUT_OF_RESOURCES;
        }
      }
    used = 0;
  };
  std::vector<std::map<ClassKey, ClassWork>> classes(3);
  for (int s = 0; s < 3 && rc == 0; ++s) {
    rc = classify_stage(s, classes[s]);
    if (rc == 0) rc = upload_stage(classes[s]);
    join();  // stage s - 1 done before stage s starts (and before its singles)
    if (rc) break;
    std::vector<std::pair<const ClassKey*, const ClassWork*>> order;
    for (auto& kv : classes[s])
      if (kv.second.nblocks > 0) order.push_back({&kv.first, &kv.second});
    std::stable_sort(order.begin(), order.end(), [](const auto& a, const auto& b) {
      return a.second->nblocks > b.second->nblocks;
    });
    const bool side = naux > 0 && order.size() > 1;
    if (side && hipEventRecord(fork, st) != hipSuccess) {
      *err = "batch: fork event record failed";
      rc = CLK_OUT_OF_RESOURCES;
      break;
    }
    for (size_t i = 0; i < order.size() && rc == 0; ++i) {
      const ClassWork& w = *order[i].second;
      if (w.nblocks > 0x7fffffffULL) {
        *err = "batch too large for one launch";
        rc = CLK_INVALID_VALUE;
        break;
      }
      hipStream_t q = st;
      if (side) {
        const int a = (int)(i % (size_t)naux);
        q = aux[a];
        if (a >= used) {
          if (hipStreamWaitEvent(q, fork, 0) != hipSuccess) {
            *err = "batch: side stream wait failed";
            rc = CLK_OUT_OF_RESOURCES;
            break;
          }
          used = a + 1;
        }
      }
      rc = launch_class(q, *order[i].first, w, dev_tables, err);
      if (rc) break;
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) {
        *err = std::string("batch launch failed: ") + hipGetErrorString(e);
        rc = CLK_OUT_OF_RESOURCES;
        break;
      }
    }
    // passes without a batched class run as single launches on st, after
    // their stage's classes
    bool any_single = false;
    for (const Single& sg : singles) any_single |= sg.stage == s;
    if (any_single) {
      join();
      for (const Single& sg : singles) {
        if (sg.stage != s || rc) continue;
        rc = launch_separable(sg.pass.op, st, sg.pass.dst, sg.pass.src, sg.pass.dim, sg.pass.N, sg.pass.s, err);
      }
    }
  }
  join();  // even after a failed launch, so st never runs ahead of side work
  // float -> float medians: the NaN / -0 fix-up per job, one flag buffer
  // reused in stream order
  int64_t flag_bytes = 0;
  for (auto& mj : medians) {
    const clk_filter_job& jb = mj.first;
    if (jb.src.dtype == CLK_FLOAT && jb.dst.dtype == CLK_FLOAT) {
      const int64_t b = nanzero_flag_bytes(jb.dst);
      if (b > flag_bytes) flag_bytes = b;
    }
  }
  if (rc == 0 && flag_bytes > 0) {
    void* flags = nullptr;
    if (scratch.get(0, (size_t)flag_bytes, &flags) != hipSuccess) {
      *err = "median brick flags: allocation failed";
      return CLK_MEM_OBJECT_ALLOCATION_FAILURE;
    }
    for (auto& mj : medians) {
      const clk_filter_job& jb = mj.first;
      rc = launch_median_nanzero(st, jb.dst, jb.src, jb.radius[0], jb.radius[1], jb.radius[2], false, flags, err);
      if (rc) break;
    }
  }
  return rc;
}

}  // namespace clk
