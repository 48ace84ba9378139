// Internal interface between the C-ABI layer (clkmi_api.cpp) and the kernel
// launchers (sep_kernels.hip, median_kernels.hip).
#pragma once
#include <memory>
#include <mutex>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/clkmi.h"

namespace clk {

constexpr int kMaxTaps = CLK_GAUSSIAN_MAX_TAPS;

// Parameters of one single-axis separable pass.  The volume is viewed as
// [outer][L][inner]: `inner` elements are contiguous, the filtered axis has
// length L and stride `inner`.  dim 0 -> (inner=1, L=W, outer=H*D),
// dim 1 -> (W, H, D), dim 2 -> (W*H, D, 1), any other dim -> (V, 1, 1) so
// that every clamped tap reads the centre voxel.
struct SepParams {
  const void* src;
  void* dst;
  int64_t inner;
  int64_t L;
  int64_t outer;
  int64_t ncolgroups;  // strided pass: column groups of COLS per slab
  int64_t ncols_total; // strided pass: outer * ncolgroups
  int64_t ntasks;      // x pass: rows * segments
  uint32_t nblk;       // strided pass: blocks covering ncols_total
  int32_t nseg;        // segments along the axis (per column group / row)
  int32_t seg;         // outputs per segment along the axis (multiple of K)
  int32_t c;           // half width, >= 0
  int32_t ntaps;       // 2c+1
  float div;           // Gaussian: sum of weights; mean: tap count
  float rdiv;          // RN(1/div)
  int32_t lds_stride;  // x pass: 4-byte words of LDS per wave
  int32_t vec_ld;      // x pass: rows are aligned for vector loads
  int32_t vec_st;      // x pass: rows are aligned for vector stores
  int32_t fastdiv;     // numerators never need the guarded divide
  int32_t order;       // strided pass block order: bit 0 = segments of a column
                       // block adjacent, bit 1 = contiguous ranges per XCD
  uint32_t nseg_m;     // n / nseg = (n * nseg_m) >> (31 + nseg_l) for n < 2^31
  int32_t nseg_l;
  float w[kMaxTaps];   // Gaussian weights in tap order v = -c..c
};

// Sets p.nseg and the constants of an exact division by it: with
// l = ceil(log2 nseg) and m = ceil(2^(31+l) / nseg) (m < 2^32 for
// nseg < 2^31), floor(n * m / 2^(31+l)) = floor(n / nseg) for every
// 0 <= n < 2^31 (Granlund & Montgomery, PLDI 1994, Thm 4.2 with N = 31).
inline void set_nseg(SepParams& p, int64_t nseg) {
  p.nseg = (int32_t)nseg;
  int l = 0;
  while ((int64_t(1) << l) < nseg) ++l;
  p.nseg_l = l;
  p.nseg_m = (uint32_t)(((uint64_t(1) << (31 + l)) + (uint64_t)nseg - 1) / (uint64_t)nseg);
}

struct Status {
  int32_t code = 0;
  std::string msg;
};

// Reference-semantics Gaussian weight exp(x) for one tap, computed on the
// host once per launch.
float expf_weight(float x);

// Single-axis separable launch; op is Op (clk_device.h).  Arguments are
// already validated by the caller (types, sizes, no overlap).  With
// allow_tune false the launch never times candidates (it uses a shape the
// autotuner already chose, or the default): no events, no synchronisation.
int32_t launch_separable(int op, hipStream_t st, const clk_image& dst,
                         const clk_image& src, int32_t dim, int32_t N, float s,
                         std::string* err, bool allow_tune = true);

// Tuned launch shapes of single launches: write them to / merge them from
// a text file (see clk_autotune_save / clk_autotune_load).
int32_t autotune_save(const char* path);
int32_t autotune_load(const char* path);

// Fill SepParams for one single-axis launch (weights, divisor, axis view).
// Sets *fill_nan when the launch reduces to writing CONVERT(NaN) everywhere.
int32_t prepare_separable(int op, const clk_image& dst, const clk_image& src, int32_t dim, int32_t N,
                          float s, SepParams& p, int* axis_is_x, bool* fill_nan, std::string* err);
int32_t launch_fill_nan(hipStream_t st, const clk_image& dst, std::string* err);

// Stream-ordered temporaries (free them with hipFreeAsync) from a memory pool
// of the library's own per device, which keeps freed memory for the next
// call instead of handing it back to the driver at every synchronisation as
// the default pool does (release threshold 0).  clk_release_temporaries
// trims it.
hipError_t temp_alloc(void** p, size_t bytes, hipStream_t st);

// The temporaries of one call (host chain, batch).  Normally they are
// device blocks kept per (device, stream, slot) and reused by the next call
// on that stream: work on one stream runs in issue order, so a call's
// temporaries are free once its kernels are queued behind it.  A larger
// request replaces the block (the old one is freed in stream order).  A new
// allocation per call of a new size cost about 0.45 ms of host time.  The
// scope holds that stream's lock from construction until destruction, so
// two host threads on one stream cannot interleave launches that share a
// block; calls on other streams are not blocked.
//
// Streams on which one handle does not name one stream in issue order get
// transient blocks instead, allocated per call and freed in stream order
// when the scope ends: a stream under graph capture (the graph then owns
// its allocations; a kept block would be baked into the graph and could be
// freed or shared by later eager calls) and hipStreamPerThread (one handle,
// a different stream in every host thread).
class CallScratch {
 public:
  explicit CallScratch(hipStream_t st);
  ~CallScratch();
  CallScratch(const CallScratch&) = delete;
  CallScratch& operator=(const CallScratch&) = delete;
  hipError_t get(int slot, size_t bytes, void** p);
  bool transient() const { return transient_; }
  // Side streams of this (device, stream) for independent launches of one
  // call (created on first use, kept with the blocks, default priority),
  // with one event each and one for the fork: 0 when the call must stay on
  // its own stream (transient mode, or creation failed).
  int aux(hipStream_t** streams, hipEvent_t** joins, hipEvent_t* fork);

 private:
  hipStream_t st_;
  bool transient_ = false;
  void* blocks_ = nullptr;  // the stream's kept blocks (cached mode)
  std::unique_lock<std::recursive_mutex> lock_;
  void* owned_[3] = {nullptr, nullptr, nullptr};  // transient mode
};
constexpr int kScratchSlots = 3;  // 0: host chain / median brick flags, 1: batch temporaries, 2: batch tables
constexpr int kAuxStreams = 4;    // side streams per stream (one per hardware queue a process gets by default)

// Kept blocks of one stream on the current device: bytes held, and release
// (synchronises the stream, frees the blocks, trims the pool).
int64_t kept_bytes(hipStream_t st);
hipError_t release_kept(hipStream_t st);
// Memory the library's pool holds on the current device (kept blocks and
// freed memory not yet returned to the driver).
int64_t pool_reserved_bytes();

// The fixed device array of clk_filter_3d_noalloc (noalloc.hip): its size,
// and its address on the current device (resolved once per device).
constexpr size_t kNoAllocScratchBytes = CLK_NOALLOC_SCRATCH_BYTES;
hipError_t noalloc_scratch(void** ptr);

// detect_maxima_3d (detect = true; dst and src of one size) or
// maximum_z_projection (dst W x H x 1); arguments validated by the caller.
int32_t launch_consumer(bool detect, hipStream_t st, const clk_image& dst, const clk_image& src, std::string* err);

// minimum/maximum/mean/median_sphere_3d (op 0..3); arguments validated by
// the caller.
int32_t launch_sphere(int op, hipStream_t st, const clk_image& dst, const clk_image& src, int32_t Nx,
                      int32_t Ny, int32_t Nz, std::string* err);

// median_box_3d launch (median_box_2d with wrap = true and Nz = 1: the 2D
// kernel's C cast of integer values); arguments validated by the caller.
int32_t launch_median(hipStream_t st, const clk_image& dst, const clk_image& src,
                      int32_t Nx, int32_t Ny, int32_t Nz, std::string* err, bool wrap = false);

// Run a batch of 3D filter jobs stage by stage, one launch per kernel class
// and stage (batch.hip).
int32_t run_batch(hipStream_t st, const clk_filter_job* jobs, int32_t njobs, std::string* err);

// Converting copy dst <- CONVERT(src) (used by the host chain for axes that
// are not filtered).
int32_t launch_convert_copy(hipStream_t st, const clk_image& dst,
                            const clk_image& src, std::string* err);

}  // namespace clk
