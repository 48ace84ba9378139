// Host side of one single-axis separable launch: maps (dim, N, s) of the
// reference kernel signature onto SepParams and picks the pass kernel.
// Also holds the small converting-copy and fill kernels.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "clk_device.h"
#include "clk_internal.h"
#include "clk_tuning.h"
#include "sep_impl.h"

namespace clk {

// exp() of one Gaussian tap argument.  The reference evaluates the OpenCL
// builtin exp per voxel and tap (gaussian_blur_separable_3d_x.cl:29); here it
// is evaluated once per launch on the host, in double precision and rounded
// to float, so the weight is the correctly rounded value in all but
// vanishingly rare cases (OpenCL 1.2 allows 3 ulp).
float expf_weight(float x) { return (float)exp((double)x); }

template <class TS, class TD>
__global__ __launch_bounds__(256) void convert_copy_kernel(const TS* __restrict__ src,
                                                           TD* __restrict__ dst, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
    dst[i] = convert_sat<TD>(static_cast<float>(src[i]));
}

// 4 consecutive pixels per lane per step and consecutive lanes on
// consecutive groups of 4, so every load and store instruction of a wave
// covers one contiguous span (a 32-byte-per-lane store as two 16-byte halves
// would leave every 64-byte line half written per instruction).  A thread
// takes kCopySteps steps 1024 pixels apart, their loads issued before any
// store; pointers are aligned to 4 pixels, a partial last group goes pixel
// by pixel.
constexpr int kCopySteps = 4;

template <class TS, class TD>
__global__ __launch_bounds__(256) void convert_copy4_kernel(const TS* __restrict__ src,
                                                            TD* __restrict__ dst, int64_t n) {
  const int64_t base = (int64_t)blockIdx.x * (1024 * kCopySteps) + (int64_t)threadIdx.x * 4;
  PixVec<TS, 4> v[kCopySteps];
#pragma unroll
  for (int u = 0; u < kCopySteps; ++u) {
    const int64_t i = base + (int64_t)u * 1024;
    if (i + 4 <= n) v[u] = load_nt<TS, 4>(src + i);
  }
#pragma unroll
  for (int u = 0; u < kCopySteps; ++u) {
    const int64_t i = base + (int64_t)u * 1024;
    if (i + 4 <= n) {
      TD o[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) o[j] = convert_sat<TD>(static_cast<float>(v[u].v[j]));
      store_nt<TD, 4>(dst + i, o);
    } else {
      for (int64_t j = i; j < n; ++j) dst[j] = convert_sat<TD>(static_cast<float>(src[j]));
    }
  }
}

template <class TD>
__global__ __launch_bounds__(256) void fill_kernel(TD* __restrict__ dst, int64_t n, TD value) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dst[i] = value;
}

static uint32_t grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (uint32_t)b;
}

template <class TS, class TD>
static int32_t copy_typed(hipStream_t st, const clk_image& dst, const clk_image& src, int64_t n) {
  const int64_t blocks = (n + 1024 * kCopySteps - 1) / (1024 * kCopySteps);
  const bool vec = (uintptr_t)src.ptr % (4 * sizeof(TS)) == 0 && (uintptr_t)dst.ptr % (4 * sizeof(TD)) == 0 &&
                   blocks <= 0x7fffffffLL;
  if (vec)
    hipLaunchKernelGGL((convert_copy4_kernel<TS, TD>), dim3((uint32_t)blocks), dim3(256), 0, st,
                       static_cast<const TS*>(src.ptr), static_cast<TD*>(dst.ptr), n);
  else
    hipLaunchKernelGGL((convert_copy_kernel<TS, TD>), dim3(grid_for(n)), dim3(256), 0, st,
                       static_cast<const TS*>(src.ptr), static_cast<TD*>(dst.ptr), n);
  return 0;
}

template <class TS>
static int32_t copy_dst(hipStream_t st, const clk_image& dst, const clk_image& src, int64_t n) {
  switch (dst.dtype) {
    case CLK_CHAR:   return copy_typed<TS, int8_t>(st, dst, src, n);
    case CLK_UCHAR:  return copy_typed<TS, uint8_t>(st, dst, src, n);
    case CLK_SHORT:  return copy_typed<TS, int16_t>(st, dst, src, n);
    case CLK_USHORT: return copy_typed<TS, uint16_t>(st, dst, src, n);
    case CLK_FLOAT:  return copy_typed<TS, float>(st, dst, src, n);
  }
  return CLK_INVALID_VALUE;
}

int32_t launch_convert_copy(hipStream_t st, const clk_image& dst, const clk_image& src,
                            std::string* err) {
  const int64_t n = dst.width * dst.height * dst.depth;
  if (n == 0) return 0;
  int32_t rc = CLK_INVALID_VALUE;
  switch (src.dtype) {
    case CLK_CHAR:   rc = copy_dst<int8_t>(st, dst, src, n); break;
    case CLK_UCHAR:  rc = copy_dst<uint8_t>(st, dst, src, n); break;
    case CLK_SHORT:  rc = copy_dst<int16_t>(st, dst, src, n); break;
    case CLK_USHORT: rc = copy_dst<uint16_t>(st, dst, src, n); break;
    case CLK_FLOAT:  rc = copy_dst<float>(st, dst, src, n); break;
  }
  if (rc != 0) { *err = "invalid pixel type"; return rc; }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) { *err = std::string("kernel launch failed: ") + hipGetErrorString(e); return CLK_OUT_OF_RESOURCES; }
  return 0;
}

// Every voxel gets CONVERT_dst(NaN): what the reference computes for mean and
// Gaussian when N < 0 leaves the tap loop empty (0/0).
static int32_t fill_nan(hipStream_t st, const clk_image& dst, std::string* err) {
  const int64_t n = dst.width * dst.height * dst.depth;
  if (n == 0) return 0;
  const uint32_t g = grid_for(n);
  switch (dst.dtype) {
    case CLK_CHAR:   hipLaunchKernelGGL((fill_kernel<int8_t>), dim3(g), dim3(256), 0, st, (int8_t*)dst.ptr, n, (int8_t)0); break;
    case CLK_UCHAR:  hipLaunchKernelGGL((fill_kernel<uint8_t>), dim3(g), dim3(256), 0, st, (uint8_t*)dst.ptr, n, (uint8_t)0); break;
    case CLK_SHORT:  hipLaunchKernelGGL((fill_kernel<int16_t>), dim3(g), dim3(256), 0, st, (int16_t*)dst.ptr, n, (int16_t)0); break;
    case CLK_USHORT: hipLaunchKernelGGL((fill_kernel<uint16_t>), dim3(g), dim3(256), 0, st, (uint16_t*)dst.ptr, n, (uint16_t)0); break;
    case CLK_FLOAT:  hipLaunchKernelGGL((fill_kernel<float>), dim3(g), dim3(256), 0, st, (float*)dst.ptr, n, __builtin_nanf("")); break;
    default: *err = "invalid dst pixel type"; return CLK_INVALID_VALUE;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) { *err = std::string("kernel launch failed: ") + hipGetErrorString(e); return CLK_OUT_OF_RESOURCES; }
  return 0;
}

int32_t prepare_separable(int op, const clk_image& dst, const clk_image& src, int32_t dim, int32_t N,
                          float s, SepParams& p, int* axis_is_x, bool* fill_nan_out, std::string* err) {
  const int64_t W = dst.width, H = dst.height, D = dst.depth;
  *fill_nan_out = false;
  int c = (N - 1) / 2;  // C division truncates toward zero, as the kernels' (N-1)/2
  if (c < 0) {
    // empty tap loop: min/max keep the centre (== a one-tap window);
    // mean and Gaussian divide 0 by 0.
    if (op == OP_MIN || op == OP_MAX) {
      c = 0;
    } else {
      *fill_nan_out = true;
      return 0;
    }
  }
  if (op == OP_GAUSS && 2 * (int64_t)c + 1 > kMaxTaps) {
    *err = "gaussian_blur_separable_3d: N exceeds CLK_GAUSSIAN_MAX_TAPS";
    return CLK_INVALID_VALUE;
  }
  memset(&p, 0, sizeof(p));
  p.src = src.ptr;
  p.dst = dst.ptr;
  p.c = c;
  p.ntaps = 2 * c + 1;
  *axis_is_x = 0;
  switch (dim) {
    case 0: p.inner = 1; p.L = W; p.outer = H * D; *axis_is_x = 1; break;
    case 1: p.inner = W; p.L = H; p.outer = D; break;
    case 2: p.inner = W * H; p.L = D; p.outer = 1; break;
    default: p.inner = W * H * D; p.L = 1; p.outer = 1; break;  // dir = (0,0,0)
  }
  if (op == OP_GAUSS) {
    // gaussian_blur_separable_3d_x.cl:24,28-31 in float: n = -2*s*s,
    // h = exp((v*v)/n), hsum accumulated in tap order.
    const float n = -2 * s * s;
    float hsum = 0.f;
    for (int v = -c; v <= c; ++v) {
      const float h = expf_weight((float)(v * v) / n);
      p.w[v + c] = h;
      hsum += h;
    }
    p.div = hsum;
    p.rdiv = 1.0f / hsum;
    // integer pixels: every nonzero numerator is at least the smallest weight
    float hmin = 1.f;
    for (int v = 0; v <= 2 * c; ++v) hmin = (p.w[v] < hmin) ? p.w[v] : hmin;
    p.fastdiv = (src.dtype != CLK_FLOAT) && (hmin >= 0x1p-48f);
  } else if (op == OP_MEAN) {
    p.div = (float)(2 * c + 1);  // mean_separable_3d_x.cl:26 counts the taps in float
    p.rdiv = 1.0f / p.div;
    // integer sums: 0 or >= 1 in magnitude (the kernels also assume this
    // flag for integer sources at compile time)
    p.fastdiv = (src.dtype != CLK_FLOAT);
  }
  return 0;
}

int32_t launch_fill_nan(hipStream_t st, const clk_image& dst, std::string* err) {
  return fill_nan(st, dst, err);
}

namespace {

// One launch of the pass kernel with the current tuning (prepare + pick).
int32_t launch_once(int op, hipStream_t st, const clk_image& dst, const clk_image& src, int32_t dim,
                    int32_t N, float s, SepParams& p, std::string* err) {
  int axis_is_x = 0;
  bool nan_fill = false;
  int32_t rc = prepare_separable(op, dst, src, dim, N, s, p, &axis_is_x, &nan_fill, err);
  if (rc) return rc;
  if (nan_fill) return fill_nan(st, dst, err);
  switch (op) {
    case OP_GAUSS: return launch_sep_gauss(st, p, axis_is_x, src.dtype, dst.dtype, err);
    case OP_MEAN:  return launch_sep_mean(st, p, axis_is_x, src.dtype, dst.dtype, err);
    case OP_MIN:   return launch_sep_min(st, p, axis_is_x, src.dtype, dst.dtype, err);
    case OP_MAX:   return launch_sep_max(st, p, axis_is_x, src.dtype, dst.dtype, err);
  }
  *err = "invalid op";
  return CLK_INVALID_VALUE;
}

// ---- first-launch tuning of launch shapes ----------------------------------
// With the autotune knob on, the first launch of each signature (device, op,
// pixel types, axis, N, volume size) times a few launch shapes on the
// caller's stream -- every candidate writes the same output, since results
// never depend on launch shape -- and later launches of that signature use
// the fastest.  Small volumes and streams under graph capture skip it.

struct TuneSig {
  int dev, op, ts, td, dim, N;
  int64_t w, h, d;
  bool operator<(const TuneSig& o) const {
    return std::tie(dev, op, ts, td, dim, N, w, h, d) < std::tie(o.dev, o.op, o.ts, o.td, o.dim, o.N, o.w, o.h, o.d);
  }
};

struct Candidate {
  TuneSetting set[5];
  int n;
};

// x axis: register window (when it applies) or LDS segments of 1/2/4 x 512
const Candidate kXCands[] = {
    {{}, 0},
    {{{TUNE_XREG_UNITS, 1}}, 1},
    {{{TUNE_XREG_UNITS, 4}}, 1},
    {{{TUNE_NO_XREG, 1}, {TUNE_X_SUB, 1}}, 2},
    {{{TUNE_NO_XREG, 1}, {TUNE_X_SUB, 2}}, 2},
    {{{TUNE_NO_XREG, 1}, {TUNE_X_SUB, 2}, {TUNE_X_TASKS, 2}}, 3},
    {{{TUNE_NO_XREG, 1}, {TUNE_X_SUB, 1}, {TUNE_X_TASKS, 2}}, 3},
    {{{TUNE_NO_XREG, 1}, {TUNE_X_SUB, 4}}, 2},
    {{{TUNE_XREG_MAXC, 16}, {TUNE_XREG_UNITS, 4}}, 2},
    {{{TUNE_XREG_MAXC, 16}, {TUNE_XREG_UNITS, 2}}, 2},
};
// y / z: coordinate descent over blocks per slot, block order, column
// width and the ring pass's step shape (then blocks per slot again, as the
// best count depends on order)
struct TuneDim {
  TuneKey key;
  int n;
  int32_t vals[6];
};
const TuneDim kYZDims[] = {
    {TUNE_YRING, 2, {1, 2}},  // float -> float only: the lockstep ring pass or not
    {TUNE_STRIDED_ROUNDS, 6, {2, 4, 8, 16, 32, 64}},
    {TUNE_STRIDED_ORDER, 4, {0, 1, 2, 3}},
    {TUNE_YSTATIC_COLS, 3, {0, 1, 2}},
    {TUNE_YRING_DEEP, 2, {0, 1}},  // float -> float only: loads 1 or 4-8 steps ahead
    {TUNE_STRIDED_ROUNDS, 6, {2, 4, 8, 16, 32, 64}},
};
constexpr int64_t kTuneMinVoxels = int64_t(1) << 22;

std::mutex g_tune_mu;
std::map<TuneSig, Candidate> g_tuned;

// Times every candidate in four rounds (the first warms up) and sets *bi
// to the one with the smallest of its timed launches; rc != 0 when a launch
// failed.
int32_t time_candidates(int op, hipStream_t st, const clk_image& dst, const clk_image& src, int32_t dim,
                        int32_t N, float s, SepParams& p, const Candidate* cands, int n, int* bi,
                        std::string* err) {
  constexpr int kTimed = 3;
  std::vector<hipEvent_t> ev(2 * n * kTimed, nullptr);
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) {
      for (auto& f : ev) if (f) (void)hipEventDestroy(f);
      *err = "autotune: event creation failed";
      return CLK_OUT_OF_RESOURCES;
    }
  int32_t rc = 0;
  for (int pass = 0; pass <= kTimed && rc == 0; ++pass)
    for (int i = 0; i < n && rc == 0; ++i) {
      ScopedTuning scope(cands[i].set, cands[i].n);
      hipEvent_t* e = pass > 0 ? &ev[2 * (i * kTimed + pass - 1)] : nullptr;
      if (e) (void)hipEventRecord(e[0], st);
      rc = launch_once(op, st, dst, src, dim, N, s, p, err);
      if (e) (void)hipEventRecord(e[1], st);
    }
  if (rc == 0 && hipEventSynchronize(ev[2 * ((n - 1) * kTimed + kTimed - 1) + 1]) != hipSuccess) {
    *err = "autotune: synchronisation failed";
    rc = CLK_OUT_OF_RESOURCES;
  }
  if (rc == 0) {
    float best_ms = -1.f;
    *bi = 0;
    for (int i = 0; i < n; ++i)
      for (int t = 0; t < kTimed; ++t) {
        float ms = 0.f;
        const int k = 2 * (i * kTimed + t);
        if (hipEventElapsedTime(&ms, ev[k], ev[k + 1]) != hipSuccess) continue;
        if (best_ms < 0.f || ms < best_ms) {
          best_ms = ms;
          *bi = i;
        }
      }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return rc;
}

int32_t tune_signature(int op, hipStream_t st, const clk_image& dst, const clk_image& src, int32_t dim,
                       int32_t N, float s, SepParams& p, bool axis_x, Candidate* best, std::string* err) {
  int bi = 0;
  if (axis_x) {
    const int n = (int)(sizeof(kXCands) / sizeof(kXCands[0]));
    const int32_t rc = time_candidates(op, st, dst, src, dim, N, s, p, kXCands, n, &bi, err);
    if (rc == 0) *best = kXCands[bi];
    return rc;
  }
  // current point: rounds, order, cols, ring, ring steps (the launcher's defaults)
  const bool float_pair = src.dtype == CLK_FLOAT && dst.dtype == CLK_FLOAT;
  Candidate cur{{{TUNE_STRIDED_ROUNDS, 16},
                 {TUNE_STRIDED_ORDER, 0},
                 {TUNE_YSTATIC_COLS, 0},
                 {TUNE_YRING, 0},
                 {TUNE_YRING_DEEP, 0}},
                5};
  for (const TuneDim& d : kYZDims) {
    // no ring pass for other pixel types
    if ((d.key == TUNE_YRING || d.key == TUNE_YRING_DEEP) && !float_pair) continue;
    Candidate c[6];
    int slot = 0;
    while (cur.set[slot].key != d.key) ++slot;
    for (int i = 0; i < d.n; ++i) {
      c[i] = cur;
      c[i].set[slot].value = d.vals[i];
    }
    const int32_t rc = time_candidates(op, st, dst, src, dim, N, s, p, c, d.n, &bi, err);
    if (rc) return rc;
    cur = c[bi];
  }
  *best = cur;
  return 0;
}

}  // namespace

int32_t autotune_save(const char* path) {
  FILE* f = fopen(path, "w");
  if (!f) return CLK_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  fprintf(f, "# clkmi launch shapes: op src dst dim N width height depth : key=value ...\n");
  for (const auto& kv : g_tuned) {
    const TuneSig& t = kv.first;
    fprintf(f, "%d %d %d %d %d %lld %lld %lld :", t.op, t.ts, t.td, t.dim, t.N, (long long)t.w,
            (long long)t.h, (long long)t.d);
    for (int i = 0; i < kv.second.n; ++i)
      fprintf(f, " %s=%d", tuning_name(kv.second.set[i].key), kv.second.set[i].value);
    fprintf(f, "\n");
  }
  return fclose(f) == 0 ? 0 : CLK_INVALID_VALUE;
}

int32_t autotune_load(const char* path) {
  FILE* f = fopen(path, "r");
  if (!f) return CLK_INVALID_VALUE;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  char line[512];
  int32_t rc = 0;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  while (fgets(line, sizeof(line), f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    TuneSig t{};
    long long w = 0, h = 0, d = 0;
    int used = 0;
    if (sscanf(line, "%d %d %d %d %d %lld %lld %lld :%n", &t.op, &t.ts, &t.td, &t.dim, &t.N, &w, &h, &d,
               &used) < 8 || used == 0) {
      rc = CLK_INVALID_VALUE;
      break;
    }
    t.dev = dev;
    t.w = w;
    t.h = h;
    t.d = d;
    Candidate c{};
    const char* q = line + used;
    char key[64];
    int val = 0, adv = 0;
    while (c.n < (int)(sizeof(c.set) / sizeof(c.set[0])) && sscanf(q, " %63[a-z_]=%d%n", key, &val, &adv) == 2) {
      int k = 0;
      while (k < TUNE_COUNT && strcmp(tuning_name(k), key) != 0) ++k;
      if (k == TUNE_COUNT) {
        rc = CLK_INVALID_VALUE;
        break;
      }
      c.set[c.n++] = TuneSetting{(TuneKey)k, (int32_t)val};
      q += adv;
    }
    if (rc) break;
    g_tuned[t] = c;
  }
  fclose(f);
  return rc;
}

int32_t launch_separable(int op, hipStream_t st, const clk_image& dst, const clk_image& src,
                         int32_t dim, int32_t N, float s, std::string* err, bool allow_tune) {
  const int64_t W = dst.width, H = dst.height, D = dst.depth;
  if (W * H * D == 0) return 0;
  // the launch parameters (weights included, ~2 KB) of this thread's launch
  thread_local SepParams tl_params;
  SepParams& p = tl_params;
  int32_t rc = 0;
  bool done = false;
  if (tuning(TUNE_AUTOTUNE) == 1 && W * H * D >= kTuneMinVoxels && N > 1) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    int dev = 0;
    if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone &&
        hipGetDevice(&dev) == hipSuccess) {
      const TuneSig sig{dev, op, src.dtype, dst.dtype, dim, N, W, H, D};
      Candidate c{};
      bool have = false;
      {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(sig);
        if (it != g_tuned.end()) {
          c = it->second;
          have = true;
        }
      }
      if (!have && allow_tune) {
        rc = tune_signature(op, st, dst, src, dim, N, s, p, dim == 0, &c, err);
        if (rc == 0) {
          std::lock_guard<std::mutex> lk(g_tune_mu);
          g_tuned[sig] = c;
          if (getenv("CLK_AUTOTUNE_LOG")) {
            fprintf(stderr, "clkmi autotune: op=%d src=%d dst=%d dim=%d N=%d %lldx%lldx%lld ->", op, src.dtype,
                    dst.dtype, dim, N, (long long)W, (long long)H, (long long)D);
            for (int i = 0; i < c.n; ++i) fprintf(stderr, " %s=%d", tuning_name(c.set[i].key), c.set[i].value);
            fprintf(stderr, "\n");
          }
        }
        done = true;  // the candidates already wrote dst
      }
      if (!done && have) {
        ScopedTuning scope(c.set, c.n);
        rc = launch_once(op, st, dst, src, dim, N, s, p, err);
        done = true;
      }
    }
  }
  if (!done) rc = launch_once(op, st, dst, src, dim, N, s, p, err);
  return rc;
}

}  // namespace clk
