// Single-axis separable passes for gfx950: gaussian_blur_separable_3d,
// mean_separable_3d, minimum_separable_3d, maximum_separable_3d.
//
// Every output voxel gets exactly the reference's per-voxel arithmetic:
//   Gaussian  res = fma(h_v, x_v, res) for v = -c..c in order, then res/hsum
//             (gaussian_blur_separable_3d_x.cl:26-33)
//   mean      res = res + x_v in order, then res/count (mean_separable_3d_x.cl:22-28)
//   min/max   res = centre; res = min/max(res, x_v) in order
//             (minimum_separable_3d_x.cl:19-22, maximum_separable_3d_x.cl:19-22)
// with clamp-to-edge taps and CONVERT_dst on the way out.  What changes is
// how the taps reach the ALUs:
//
//  * Each thread produces K consecutive outputs along the filtered axis.
//    The window is streamed through a ring of K registers: tap v of output q
//    reads ring slot (q+v) mod K, and after tap v slot (v mod K) is refilled
//    with input v+K.  The tap loop is unrolled by K so every ring index is a
//    compile-time register, and each refill is issued K taps before its use
//    (a software pipeline).  When the window fits the ring (N <= K) the ring
//    is carried into the next block of K outputs, so along a segment every
//    input is loaded exactly once.
//  * y/z passes (and an out-of-range dim): lanes cover consecutive x
//    columns, COLS columns per lane with one vector load, so every load of a
//    wave is one contiguous run; the axis position is wave-uniform, and
//    blocks of K outputs away from the volume border use unclamped
//    addresses.
//  * x pass: waves walk row segments of 512 outputs.  A segment is fetched
//    with one vector load per lane (plus c halo values on each side) while
//    the previous segment is being filtered, staged in LDS as 32-bit values
//    with one pad word every 8 so that the lane-strided ring reads are free
//    of bank conflicts.
//  * min/max of integer pixels run on 32-bit integers (exact, and no NaN
//    handling); Gaussian/mean divide by the launch-uniform normaliser with a
//    reciprocal and one FMA correction, which is the IEEE quotient.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdlib.h>

#include <mutex>
#include <string>
#include <type_traits>
#include <map>
#include <utility>

#include "clk_device.h"
#include "clk_internal.h"
#include "clk_tuning.h"

namespace clk {

// n / p.nseg for n < 2^31 by the multiplier set_nseg computed; for a
// wave-uniform n this stays on the scalar unit
__device__ __forceinline__ uint32_t div_nseg(const SepParams& p, uint32_t n) {
  return (uint32_t)(((uint64_t)n * p.nseg_m) >> (31 + p.nseg_l));
}

template <int OP, class TS>
using vt_t = std::conditional_t<((OP == OP_MIN || OP == OP_MAX) && Pix<TS>::kInt), int32_t, float>;

template <bool B> struct BoolC { static constexpr bool value = B; };
template <int I> struct IntC { static constexpr int value = I; };

template <class TS, class VT, int COLS>
__device__ __forceinline__ void load_vt(const TS* __restrict__ p, VT (&out)[COLS]) {
  if constexpr (COLS == 1) {
    out[0] = static_cast<VT>(gmem(p)[0]);
  } else {
    const PixVec<TS, COLS> v = gload_vec<TS, COLS>(p);
#pragma unroll
    for (int i = 0; i < COLS; ++i) out[i] = static_cast<VT>(v.v[i]);
  }
}

// Loads of the static-width y/z pass: float sources stream with
// non-temporal loads (measured faster for float32 passes over volumes of
// gigabytes); 8/16-bit sources keep cached loads, whose L2 hits on the halo
// rows of neighbouring segments they need.  The generic strided body, which
// the batch engine runs over many small volumes, keeps cached loads for all
// types (non-temporal ones made a batch of 256 small volumes 12% slower).
template <class TS, class VT, int COLS>
__device__ __forceinline__ void load_vt_stream(const TS* __restrict__ p, VT (&out)[COLS]) {
  if constexpr (!Pix<TS>::kInt) {
    const PixVec<TS, COLS> v = load_nt<TS, COLS>(p);
#pragma unroll
    for (int i = 0; i < COLS; ++i) out[i] = static_cast<VT>(v.v[i]);
  } else {
    load_vt<TS, VT, COLS>(p, out);
  }
}

template <int OP, bool kFloatSrc, class VT>
__device__ __forceinline__ VT tap_apply(VT acc, VT x, float h) {
  if constexpr (OP == OP_GAUSS) {
    return __builtin_fmaf(h, x, acc);
  } else if constexpr (OP == OP_MEAN) {
    return acc + x;
  } else if constexpr (OP == OP_MIN) {
    // OpenCL min(x, y) = y < x ? y : x; on integers this is the plain min.
    if constexpr (kFloatSrc) return (x < acc) ? x : acc;
    else return x < acc ? x : acc;
  } else {
    // OpenCL max(x, y) = x < y ? y : x.
    if constexpr (kFloatSrc) return (acc < x) ? x : acc;
    else return acc < x ? x : acc;
  }
}

// Saturating int32 -> pixel conversion for integer-valued min/max results.
template <class TD>
__device__ __forceinline__ TD clamp_int(int32_t v) {
  if constexpr (!Pix<TD>::kInt) {
    return static_cast<float>(v);
  } else {
    constexpr int32_t lo = static_cast<int32_t>(Pix<TD>::lo), hi = static_cast<int32_t>(Pix<TD>::hi);
    return static_cast<TD>(v < lo ? lo : (v > hi ? hi : v));
  }
}

// The divide the launch planned (p.fastdiv), resolved at compile time where
// the pixel type fixes it: float sources always take the guarded divide,
// box means of integer sources never need it, and Gaussians of integer
// sources follow the flag (set from the smallest weight).
template <int OP, class TS>
__device__ __forceinline__ float div_planned_dr(float a, float div, float rdiv, int32_t fastdiv) {
  if constexpr (!Pix<TS>::kInt) return div_uniform(a, div, rdiv);
  else if constexpr (OP == OP_MEAN) return div_uniform_fast(a, div, rdiv);
  else return fastdiv ? div_uniform_fast(a, div, rdiv) : div_uniform(a, div, rdiv);
}

template <int OP, class TS>
__device__ __forceinline__ float div_planned(float a, const SepParams& p) {
  return div_planned_dr<OP, TS>(a, p.div, p.rdiv, p.fastdiv);
}

// The pass's output value: Gaussian/mean divided as planned, min/max
// saturated, then CONVERT_dst.  (div, rdiv, fastdiv) as in SepParams.
template <int OP, class TS, class TD, class VT>
__device__ __forceinline__ TD finish_dr(VT acc, float div, float rdiv, int32_t fastdiv) {
  if constexpr (OP == OP_GAUSS || OP == OP_MEAN) {
    return convert_sat<TD>(div_planned_dr<OP, TS>(static_cast<float>(acc), div, rdiv, fastdiv));
  } else if constexpr (std::is_same<VT, int32_t>::value) {
    return clamp_int<TD>(acc);
  } else {
    return convert_sat<TD>(acc);
  }
}

template <int OP, class TS, class TD, class VT>
__device__ __forceinline__ TD finish(VT acc, const SepParams& p) {
  return finish_dr<OP, TS, TD, VT>(acc, p.div, p.rdiv, p.fastdiv);
}

// ---- y / z (strided) pass ---------------------------------------------------
template <int OP, class TS, class TD, int K, int COLS>
__device__ __forceinline__ void sep_strided_body(const SepParams& p, const uint32_t vblk) {
  using VT = vt_t<OP, TS>;
  constexpr bool kFloatSrc = !Pix<TS>::kInt;
  const uint32_t sid = vblk / p.nblk;
  const uint32_t b = vblk - sid * p.nblk;
  const int64_t g = (int64_t)b * 256 + threadIdx.x;
  if (g >= p.ncols_total) return;
  const int64_t o = g / p.ncolgroups;
  const int64_t cg = g - o * p.ncolgroups;
  const int64_t eoff = o * (p.L * p.inner) + cg * COLS;
  const char* const ls = static_cast<const char*>(p.src) + eoff * (int64_t)sizeof(TS);
  char* const ld = static_cast<char*>(p.dst) + eoff * (int64_t)sizeof(TD);
  const int64_t ss = p.inner * (int64_t)sizeof(TS);
  const int64_t ds = p.inner * (int64_t)sizeof(TD);
  const int64_t Lm1 = p.L - 1;
  const int c = p.c;
  const int ntaps = p.ntaps;
  const bool carry = ntaps <= K;
  const int64_t a0 = (int64_t)sid * p.seg;
  const int64_t a1 = (a0 + p.seg < p.L) ? a0 + p.seg : p.L;

  VT win[K][COLS];
  VT acc[K][COLS];
  auto ldc = [&](int64_t pos, VT (&dst)[COLS]) {
    load_vt<TS, VT, COLS>(reinterpret_cast<const TS*>(ls + clamp64(pos, 0, Lm1) * ss), dst);
  };

  // One block of K outputs starting at p0.  Loads in it read positions
  // p0-c .. p0+c+K; CLAMP=false when all of them are inside the axis.
  auto block = [&](int64_t p0, auto clampTag) {
    constexpr bool CLAMP = decltype(clampTag)::value;
    int64_t nxt = p0 - c + K;  // position of the next ring refill
    auto refill = [&](VT (&dst)[COLS]) {
      if constexpr (CLAMP) ldc(nxt, dst);
      else load_vt<TS, VT, COLS>(reinterpret_cast<const TS*>(ls + nxt * ss), dst);
      ++nxt;
    };
#pragma unroll
    for (int q = 0; q < K; ++q) {
      if constexpr (OP == OP_GAUSS || OP == OP_MEAN) {
#pragma unroll
        for (int j = 0; j < COLS; ++j) acc[q][j] = 0.f;
      } else if constexpr (kFloatSrc) {
        ldc(p0 + q, acc[q]);  // centre first, as the reference initialises res
      } else {
#pragma unroll
        for (int j = 0; j < COLS; ++j) acc[q][j] = win[q][j];  // any window value (exact)
      }
    }
    int v = 0;
    for (; v + K <= ntaps; v += K) {
#pragma unroll
      for (int r = 0; r < K; ++r) {
        const float h = (OP == OP_GAUSS) ? p.w[v + r] : 0.f;
#pragma unroll
        for (int q = 0; q < K; ++q)
#pragma unroll
          for (int j = 0; j < COLS; ++j)
            acc[q][j] = tap_apply<OP, kFloatSrc, VT>(acc[q][j], win[(q + r) % K][j], h);
        refill(win[r]);
      }
    }
#pragma unroll
    for (int r = 0; r < K - 1; ++r) {
      if (v + r < ntaps) {
        const float h = (OP == OP_GAUSS) ? p.w[v + r] : 0.f;
#pragma unroll
        for (int q = 0; q < K; ++q)
#pragma unroll
          for (int j = 0; j < COLS; ++j)
            acc[q][j] = tap_apply<OP, kFloatSrc, VT>(acc[q][j], win[(q + r) % K][j], h);
        refill(win[r]);
      }
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
      if (p0 + q < a1) {
        TD out[COLS];
#pragma unroll
        for (int j = 0; j < COLS; ++j) out[j] = finish<OP, TS, TD, VT>(acc[q][j], p);
        store_nt<TD, COLS>(reinterpret_cast<TD*>(ld + (p0 + q) * ds), out);
      }
    }
  };

#pragma unroll
  for (int t = 0; t < K; ++t) ldc(a0 - c + t, win[t]);
  for (int64_t p0 = a0; p0 < a1; p0 += K) {
    if (p0 - c >= 0 && p0 + c + K <= Lm1) block(p0, BoolC<false>{});
    else block(p0, BoolC<true>{});
    const int64_t pn = p0 + K;
    if (pn < a1) {
      // After the taps slot s holds input s+K (s < ntaps); with a carried
      // ring only the slots the taps did not refill need loading.
#pragma unroll
      for (int s = 0; s < K; ++s)
        if (!carry || s >= ntaps) ldc(pn - c + s, win[s]);
    }
  }
}

template <int OP, class TS, class TD, int K, int COLS>
__global__ __launch_bounds__(256) void sep_strided_kernel(const SepParams p) {
  sep_strided_body<OP, TS, TD, K, COLS>(p, blockIdx.x);
}

// ---- x pass (contiguous axis, LDS-staged row segments) ----------------------
constexpr int kXK = 8;            // outputs per lane per sub-segment
constexpr int kXSeg = 64 * kXK;   // outputs per sub-segment (one wave-wide vector load)
constexpr int kMaxCX = 1024;      // larger half widths along x use the strided kernel

__host__ __device__ __forceinline__ int lds_pad(int e) { return e + e / kXK; }
}  // namespace clk
#include "sep_xreg.h"
#include "sep_ystatic.h"
#include "sep_yring.h"
#include "sep_ydyn.h"
namespace clk {

// The taps of 8 consecutive outputs whose windows start at lb (LDS, one
// pad word per 8): a ring of 8 window registers, refilled as the taps
// advance, so each LDS value is read once per lane.
template <int OP, class TS, class VT>
__device__ __forceinline__ void x_ring(const VT* lb, const int c, const int ntaps, const SepParams& p,
                                       VT (&acc)[kXK]) {
  constexpr bool kFloatSrc = !Pix<TS>::kInt;
  constexpr int K = kXK;
  VT win[K];
#pragma unroll
  for (int t = 0; t < K; ++t) win[t] = lb[lds_pad(t)];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    if constexpr (OP == OP_GAUSS || OP == OP_MEAN) acc[q] = 0.f;
    else if constexpr (kFloatSrc) acc[q] = lb[lds_pad(q + c)];
    else acc[q] = win[q];
  }
  int vv = 0;
  for (; vv + K <= ntaps; vv += K) {
    const VT* lv = lb + (K + 1) * (vv / K);  // lds_pad(vv + r + K) = (K+1)*(vv/K) + r + K + 1
#pragma unroll
    for (int r = 0; r < K; ++r) {
      const float h = (OP == OP_GAUSS) ? p.w[vv + r] : 0.f;
#pragma unroll
      for (int q = 0; q < K; ++q) acc[q] = tap_apply<OP, kFloatSrc, VT>(acc[q], win[(q + r) % K], h);
      win[r] = lv[r + K + 1];
    }
  }
#pragma unroll
  for (int r = 0; r < K - 1; ++r) {
    if (vv + r < ntaps) {
      const float h = (OP == OP_GAUSS) ? p.w[vv + r] : 0.f;
#pragma unroll
      for (int q = 0; q < K; ++q) acc[q] = tap_apply<OP, kFloatSrc, VT>(acc[q], win[(q + r) % K], h);
      if (vv + r + 1 < ntaps) win[r] = lb[lds_pad(vv + r + K)];
    }
  }
}

// One wave filters SUB*512 consecutive outputs of one row: SUB vector loads
// per lane (all in flight together) plus c halo pixels per side, staged in
// LDS, then SUB blocks of 8 outputs per lane through the register ring.  The
// grid is one wave per task, in address order (no persistent loop).
template <int OP, class TS, class TD, int SUB>
__device__ __forceinline__ void sep_x_body(const SepParams& p, const uint32_t vblk, uint32_t* lds_raw) {
  using VT = vt_t<OP, TS>;
  constexpr int K = kXK;
  constexpr int TSEG = kXSeg * SUB;
  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  VT* const buf = reinterpret_cast<VT*>(lds_raw) + wv * p.lds_stride;
  const int64_t W = p.L;
  const int c = p.c;
  const int ntaps = p.ntaps;
  const uint32_t task = vblk * 4u + (uint32_t)wv;
  const bool valid = task < (uint32_t)p.ntasks;
  const uint32_t row = div_nseg(p, task);
  const int64_t x0 = (int64_t)(task - row * (uint32_t)p.nseg) * TSEG;
  const CLK_GLOBAL TS* __restrict__ rp = gmem(static_cast<const TS*>(p.src)) + (int64_t)row * W;

  if (valid) {
    PixVec<TS, K> v[SUB];
#pragma unroll
    for (int m = 0; m < SUB; ++m) {
      const int64_t xi = x0 + kXSeg * m + K * lane;
      if (p.vec_ld && xi + K <= W) {
        v[m] = load_nt<TS, K>(gen(rp + xi));
      } else {
#pragma unroll
        for (int j = 0; j < K; ++j) v[m].v[j] = rp[xi + j < W ? xi + j : W - 1];
      }
    }
    for (int e = lane; e < c; e += 64) {
      buf[lds_pad(e)] = static_cast<VT>(rp[clamp64(x0 - c + e, 0, W - 1)]);
      const int64_t xr = x0 + TSEG + e;
      buf[lds_pad(c + TSEG + e)] = static_cast<VT>(rp[xr < W ? xr : W - 1]);
    }
#pragma unroll
    for (int m = 0; m < SUB; ++m)
#pragma unroll
      for (int j = 0; j < K; ++j)
        buf[lds_pad(c + kXSeg * m + K * lane + j)] = static_cast<VT>(v[m].v[j]);
  }
  __syncthreads();
  if (!valid) return;
#pragma unroll
  for (int m = 0; m < SUB; ++m) {
    // lds_pad(512*m + K*lane + t) = 576*m + (K+1)*lane + lds_pad(t)
    const VT* lb = buf + (kXSeg + kXSeg / K) * m + (K + 1) * lane;
    VT acc[K];
    x_ring<OP, TS, VT>(lb, c, ntaps, p, acc);
    const int64_t xs = x0 + kXSeg * m + (int64_t)K * lane;
    if (xs < W) {
      CLK_GLOBAL TD* __restrict__ dp = gmem(static_cast<TD*>(p.dst)) + (int64_t)row * W + xs;
      TD out[K];
#pragma unroll
      for (int q = 0; q < K; ++q) out[q] = finish<OP, TS, TD, VT>(acc[q], p);
      if (p.vec_st && xs + K <= W) {
        store_nt<TD, K>(gen(dp), out);
      } else {
#pragma unroll
        for (int q = 0; q < K; ++q)
          if (xs + q < W) dp[q] = out[q];
      }
    }
  }
}

template <int OP, class TS, class TD, int SUB>
__global__ __launch_bounds__(256) void sep_x_kernel(const SepParams p) {
  extern __shared__ uint32_t lds_raw[];
  sep_x_body<OP, TS, TD, SUB>(p, blockIdx.x, lds_raw);
}

// x pass for rows of at most 512 pixels, several rows per wave.  A task is
// k = nseg consecutive rows, staged in LDS as padded rows (c edge pixels,
// the row, c edge pixels; k (W + 2c) <= 512 + 2c words, one pad word per 8).
// Lane l computes the 8 padded positions c + 8l .. c + 8l + 7 with the
// register ring; positions that fall in a halo are computed and dropped.
// Each wave uses its own LDS rows, so waves synchronise with themselves.
template <int OP, class TS, class TD>
__device__ __forceinline__ void sep_xpack_body(const SepParams& p, const uint32_t vblk, uint32_t* lds_raw) {
  using VT = vt_t<OP, TS>;
  constexpr int K = kXK;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const uint32_t task = vblk * 4u + (uint32_t)wv;
  if (task >= (uint32_t)p.ntasks) return;
  VT* const buf = reinterpret_cast<VT*>(lds_raw) + wv * p.lds_stride;
  const int W = (int)p.L;
  const int c = p.c;
  const int PW = W + 2 * c;
  const int64_t row0 = (int64_t)task * p.nseg;
  const int k = (int)(p.outer - row0 < p.nseg ? p.outer - row0 : p.nseg);
  const CLK_GLOBAL TS* __restrict__ rp0 = gmem(static_cast<const TS*>(p.src)) + row0 * W;
  for (int r = 0; r < k; ++r) {
    const CLK_GLOBAL TS* __restrict__ rp = rp0 + (int64_t)r * W;
    for (int e = lane; e < PW; e += 64) {
      int x = e - c;
      x = x < 0 ? 0 : (x >= W ? W - 1 : x);
      buf[lds_pad(r * PW + e)] = static_cast<VT>(rp[x]);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  VT acc[K];
  x_ring<OP, TS, VT>(buf + (K + 1) * lane, c, p.ntaps, p, acc);
  // position i = c + 8 lane + q is x = (i mod PW) - c of padded row i / PW
  const int i0 = c + K * lane;
  int r = i0 / PW;
  int x = i0 - r * PW - c;
  CLK_GLOBAL TD* __restrict__ dp = gmem(static_cast<TD*>(p.dst)) + row0 * W;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    if (x >= W + c) {
      ++r;
      x -= PW;
    }
    if (r < k && x >= 0 && x < W) dp[(int64_t)r * W + x] = finish<OP, TS, TD, VT>(acc[q], p);
    ++x;
  }
}

template <int OP, class TS, class TD>
__global__ __launch_bounds__(256) void sep_xpack_kernel(const SepParams p) {
  extern __shared__ uint32_t lds_raw[];
  sep_xpack_body<OP, TS, TD>(p, blockIdx.x, lds_raw);
}

// Rows per packed task and the LDS words per wave for half width c (W <= 512).
inline void plan_xpack(SepParams& p) {
  const int64_t PW = p.L + 2 * (int64_t)p.c;
  int64_t k = (kXSeg + 2 * (int64_t)p.c) / PW;
  if (k < 1) k = 1;
  p.nseg = (int32_t)k;
  p.ntasks = (p.outer + k - 1) / k;
  p.lds_stride = lds_pad(kXSeg + 2 * p.c + kXK) + 1;
}

}  // namespace clk

#include "sep_xstatic.h"

namespace clk {

// ---- host-side launch of one (op, src type, dst type) ----------------------

// Resident blocks of a kernel (launched with `threads` per block) on the
// current device, cached per kernel.
inline int resident_blocks(const void* kernel, size_t shmem, int threads = 256) {
  static std::mutex mu;
  static std::map<std::pair<const void*, size_t>, int> cache;
  static int cus = 0;
  std::lock_guard<std::mutex> lk(mu);
  if (cus == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
  }
  const auto key = std::make_pair(kernel, shmem);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int bpc = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, kernel, threads, shmem) != hipSuccess || bpc <= 0)
    bpc = 1;
  cache[key] = bpc * cus;
  return bpc * cus;
}

// Blocks of 256 threads per resident slot that a launch aims for; a few
// rounds of equal-sized blocks keep every CU busy to the end.  CLK_ROUNDS
// and CLK_STRIDED_VARIANT override the defaults (used when tuning).
inline int strided_rounds() { const int r = tuning(TUNE_STRIDED_ROUNDS); return r > 0 ? r : 16; }
inline bool no_xreg() { return tuning(TUNE_NO_XREG) != 0; }
inline int strided_variant() { return tuning(TUNE_STRIDED_VARIANT); }

// Segments along the axis for a strided launch of nblk column blocks: about
// target_blocks blocks in all, each segment a multiple of K and at least
// twice the filter long (a segment re-reads 2c halo inputs).
template <int K>
inline void plan_segments(SepParams& p, int64_t nblk, int64_t target_blocks) {
  int64_t nseg = (target_blocks + nblk - 1) / nblk;
  int64_t minseg = 2 * (int64_t)p.ntaps;
  if (minseg < K) minseg = K;
  int64_t maxseg = (p.L + minseg - 1) / minseg;
  if (nseg > maxseg) nseg = maxseg;
  if (nseg < 1) nseg = 1;
  int64_t seg = (p.L + nseg - 1) / nseg;
  seg = (seg + K - 1) / K * K;
  nseg = (p.L + seg - 1) / seg;
  p.seg = (int32_t)seg;
  set_nseg(p, nseg);
  p.nblk = (uint32_t)nblk;
}

template <int OP, class TS, class TD, int K, int COLS>
int32_t launch_strided(hipStream_t st, SepParams& p, std::string* err) {
  p.ncolgroups = p.inner / COLS;
  p.ncols_total = p.outer * p.ncolgroups;
  const int64_t nblk = (p.ncols_total + 255) / 256;
  const auto kern = sep_strided_kernel<OP, TS, TD, K, COLS>;
  plan_segments<K>(p, nblk, (int64_t)resident_blocks((const void*)kern, 0) * strided_rounds());
  const int64_t blocks = nblk * p.nseg;
  if (blocks > 0x7fffffffLL) {
    *err = "volume too large for the strided pass grid";
    return CLK_INVALID_VALUE;
  }
  hipLaunchKernelGGL(kern, dim3((uint32_t)blocks), dim3(256), 0, st, p);
  return 0;
}

template <int OP, class TS, class TD, int C, int COLS, int K>
int32_t launch_ystatic_k(hipStream_t st, SepParams& p, std::string* err) {
  p.ncolgroups = p.inner / COLS;
  p.ncols_total = p.outer * p.ncolgroups;
  const int64_t nblk = (p.ncols_total + 255) / 256;
  const auto kern = sep_ystatic_kernel<OP, TS, TD, C, COLS, K>;
  plan_segments<K>(p, nblk, (int64_t)resident_blocks((const void*)kern, 0) * strided_rounds());
  const int64_t blocks = nblk * p.nseg;
  if (blocks > 0x7fffffffLL) {
    *err = "volume too large for the strided pass grid";
    return CLK_INVALID_VALUE;
  }
  p.order = tuning(TUNE_STRIDED_ORDER);
  hipLaunchKernelGGL(kern, dim3((uint32_t)blocks), dim3(256), 0, st, p);
  return 0;
}

template <int OP, class TS, class TD, int C, int COLS>
int32_t launch_ystatic_cc(hipStream_t st, SepParams& p, std::string* err) {
  return launch_ystatic_k<OP, TS, TD, C, COLS, 8>(st, p, err);
}

// The lockstep ring pass (float -> float): K per column width so that the
// ring and a step's outputs fit the 128 registers a 1024-thread block
// allows, loads one step ahead; with the yring_deep knob, K = 4 (one
// column) or 2 and loads 8 or 4 steps ahead (twice the rows in flight).
// Segments: about rounds/8 blocks per resident slot (one block per CU
// fits), i.e. whole axes for the y passes of large volumes by default.
template <int OP, class TS, class TD, int C, int COLS, int K, int D>
int32_t launch_yring_kd(hipStream_t st, SepParams& p, std::string* err) {
  p.ncolgroups = p.inner / COLS;
  p.ncols_total = p.outer * p.ncolgroups;
  const int64_t nblk = (p.ncols_total + kYRingThreads - 1) / kYRingThreads;
  const auto kern = sep_yring_kernel<OP, TS, TD, C, COLS, K, D>;
  int64_t target = (int64_t)resident_blocks((const void*)kern, 0, kYRingThreads) * strided_rounds() / 8;
  plan_segments<K>(p, nblk, target > 0 ? target : 1);
  const int64_t blocks = nblk * p.nseg;
  if (blocks > 0x7fffffffLL) {
    *err = "volume too large for the strided pass grid";
    return CLK_INVALID_VALUE;
  }
  p.order = tuning(TUNE_STRIDED_ORDER);
  hipLaunchKernelGGL(kern, dim3((uint32_t)blocks), dim3(kYRingThreads), 0, st, p);
  return 0;
}

template <int OP, class TS, class TD, int C, int COLS>
int32_t launch_yring_cc(hipStream_t st, SepParams& p, std::string* err) {
  if (tuning(TUNE_YRING_DEEP)) return launch_yring_kd<OP, TS, TD, C, COLS, COLS == 1 ? 4 : 2, COLS == 1 ? 8 : 4>(st, p, err);
  return launch_yring_kd<OP, TS, TD, C, COLS, COLS == 1 ? 16 : (COLS == 2 ? (C <= 6 ? 8 : 4) : 4), 1>(st, p, err);
}

// The lockstep ring pass for float -> float when the yring knob asks for it
// (2) or by default (0) on y passes (outer > 1), where it measured faster
// on every placement; one column per lane there, else four where they fit.
// Returns false when it does not apply.
template <int OP, class TS, class TD, int C>
bool launch_yring_c(hipStream_t st, SepParams& p, bool v4, bool v2, bool off32_1024, int32_t* rc,
                    std::string* err) {
  if constexpr (std::is_same<TS, float>::value && std::is_same<TD, float>::value) {
    const int yr = tuning(TUNE_YRING);
    if (!off32_1024 || !(yr == 2 || (yr == 0 && p.outer > 1))) return false;
    int want = tuning(TUNE_YSTATIC_COLS);
    if (want == 0) want = p.outer > 1 ? 1 : 4;
    if constexpr (C <= 6) {
      if (want == 4 && v4) { *rc = launch_yring_cc<OP, TS, TD, C, 4>(st, p, err); return true; }
    }
    if (want >= 2 && v2) { *rc = launch_yring_cc<OP, TS, TD, C, 2>(st, p, err); return true; }
    *rc = launch_yring_cc<OP, TS, TD, C, 1>(st, p, err);
    return true;
  }
  return false;
}

// Wider column groups while the window registers stay within ~112; the
// ystatic_cols knob can ask for fewer (or, where they fit, four).
template <int OP, class TS, class TD, int C>
int32_t launch_ystatic_c(hipStream_t st, SepParams& p, bool v4, bool v2, bool off32_1024, std::string* err) {
  int32_t rrc = 0;
  if (launch_yring_c<OP, TS, TD, C>(st, p, v4, v2, off32_1024, &rrc, err)) return rrc;
  constexpr int WN = 8 + 2 * C;
  const int want = tuning(TUNE_YSTATIC_COLS);
  if constexpr ((WN + 8) * 4 <= 112) {
    if (v4 && (want == 0 || want == 4)) return launch_ystatic_cc<OP, TS, TD, C, 4>(st, p, err);
  }
  if (v2 && want != 1) return launch_ystatic_cc<OP, TS, TD, C, 2>(st, p, err);
  return launch_ystatic_cc<OP, TS, TD, C, 1>(st, p, err);
}

template <int OP, class TS, class TD>
int32_t launch_ystatic(hipStream_t st, SepParams& p, bool v4, bool v2, bool off32_1024, std::string* err) {
  if constexpr (std::is_same<TS, TD>::value || std::is_same<TD, float>::value) {
    switch (p.c) {
      case 1: return launch_ystatic_c<OP, TS, TD, 1>(st, p, v4, v2, off32_1024, err);
      case 2: return launch_ystatic_c<OP, TS, TD, 2>(st, p, v4, v2, off32_1024, err);
      case 3: return launch_ystatic_c<OP, TS, TD, 3>(st, p, v4, v2, off32_1024, err);
      case 4: return launch_ystatic_c<OP, TS, TD, 4>(st, p, v4, v2, off32_1024, err);
      case 5: return launch_ystatic_c<OP, TS, TD, 5>(st, p, v4, v2, off32_1024, err);
      case 6: return launch_ystatic_c<OP, TS, TD, 6>(st, p, v4, v2, off32_1024, err);
      case 7: return launch_ystatic_c<OP, TS, TD, 7>(st, p, v4, v2, off32_1024, err);
      case 8: return launch_ystatic_c<OP, TS, TD, 8>(st, p, v4, v2, off32_1024, err);
    }
    if constexpr (OP == OP_GAUSS) {
      switch (p.c) {
        case 9: return launch_ystatic_c<OP, TS, TD, 9>(st, p, v4, v2, off32_1024, err);
        case 10: return launch_ystatic_c<OP, TS, TD, 10>(st, p, v4, v2, off32_1024, err);
        case 11: return launch_ystatic_c<OP, TS, TD, 11>(st, p, v4, v2, off32_1024, err);
        case 12: return launch_ystatic_c<OP, TS, TD, 12>(st, p, v4, v2, off32_1024, err);
      }
    }
  }
  *err = "internal: no static-width strided kernel for this half width";
  return CLK_INVALID_VALUE;
}

// Units of 64 chunks per wave for the register-window x pass: 4 (loads of
// the next unit in flight while one is computed) when that still leaves
// several waves per SIMD slot, else 1; the xreg_units knob overrides.
inline int xreg_units(const SepParams& p) {
  const int want = tuning(TUNE_XREG_UNITS);
  if (want == 1 || want == 2 || want == 4) return want;
  const int64_t units = (p.outer * (p.L / 8) + 63) / 64;
  return units >= 4 * 65536 ? 4 : 1;
}

template <int OP, class TS, class TD, int C>
void launch_xreg_c(hipStream_t st, const SepParams& p) {
  const int R = xreg_units(p);
  const uint32_t blocks = (uint32_t)xreg_blocks(p, R);
  if (R == 4) hipLaunchKernelGGL((sep_xreg_kernel<OP, TS, TD, C, 4>), dim3(blocks), dim3(256), 0, st, p);
  else if (R == 2) hipLaunchKernelGGL((sep_xreg_kernel<OP, TS, TD, C, 2>), dim3(blocks), dim3(256), 0, st, p);
  else hipLaunchKernelGGL((sep_xreg_kernel<OP, TS, TD, C, 1>), dim3(blocks), dim3(256), 0, st, p);
}

template <int OP, class TS, class TD>
int32_t launch_xreg(hipStream_t st, const SepParams& p, std::string* err) {
  if constexpr (std::is_same<TS, TD>::value || std::is_same<TD, float>::value) {
    switch (p.c) {
      case 1: launch_xreg_c<OP, TS, TD, 1>(st, p); return 0;
      case 2: launch_xreg_c<OP, TS, TD, 2>(st, p); return 0;
      case 3: launch_xreg_c<OP, TS, TD, 3>(st, p); return 0;
      case 4: launch_xreg_c<OP, TS, TD, 4>(st, p); return 0;
      case 5: launch_xreg_c<OP, TS, TD, 5>(st, p); return 0;
      case 6: launch_xreg_c<OP, TS, TD, 6>(st, p); return 0;
      case 7: launch_xreg_c<OP, TS, TD, 7>(st, p); return 0;
      case 8: launch_xreg_c<OP, TS, TD, 8>(st, p); return 0;
    }
    if constexpr (OP == OP_GAUSS) {
      switch (p.c) {
        case 9: launch_xreg_c<OP, TS, TD, 9>(st, p); return 0;
        case 10: launch_xreg_c<OP, TS, TD, 10>(st, p); return 0;
        case 11: launch_xreg_c<OP, TS, TD, 11>(st, p); return 0;
        case 12: launch_xreg_c<OP, TS, TD, 12>(st, p); return 0;
        case 13: launch_xreg_c<OP, TS, TD, 13>(st, p); return 0;
        case 14: launch_xreg_c<OP, TS, TD, 14>(st, p); return 0;
        case 15: launch_xreg_c<OP, TS, TD, 15>(st, p); return 0;
        case 16: launch_xreg_c<OP, TS, TD, 16>(st, p); return 0;
      }
    }
  }
  *err = "internal: no register-window x kernel for this half width";
  return CLK_INVALID_VALUE;
}

// LDS x pass with the half width fixed at compile time, for the Gaussian
// widths the register pass leaves to LDS on long rows (9..16), one or two
// 512-output segments per wave.  Returns false when no such kernel exists.
template <int OP, class TS, class TD, int C>
bool launch_x_static_c(hipStream_t st, const SepParams& p, int sub, size_t shmem) {
  const int R = tuning(TUNE_X_TASKS) == 2 ? 2 : 1;
  const uint32_t blocks = (uint32_t)((p.ntasks + 4 * R - 1) / (4 * R));
  if (sub == 2 && R == 2) hipLaunchKernelGGL((sep_xs_kernel<OP, TS, TD, C, 2, 2>), dim3(blocks), dim3(256), shmem, st, p);
  else if (sub == 2) hipLaunchKernelGGL((sep_xs_kernel<OP, TS, TD, C, 2, 1>), dim3(blocks), dim3(256), shmem, st, p);
  else if (sub == 1 && R == 2) hipLaunchKernelGGL((sep_xs_kernel<OP, TS, TD, C, 1, 2>), dim3(blocks), dim3(256), shmem, st, p);
  else if (sub == 1) hipLaunchKernelGGL((sep_xs_kernel<OP, TS, TD, C, 1, 1>), dim3(blocks), dim3(256), shmem, st, p);
  else return false;
  return true;
}

template <int OP, class TS, class TD>
bool launch_x_static(hipStream_t st, const SepParams& p, int sub, size_t shmem) {
  if constexpr (OP == OP_GAUSS && (std::is_same<TS, TD>::value || std::is_same<TD, float>::value)) {
    if (tuning(TUNE_X_DYNAMIC)) return false;
    switch (p.c) {
      case 9: return launch_x_static_c<OP, TS, TD, 9>(st, p, sub, shmem);
      case 10: return launch_x_static_c<OP, TS, TD, 10>(st, p, sub, shmem);
      case 11: return launch_x_static_c<OP, TS, TD, 11>(st, p, sub, shmem);
      case 12: return launch_x_static_c<OP, TS, TD, 12>(st, p, sub, shmem);
      case 13: return launch_x_static_c<OP, TS, TD, 13>(st, p, sub, shmem);
      case 14: return launch_x_static_c<OP, TS, TD, 14>(st, p, sub, shmem);
      case 15: return launch_x_static_c<OP, TS, TD, 15>(st, p, sub, shmem);
      case 16: return launch_x_static_c<OP, TS, TD, 16>(st, p, sub, shmem);
    }
  }
  return false;
}

template <int OP, class TS, class TD>
int32_t launch_sep_typed(hipStream_t st, SepParams& p, int axis_is_x, std::string* err) {
  const int64_t V = p.outer * p.L * p.inner;
  if (V == 0) return 0;
  int32_t rc = 0;
  constexpr bool kXRegPair = std::is_same<TS, TD>::value || std::is_same<TD, float>::value;
  // wider windows (Gaussian only) on short rows, where a 512-output LDS
  // segment would leave lanes idle; long rows stage through LDS
  const int xreg_cap = (OP == OP_GAUSS) ? 16 : 8;
  const int xreg_want = tuning(TUNE_XREG_MAXC);
  const int xreg_maxc = xreg_want > 0 ? (xreg_want < xreg_cap ? xreg_want : xreg_cap)
                                      : ((OP == OP_GAUSS && p.L <= 256) ? 16 : 8);
  if (kXRegPair && axis_is_x && p.c >= 1 && p.c <= xreg_maxc && p.L % kXK == 0 &&
      (uintptr_t)p.src % (kXK * sizeof(TS)) == 0 && (uintptr_t)p.dst % (kXK * sizeof(TD)) == 0 &&
      !no_xreg()) {
    if (xreg_blocks(p, 1) > 0x7fffffffLL) {
      *err = "volume too large for the x pass grid";
      return CLK_INVALID_VALUE;
    }
    rc = launch_xreg<OP, TS, TD>(st, p, err);
  } else if (axis_is_x && p.L <= 256 && p.c <= 256 && tuning(TUNE_NO_XPACK) == 0) {
    // short rows: several rows per wave
    plan_xpack(p);
    if (p.ntasks > 0x7ffffff0LL) {
      *err = "volume too large for the x pass grid";
      return CLK_INVALID_VALUE;
    }
    const size_t shmem = (size_t)4 * p.lds_stride * 4;
    hipLaunchKernelGGL((sep_xpack_kernel<OP, TS, TD>), dim3((uint32_t)((p.ntasks + 3) / 4)), dim3(256), shmem, st,
                       p);
  } else if (axis_is_x && p.c <= kMaxCX) {
    const int64_t W = p.L;
    // larger tasks for long rows and narrow halos (more loads in flight per
    // wave, less per-task work); CLK_XSUB overrides when tuning
    int sub = tuning(TUNE_X_SUB);
    if (sub != 1 && sub != 2 && sub != 4) sub = (W >= 2 * kXSeg && p.c <= 256) ? 2 : 1;
    const int tseg = kXSeg * sub;
    set_nseg(p, (W + tseg - 1) / tseg);
    p.ntasks = p.outer * p.nseg;
    if (p.ntasks > 0x7ffffff0LL) {
      *err = "volume too large for the x pass grid";
      return CLK_INVALID_VALUE;
    }
    p.lds_stride = lds_pad(tseg + 2 * p.c) + 1;  // +1: the last ring refill reads one past
    p.vec_ld = (W % kXK == 0) && ((uintptr_t)p.src % (kXK * sizeof(TS)) == 0);
    p.vec_st = (W % kXK == 0) && ((uintptr_t)p.dst % (kXK * sizeof(TD)) == 0);
    const size_t shmem = (size_t)4 * p.lds_stride * 4;
    const uint32_t blocks = (uint32_t)((p.ntasks + 3) / 4);
    if (!launch_x_static<OP, TS, TD>(st, p, sub, shmem)) {
      if (sub == 4) hipLaunchKernelGGL((sep_x_kernel<OP, TS, TD, 4>), dim3(blocks), dim3(256), shmem, st, p);
      else if (sub == 2) hipLaunchKernelGGL((sep_x_kernel<OP, TS, TD, 2>), dim3(blocks), dim3(256), shmem, st, p);
      else hipLaunchKernelGGL((sep_x_kernel<OP, TS, TD, 1>), dim3(blocks), dim3(256), shmem, st, p);
    }
  } else {
    // vector columns when the inner extent and both pointers allow it
    const bool v4 = (p.inner % 4 == 0) && ((uintptr_t)p.src % (4 * sizeof(TS)) == 0) &&
                    ((uintptr_t)p.dst % (4 * sizeof(TD)) == 0);
    const bool v2 = (p.inner % 2 == 0) && ((uintptr_t)p.src % (2 * sizeof(TS)) == 0) &&
                    ((uintptr_t)p.dst % (2 * sizeof(TD)) == 0);
    // static-width kernel: fast pairs, half width 1..8 (Gaussian 1..12),
    // block offsets within 32 bits, axis length within int
    constexpr bool kStaticPair = std::is_same<TS, TD>::value || std::is_same<TD, float>::value;
    const int smaxc = (OP == OP_GAUSS) ? 12 : 8;
    const int64_t maxbytes = (int64_t)(sizeof(TS) > sizeof(TD) ? sizeof(TS) : sizeof(TD));
    // a block's lanes lie in at most ceil(256 / column groups) + 1 slabs
    // (column groups >= inner / 4), so the largest per-lane offset is below
    // (that many slabs - 1) * slab + inner pixels
    const int64_t ncg_min = p.inner / 4 > 0 ? p.inner / 4 : 1;
    int64_t extra = (256 + ncg_min - 1) / ncg_min;
    if (extra > p.outer - 1) extra = p.outer - 1;
    const bool off32 = (extra * p.L * p.inner + p.inner) * maxbytes < (int64_t(1) << 31);
    // the same bound for the 1024 lanes of a lockstep ring block
    int64_t extra1k = (kYRingThreads + ncg_min - 1) / ncg_min;
    if (extra1k > p.outer - 1) extra1k = p.outer - 1;
    const bool off32_1024 = (extra1k * p.L * p.inner + p.inner) * maxbytes < (int64_t(1) << 31);
    if (kStaticPair && p.c >= 1 && p.c <= smaxc && off32 && p.L < (int64_t(1) << 30) &&
        strided_variant() == 0) {
      rc = launch_ystatic<OP, TS, TD>(st, p, v4, v2, off32_1024, err);
      if (rc) return rc;
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) {
        *err = std::string("kernel launch failed: ") + hipGetErrorString(e);
        return CLK_OUT_OF_RESOURCES;
      }
      return 0;
    }
    int var = strided_variant();
    if (var <= 0) var = (p.ntaps <= 8) ? 1 : 2;
    if (var == 2 && v2) rc = launch_strided<OP, TS, TD, 16, 2>(st, p, err);
    else if (v4) rc = launch_strided<OP, TS, TD, 8, 4>(st, p, err);
    else rc = launch_strided<OP, TS, TD, 8, 1>(st, p, err);
  }
  if (rc) return rc;
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    *err = std::string("kernel launch failed: ") + hipGetErrorString(e);
    return CLK_OUT_OF_RESOURCES;
  }
  return 0;
}

template <int OP, class TS>
int32_t launch_sep_dst(hipStream_t st, SepParams& p, int axis_is_x, int dst_t, std::string* err) {
  switch (dst_t) {
    case CLK_CHAR:   return launch_sep_typed<OP, TS, int8_t>(st, p, axis_is_x, err);
    case CLK_UCHAR:  return launch_sep_typed<OP, TS, uint8_t>(st, p, axis_is_x, err);
    case CLK_SHORT:  return launch_sep_typed<OP, TS, int16_t>(st, p, axis_is_x, err);
    case CLK_USHORT: return launch_sep_typed<OP, TS, uint16_t>(st, p, axis_is_x, err);
    case CLK_FLOAT:  return launch_sep_typed<OP, TS, float>(st, p, axis_is_x, err);
  }
  *err = "invalid dst pixel type";
  return CLK_INVALID_VALUE;
}

template <int OP>
int32_t launch_sep_op(hipStream_t st, SepParams& p, int axis_is_x, int src_t, int dst_t,
                      std::string* err) {
  switch (src_t) {
    case CLK_CHAR:   return launch_sep_dst<OP, int8_t>(st, p, axis_is_x, dst_t, err);
    case CLK_UCHAR:  return launch_sep_dst<OP, uint8_t>(st, p, axis_is_x, dst_t, err);
    case CLK_SHORT:  return launch_sep_dst<OP, int16_t>(st, p, axis_is_x, dst_t, err);
    case CLK_USHORT: return launch_sep_dst<OP, uint16_t>(st, p, axis_is_x, dst_t, err);
    case CLK_FLOAT:  return launch_sep_dst<OP, float>(st, p, axis_is_x, dst_t, err);
  }
  *err = "invalid src pixel type";
  return CLK_INVALID_VALUE;
}

// One entry per op, defined in sep_<op>.hip so the four ops compile in
// parallel.
int32_t launch_sep_gauss(hipStream_t, SepParams&, int, int, int, std::string*);
int32_t launch_sep_mean(hipStream_t, SepParams&, int, int, int, std::string*);
int32_t launch_sep_min(hipStream_t, SepParams&, int, int, int, std::string*);
int32_t launch_sep_max(hipStream_t, SepParams&, int, int, int, std::string*);

}  // namespace clk
