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
//    compile-time register; each input is loaded once per K outputs, and the
//    refill is issued K taps before its use (a software pipeline).
//  * y/z passes (and an out-of-range dim): lanes cover consecutive x
//    columns, COLS columns per lane with one vector load, so every load of a
//    wave is a contiguous run; the axis position is wave-uniform and its
//    clamp is scalar work.
//  * x pass: a wave stages its row segment (64*K outputs + 2c halo) in LDS
//    once, converted to float, with one pad word every K elements so that the
//    lane-strided ring reads are free of bank conflicts.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "clk_device.h"
#include "clk_internal.h"

namespace clk {

template <int OP, bool kFloatSrc>
__device__ __forceinline__ float tap_apply(float acc, float x, float h) {
  if constexpr (OP == OP_GAUSS) {
    return __builtin_fmaf(h, x, acc);
  } else if constexpr (OP == OP_MEAN) {
    return acc + x;
  } else if constexpr (OP == OP_MIN) {
    // OpenCL min(x, y) = y < x ? y : x.  For integer sources fminf is the
    // same value (no NaN, no signed zero).
    if constexpr (kFloatSrc) return (x < acc) ? x : acc;
    else return fminf(acc, x);
  } else {
    // OpenCL max(x, y) = x < y ? y : x.
    if constexpr (kFloatSrc) return (acc < x) ? x : acc;
    else return fmaxf(acc, x);
  }
}

template <int OP, class TD>
__device__ __forceinline__ TD finish(float acc, const SepParams& p) {
  if constexpr (OP == OP_GAUSS || OP == OP_MEAN) {
    return convert_sat<TD>(div_uniform(acc, p.div, p.rdiv));
  } else {
    return convert_sat<TD>(acc);
  }
}

// ---- y / z (strided) pass ---------------------------------------------------
template <int OP, class TS, class TD, int K, int COLS>
__global__ __launch_bounds__(256) void sep_strided_kernel(const SepParams p) {
  constexpr bool kFloatSrc = !Pix<TS>::kInt;
  const uint32_t nblk = (uint32_t)((p.outer * p.ncolgroups + 255) / 256);
  const uint32_t sid = blockIdx.x / nblk;
  const uint32_t b = blockIdx.x - sid * nblk;
  const int64_t g = (int64_t)b * 256 + threadIdx.x;
  if (g >= p.outer * p.ncolgroups) return;
  const int64_t o = g / p.ncolgroups;
  const int64_t cg = g - o * p.ncolgroups;
  const int64_t inner = p.inner;
  const int64_t base = o * (p.L * inner) + cg * COLS;
  const TS* __restrict__ sp = static_cast<const TS*>(p.src) + base;
  TD* __restrict__ dp = static_cast<TD*>(p.dst) + base;
  const int64_t Lm1 = p.L - 1;
  const int c = p.c;
  const int ntaps = p.ntaps;
  const int64_t a0 = (int64_t)sid * p.seg;
  const int64_t a1 = (a0 + p.seg < p.L) ? a0 + p.seg : p.L;

  for (int64_t p0 = a0; p0 < a1; p0 += K) {
    float win[K][COLS];
    float acc[K][COLS];
#pragma unroll
    for (int t = 0; t < K; ++t)
      load_widen<TS, COLS>(sp + clamp64(p0 - c + t, 0, Lm1) * inner, win[t]);
#pragma unroll
    for (int q = 0; q < K; ++q) {
      if constexpr (OP == OP_GAUSS || OP == OP_MEAN) {
#pragma unroll
        for (int j = 0; j < COLS; ++j) acc[q][j] = 0.f;
      } else if constexpr (kFloatSrc) {
        // centre value first, exactly as the reference initialises res
        load_widen<TS, COLS>(sp + clamp64(p0 + q, 0, Lm1) * inner, acc[q]);
      } else {
#pragma unroll
        for (int j = 0; j < COLS; ++j) acc[q][j] = win[q][j];
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
            acc[q][j] = tap_apply<OP, kFloatSrc>(acc[q][j], win[(q + r) % K][j], h);
        load_widen<TS, COLS>(sp + clamp64(p0 - c + v + r + K, 0, Lm1) * inner, win[r]);
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
            acc[q][j] = tap_apply<OP, kFloatSrc>(acc[q][j], win[(q + r) % K][j], h);
        if (v + r + 1 < ntaps)
          load_widen<TS, COLS>(sp + clamp64(p0 - c + v + r + K, 0, Lm1) * inner, win[r]);
      }
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
      if (p0 + q < a1) {
        TD out[COLS];
#pragma unroll
        for (int j = 0; j < COLS; ++j) out[j] = finish<OP, TD>(acc[q][j], p);
        store_vec<TD, COLS>(dp + (p0 + q) * inner, out);
      }
    }
  }
}

// ---- x pass (contiguous axis, LDS-staged rows) ------------------------------
template <int K>
__host__ __device__ __forceinline__ int lds_pad(int e) { return e + e / K; }

template <int OP, class TS, class TD, int K>
__global__ __launch_bounds__(256) void sep_x_kernel(const SepParams p) {
  constexpr bool kFloatSrc = !Pix<TS>::kInt;
  extern __shared__ float lds[];
  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t task = (int64_t)blockIdx.x * 4 + wv;
  const bool active = task < p.ntasks;
  const int64_t row = task / p.nseg;
  const int64_t seg = task - row * p.nseg;
  const int64_t W = p.L;
  const int c = p.c;
  const int ntaps = p.ntaps;
  const int span = 64 * K + 2 * c;
  float* buf = lds + wv * p.lds_stride;
  const int64_t x0 = seg * (64 * K);
  if (active) {
    const TS* __restrict__ rp = static_cast<const TS*>(p.src) + row * W;
    for (int e = lane; e < span; e += 64) {
      const int64_t x = clamp64(x0 - c + e, 0, W - 1);
      buf[lds_pad<K>(e)] = static_cast<float>(rp[x]);
    }
  }
  __syncthreads();
  if (!active) return;
  const float* lb = buf + (K + 1) * lane;  // lds_pad(K*lane + t) = (K+1)*lane + lds_pad(t)

  float win[K];
  float acc[K];
#pragma unroll
  for (int t = 0; t < K; ++t) win[t] = lb[lds_pad<K>(t)];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    if constexpr (OP == OP_GAUSS || OP == OP_MEAN) acc[q] = 0.f;
    else if constexpr (kFloatSrc) acc[q] = lb[lds_pad<K>(q + c)];
    else acc[q] = win[q];
  }
  int v = 0;
  for (; v + K <= ntaps; v += K) {
    const float* lv = lb + (K + 1) * (v / K);  // lds_pad(v + r + K) = (K+1)*(v/K) + r + K + 1
#pragma unroll
    for (int r = 0; r < K; ++r) {
      const float h = (OP == OP_GAUSS) ? p.w[v + r] : 0.f;
#pragma unroll
      for (int q = 0; q < K; ++q) acc[q] = tap_apply<OP, kFloatSrc>(acc[q], win[(q + r) % K], h);
      win[r] = lv[r + K + 1];
    }
  }
#pragma unroll
  for (int r = 0; r < K - 1; ++r) {
    if (v + r < ntaps) {
      const float h = (OP == OP_GAUSS) ? p.w[v + r] : 0.f;
#pragma unroll
      for (int q = 0; q < K; ++q) acc[q] = tap_apply<OP, kFloatSrc>(acc[q], win[(q + r) % K], h);
      if (v + r + 1 < ntaps) win[r] = lb[lds_pad<K>(v + r + K)];
    }
  }
  const int64_t xs = x0 + (int64_t)K * lane;
  TD* __restrict__ dp = static_cast<TD*>(p.dst) + row * W + xs;
  TD out[K];
#pragma unroll
  for (int q = 0; q < K; ++q) out[q] = finish<OP, TD>(acc[q], p);
  if (p.vec_ok && xs + K <= W) {
    store_vec<TD, K>(dp, out);
  } else {
#pragma unroll
    for (int q = 0; q < K; ++q)
      if (xs + q < W) dp[q] = out[q];
  }
}

// ---- host-side launch of one (op, src type, dst type) ----------------------
constexpr int kStridedK = 8;
constexpr int kXK = 8;
constexpr int kMaxCX = 1024;  // larger half widths along x use the strided kernel

template <int OP, class TS, class TD>
int32_t launch_sep_typed(hipStream_t st, SepParams& p, int axis_is_x, std::string* err) {
  const int64_t V = p.outer * p.L * p.inner;
  if (V == 0) return 0;
  if (axis_is_x && p.c <= kMaxCX) {
    constexpr int K = kXK;
    const int64_t W = p.L;
    const int64_t rows = p.outer;
    p.nseg = (int32_t)((W + 64 * K - 1) / (64 * K));
    p.ntasks = rows * p.nseg;
    const int span = 64 * K + 2 * p.c;
    // one spare word: the last ring refill of the last lane reads one
    // element past the window (its value is never used)
    p.lds_stride = lds_pad<K>(span) + 1;
    p.vec_ok = (W % K == 0) && ((uintptr_t)p.dst % (K * sizeof(TD)) == 0);
    const int64_t blocks = (p.ntasks + 3) / 4;
    const size_t shmem = (size_t)4 * p.lds_stride * sizeof(float);
    if (blocks > 0x7fffffffLL) { *err = "volume too large for the x pass grid"; return CLK_INVALID_VALUE; }
    hipLaunchKernelGGL((sep_x_kernel<OP, TS, TD, K>), dim3((uint32_t)blocks), dim3(256), shmem, st, p);
  } else {
    constexpr int K = kStridedK;
    const bool vec4 = (p.inner % 4 == 0) && ((uintptr_t)p.src % (4 * sizeof(TS)) == 0) &&
                      ((uintptr_t)p.dst % (4 * sizeof(TD)) == 0);
    const int cols = vec4 ? 4 : 1;
    p.ncolgroups = p.inner / cols;
    const int64_t threads_axis = p.outer * p.ncolgroups;
    // enough segments along the axis to give the chip ~1M threads
    const int64_t target = int64_t(1) << 20;
    int64_t nseg = (target + threads_axis - 1) / threads_axis;
    const int64_t maxseg = (p.L + K - 1) / K;
    if (nseg > maxseg) nseg = maxseg;
    if (nseg < 1) nseg = 1;
    int64_t seg = (p.L + nseg - 1) / nseg;
    seg = (seg + K - 1) / K * K;
    nseg = (p.L + seg - 1) / seg;
    p.seg = (int32_t)seg;
    p.nseg = (int32_t)nseg;
    const int64_t nblk = (threads_axis + 255) / 256;
    const int64_t blocks = nblk * nseg;
    if (blocks > 0x7fffffffLL || nblk > 0xffffffffLL) {
      *err = "volume too large for the strided pass grid";
      return CLK_INVALID_VALUE;
    }
    if (vec4)
      hipLaunchKernelGGL((sep_strided_kernel<OP, TS, TD, K, 4>), dim3((uint32_t)blocks), dim3(256), 0, st, p);
    else
      hipLaunchKernelGGL((sep_strided_kernel<OP, TS, TD, K, 1>), dim3((uint32_t)blocks), dim3(256), 0, st, p);
  }
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
