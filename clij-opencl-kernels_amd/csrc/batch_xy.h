// Fused x and y passes of the batch engine: one launch computes the y pass
// of a job straight from its source, with the x pass done on the way in
// LDS, so the x pass's float32 (or source-type) temporary never goes to
// HBM.  A 3D filter of the batch then moves src + temporary + temporary +
// dst bytes per voxel instead of src + 2 x (two temporaries) + dst.
// One compilation unit per op instantiates the classes (batch_xy_<op>.hip).
//
// Layout of one block (256 threads).  A job's volume is seen as column
// groups (z, xg): COLS consecutive x of one plane row, ng = ceil(W/COLS)
// groups per row, numbered plane after plane; a block owns 256 consecutive
// groups (one per thread) and one segment of y.  It walks y as the y/z tier
// pass does (sep_ydyn.h: a register ring of CMAX + K .. rows, static slots,
// taps outside the job's half width skipped by uniform branches), but a
// ring row is not loaded from memory: for every K rows the block stages the
// source rows of all planes it touches in LDS, padded by cx clamped pixels
// on each side of every plane row (slot (z - z0) * rstr + x + cx holds
// src[z][y][clamp(x)], rstr = ng * COLS + 2 cx rounded up to whole COLS
// vectors), and each thread reads the window of its COLS columns as COLS-word
// vectors and runs the x filter: tap t of output q is window word q + t.  The x results are rounded to the
// chain's temporary type exactly as the x pass stores them, and the y pass
// reads them as it would read the temporary, so the results are those of
// the two single launches bit for bit:
//   Gaussian  x: acc = fma(w_t, x_t, acc) for t = 0..2cx, then / hsum
//   mean      x: acc = acc + x_t in order, then / (2cx + 1)
//   min/max   integer sources only (min/max of integers is exact in any
//             order); float sources keep the separate passes, whose
//             NaN/-0 order matters.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <type_traits>

#include "../../include/clkmi.h"
#include "batch.h"
#include "clk_device.h"
#include "clk_internal.h"
#include "sep_impl.h"

namespace clk {

// f(IntC<u>{}) for the run-time u in [0, N): one inlined copy per value.
template <int N, class F>
__device__ __forceinline__ void static_switch(const int u, F&& f) {
  if constexpr (N > 0) {
    if (u == N - 1) f(IntC<N - 1>{});
    else static_switch<N - 1>(u, f);
  }
}

template <int OP, class TS, class TT, class TD, int CMAX, int COLS>
__device__ __forceinline__ void sep_xy_body(const XYParams& p, const uint32_t vblk, const int nslot,
                                            uint32_t* lds) {
  using VX = xr_vt<OP, TS>;  // x-pass values (int32 for integer min/max/mean, else float)
  using VY = xr_vt<OP, TT>;  // y-pass values, as the y pass reads the temporary
  constexpr bool kFloatX = !Pix<TS>::kInt;
  constexpr bool kFloatY = !Pix<TT>::kInt;
  constexpr int K = kXYK;
  constexpr int NT = 2 * CMAX + 1;
  constexpr int U = (2 * K + 2 * CMAX + K - 1) / K;
  constexpr int RS = U * K;
  VX* const sbuf = reinterpret_cast<VX*>(lds);                       // [K][nslot]
  VY* const xres = reinterpret_cast<VY*>(lds + (size_t)K * nslot);   // [K][COLS][256], own slots

  const int nseg = p.nseg;
  const int b = (int)vblk / nseg;
  const int sid = (int)vblk - b * nseg;
  const int tid = threadIdx.x;
  const int ng = p.ng, ngt = p.ngt;
  const int W = p.W, H = p.H;
  const int cx = __builtin_amdgcn_readfirstlane(p.cx);
  const int cy = __builtin_amdgcn_readfirstlane(p.cy);
  const int rstr = p.rstr;
  const int g0 = b * 256;
  const int g = g0 + tid;
  const bool active = g < ngt;
  const int gl = active ? g : ngt - 1;  // idle threads mirror the last group (no stores)
  const int z = gl / ng;
  const int x0 = (gl - z * ng) * COLS;
  const int z0 = g0 / ng;
  const int glast = (g0 + 255 < ngt) ? g0 + 255 : ngt - 1;
  const int nsl = (glast / ng - z0 + 1) * rstr;  // staged words of this block's rows
  const int xb = (z - z0) * rstr + x0;            // window start of output 0

  // source offsets (within a plane row's y = 0) of the slots this thread stages
  int off[kXYSlots];
  bool own[kXYSlots];
#pragma unroll
  for (int m = 0; m < kXYSlots; ++m) {
    const int i = tid + 256 * m;
    own[m] = i < nsl;
    const int ic = own[m] ? i : 0;
    const int zz = ic / rstr;
    int x = ic - zz * rstr - cx;
    x = x < 0 ? 0 : (x > W - 1 ? W - 1 : x);
    off[m] = (z0 + zz) * H * W + x;
  }
  const TS* const src = static_cast<const TS*>(p.src);

  const int tlo = CMAX - cy, thi = CMAX + cy;
  const int a0 = sid * p.seg;
  const int a1 = (a0 + p.seg < H) ? a0 + p.seg : H;

  // ring row r is source row y = a0 - CMAX + r (clamped)
  VX v[K][kXYSlots];
  auto stage_load = [&](const int r0, const int nrows) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (k >= nrows) break;
      int y = a0 - CMAX + r0 + k;
      y = y < 0 ? 0 : (y > H - 1 ? H - 1 : y);
      const TS* const rp = src + (int64_t)y * W;
#pragma unroll
      for (int m = 0; m < kXYSlots; ++m)
        if (own[m]) v[k][m] = static_cast<VX>(gmem(rp)[off[m]]);
    }
  };
  auto stage_write = [&](const int nrows) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (k >= nrows) break;
#pragma unroll
      for (int m = 0; m < kXYSlots; ++m)
        if (own[m]) sbuf[k * nslot + tid + 256 * m] = v[k][m];
    }
  };
  // x filter of staged row k (run-time) for this thread's COLS columns:
  // the window (COLS + 2cx words from the COLS-aligned slot xb) is read as
  // COLS-word vectors, so the lanes of a wave read one contiguous span, and
  // only two of them are held at a time: tap t of output q is window word
  // q + t, in vector t / COLS or the next one
  const int nch = (2 * cx + 2 * COLS - 1) / COLS;  // vectors the job's window needs
  auto xfilter = [&](const int k) {
    const VX* const wp = sbuf + k * nslot + xb;
    using Vec = PixVec<VX, COLS>;
    Vec c0 = *reinterpret_cast<const Vec*>(wp);
    Vec c1 = c0;
    if (nch > 1) c1 = *reinterpret_cast<const Vec*>(wp + COLS);
    VX acc[COLS];
#pragma unroll
    for (int q = 0; q < COLS; ++q) {
      if constexpr (OP == OP_GAUSS || OP == OP_MEAN) acc[q] = 0;
      else acc[q] = c0.v[q];  // tap 0 (integer min/max: any order)
    }
#pragma unroll
    for (int t = 0; t <= 2 * kXYMaxC; ++t) {
      if (t > 2 * cx) break;  // wave-uniform
      if (OP == OP_GAUSS || OP == OP_MEAN || t > 0) {
        const float h = (OP == OP_GAUSS) ? p.wx[t] : 0.f;
#pragma unroll
        for (int q = 0; q < COLS; ++q) {
          const int i = t % COLS + q;
          acc[q] = tap_apply<OP, kFloatX, VX>(acc[q], i < COLS ? c0.v[i] : c1.v[i - COLS], h);
        }
      }
      if ((t + 1) % COLS == 0) {
        // tap t + 1 starts in vector (t + 1) / COLS
        c0 = c1;
        const int m2 = (t + 1) / COLS + 1;
        if (m2 < nch) c1 = *reinterpret_cast<const Vec*>(wp + m2 * COLS);
      }
    }
#pragma unroll
    for (int q = 0; q < COLS; ++q)
      xres[(k * COLS + q) * 256 + tid] =
          static_cast<VY>(finish_dr<OP, TS, TT, VX>(acc[q], p.divx, p.rdivx, p.fastx));
  };
  auto take = [&](const int k, VY (&dst)[COLS]) {
#pragma unroll
    for (int q = 0; q < COLS; ++q) dst[q] = xres[(k * COLS + q) * 256 + tid];
  };

  VY ring[RS][COLS];
  // the first step's window: ring rows tlo .. K + thi - 1, K at a time
  {
    const int n0 = K + 2 * cy;
    for (int bb = 0; bb < n0; bb += K) {
      const int nr = (n0 - bb < K) ? n0 - bb : K;
      stage_load(tlo + bb, nr);
      stage_write(nr);
      __syncthreads();
      for (int k = 0; k < nr; ++k) xfilter(k);
      __syncthreads();
#pragma unroll
      for (int i = 0; i < K + 2 * CMAX; ++i)
        if (i >= tlo + bb && i < tlo + bb + nr) take(i - tlo - bb, ring[i]);
    }
  }

  TD outs[K][COLS];
  auto compute = [&](const int u) {
    VY acc[K][COLS];
#pragma unroll
    for (int q = 0; q < K; ++q)
#pragma unroll
      for (int j = 0; j < COLS; ++j) {
        if constexpr (OP == OP_GAUSS || OP == OP_MEAN) acc[q][j] = 0;
        else acc[q][j] = ring[(u * K + q + CMAX) % RS][j];  // centre first
      }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (t < tlo || t > thi) continue;  // wave-uniform
      const float h = (OP == OP_GAUSS) ? p.wy[t] : 0.f;
#pragma unroll
      for (int q = 0; q < K; ++q)
#pragma unroll
        for (int j = 0; j < COLS; ++j)
          acc[q][j] = tap_apply<OP, kFloatY, VY>(acc[q][j], ring[(u * K + q + t) % RS][j], h);
    }
#pragma unroll
    for (int q = 0; q < K; ++q)
#pragma unroll
      for (int j = 0; j < COLS; ++j) outs[q][j] = finish_dr<OP, TT, TD, VY>(acc[q][j], p.divy, p.rdivy, p.fasty);
  };
  TD* const dz = static_cast<TD*>(p.dst) + (int64_t)z * H * W + x0;
  const bool vst = p.vec_st != 0;
  auto store_row = [&](const int y, const TD (&o)[COLS]) {
    if (!active) return;
    TD* const dp = dz + (int64_t)y * W;
    if (vst) {
      store_nt<TD, COLS>(dp, o);
    } else {
#pragma unroll
      for (int j = 0; j < COLS; ++j)
        if (x0 + j < W) gmem(dp)[j] = o[j];
    }
  };

  // Steps run U at a time, unrolled, so every ring slot is a static
  // register; step (s + u) computes output rows p0 = a0 + (s + u) * K ..,
  // and its new rows, ring rows (s + u) * K + K + thi + k (k < K), are staged
  // one step ahead.
  if (a0 + K < a1) stage_load(K + thi, K);
  for (int s = 0;; s += U) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int p0 = a0 + (s + u) * K;
      if (p0 >= a1) return;  // block-uniform
      if (p0 + K < a1) {
        stage_write(K);
        __syncthreads();
        if (p0 + 2 * K < a1) stage_load((s + u) * K + 2 * K + thi, K);  // in flight over this step
        for (int k = 0; k < K; ++k) xfilter(k);
        __syncthreads();
#pragma unroll
        for (int j = CMAX; j < 2 * CMAX + K; ++j)
          if (j >= thi && j < thi + K) take(j - thi, ring[(u * K + K + j) % RS]);
        compute(u);
#pragma unroll
        for (int q = 0; q < K; ++q) store_row(p0 + q, outs[q]);
      } else {
        compute(u);
#pragma unroll
        for (int q = 0; q < K; ++q)
          if (p0 + q < a1) store_row(p0 + q, outs[q]);
      }
    }
  }
}

template <int OP, class TS, class TT, int CMAX, int COLS>
__global__ __launch_bounds__(256) void batch_xy_kernel(const XYParams* __restrict__ ps, const BatchGrid grid,
                                                       const int nslot) {
  extern __shared__ uint32_t lds[];
  const BatchTile t = batch_tile(grid);
  sep_xy_body<OP, TS, TT, TT, CMAX, COLS>(batch_param(ps, grid, t.job), t.vblk, nslot, lds);
}

// The chain's temporary type for a fused job's source: float for the
// Gaussian and the mean, the source type for (integer) min/max.
template <int OP, class TS>
using xy_tt = std::conditional_t<(OP == OP_MIN || OP == OP_MAX), TS, float>;

template <int OP, class TS>
inline bool xy_go(hipStream_t st, int tier, int cols, const XYParams* ps, const BatchGrid grid, uint32_t n,
                  size_t lds, int nslot) {
  using TT = xy_tt<OP, TS>;
#define CLK_XY(T, C)                                                                                    \
  if (tier == T && cols == C) {                                                                         \
    hipLaunchKernelGGL((batch_xy_kernel<OP, TS, TT, T, C>), dim3(n), dim3(256), lds, st, ps, grid, nslot); \
    return true;                                                                                        \
  }
  CLK_XY(5, 4) CLK_XY(5, 2) CLK_XY(5, 1) CLK_XY(8, 2) CLK_XY(8, 1) CLK_XY(16, 2) CLK_XY(16, 1)
#undef CLK_XY
  return false;
}

// Launch one fused class of op OP (source type ts); false when no kernel
// exists for it.
template <int OP>
bool launch_xy_op(hipStream_t st, int ts, int tier, int cols, const XYParams* ps, const BatchGrid grid,
                  uint32_t n, size_t lds, int nslot) {
  constexpr bool kMinMax = (OP == OP_MIN || OP == OP_MAX);
  switch (ts) {
    case CLK_CHAR: return xy_go<OP, int8_t>(st, tier, cols, ps, grid, n, lds, nslot);
    case CLK_UCHAR: return xy_go<OP, uint8_t>(st, tier, cols, ps, grid, n, lds, nslot);
    case CLK_SHORT: return xy_go<OP, int16_t>(st, tier, cols, ps, grid, n, lds, nslot);
    case CLK_USHORT: return xy_go<OP, uint16_t>(st, tier, cols, ps, grid, n, lds, nslot);
    case CLK_FLOAT:
      if constexpr (!kMinMax) return xy_go<OP, float>(st, tier, cols, ps, grid, n, lds, nslot);
      break;
  }
  return false;
}

}  // namespace clk
