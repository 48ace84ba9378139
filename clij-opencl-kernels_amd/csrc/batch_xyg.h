// Fused x and y passes of the batch engine for the Gaussian and the mean
// (float temporaries): the x pass in registers, one source row per wave,
// handed to the y pass through LDS, so the x temporary never goes to memory.
// One compilation unit per op instantiates the classes (batch_xyg_<op>.hip).
//
// A block (4 waves) owns 64 consecutive 8-pixel chunks of plane rows
// (chunks numbered per row, ceil(W/8) per row, plane after plane, as the x
// tier pass numbers them) and one segment of y, and walks it K = 4 rows a
// step:
//   x phase  wave k filters row k of the step along x exactly as the x tier
//            pass does (lane l owns chunk l, takes its neighbours' chunks by
//            DPP wave shifts, lanes 0 and 63 load the chunks beyond the wave,
//            chunks outside a row become its edge pixel; taps over the static
//            window t' = 0 .. 2*CXT with those outside the job's half width
//            skipped by uniform branches, in the reference's order), rounds
//            the 8 results to float as the x pass stores them and writes them
//            to LDS;
//   y phase  thread t owns 2 of the block's 512 columns and keeps the y/z
//            tier pass's register ring over them (sep_ydyn.h): it moves the
//            step's 4 rows in from LDS and computes, divides and stores its
//            outputs.
// So the results are those of the x and y single launches bit for bit.  The
// LDS rows are double-buffered (one barrier a step), and each wave's source
// row of the next step is loaded while the current one is filtered.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "batch.h"
#include "clk_device.h"
#include "clk_internal.h"
#include "sep_impl.h"

namespace clk {

// One lane's source for one row: its chunk and, for lanes 0 and 63, the NB
// chunks beyond the wave on its side.
template <class TS, int NB>
struct XgRow {
  Chunk<TS> own, edge[NB];
};

template <int OP, class TS, int CXT, int CY>
__device__ __forceinline__ void sep_xyg_body(const XYParams& p, const uint32_t vblk) {
  using VX = xr_vt<OP, TS>;
  constexpr bool kFloatX = !Pix<TS>::kInt;
  constexpr int NB = (CXT + 7) / 8;
  constexpr int WIN = 8 + 2 * CXT;
  constexpr int K = 4;
  constexpr int COLS = 2;
  constexpr int U = (2 * K + 2 * CY + K - 1) / K;
  constexpr int RS = U * K;
  __shared__ float xres[2][K][512];

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ng = p.ng, W = p.W, H = p.H;
  const uint32_t nunit = ((uint32_t)p.ngt + 63u) / 64u;
  const uint32_t sid = vblk / nunit;
  const uint32_t unit = vblk - sid * nunit;
  const int cx = __builtin_amdgcn_readfirstlane(p.cx);
  const int cy = __builtin_amdgcn_readfirstlane(p.cy);
  const int a0 = (int)sid * p.seg;
  const int a1 = (a0 + p.seg < H) ? a0 + p.seg : H;

  // x phase: this lane's chunk
  const int gx = (int)unit * 64 + lane;
  const int gxc = gx < p.ngt ? gx : p.ngt - 1;  // idle tail lanes mirror the last chunk
  const int zx = gxc / ng;
  const int jx = gxc - zx * ng;
  const int x0 = 8 * jx;
  const TS* const zs = static_cast<const TS*>(p.src) + (int64_t)zx * H * W;
  const bool full = x0 + 8 <= W;
  const bool edge_lane = lane == 0 || lane == 63;

  auto load_chunk = [&](const TS* rp, int xs, Chunk<TS>& c) {
    if (xs + 8 <= W) {
      load_nt_words_ua<8 * (int)sizeof(TS)>(rp + xs, c.w);
    } else {
      TS v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = gmem(rp)[(xs + i < W) ? xs + i : W - 1];
      if constexpr (sizeof(TS) == 4) {
#pragma unroll
        for (int i = 0; i < 8; ++i) c.w[i] = __builtin_bit_cast(uint32_t, v[i]);
      } else if constexpr (sizeof(TS) == 2) {
#pragma unroll
        for (int i = 0; i < 4; ++i) c.w[i] = (uint32_t)(uint16_t)v[2 * i] | ((uint32_t)(uint16_t)v[2 * i + 1] << 16);
      } else {
#pragma unroll
        for (int i = 0; i < 2; ++i)
          c.w[i] = (uint32_t)(uint8_t)v[4 * i] | ((uint32_t)(uint8_t)v[4 * i + 1] << 8) |
                   ((uint32_t)(uint8_t)v[4 * i + 2] << 16) | ((uint32_t)(uint8_t)v[4 * i + 3] << 24);
      }
    }
  };
  auto load_row = [&](int y, XgRow<TS, NB>& r) {
    y = y < 0 ? 0 : (y > H - 1 ? H - 1 : y);
    const TS* const rp = zs + (int64_t)y * W;
    if (full) load_nt_words_ua<8 * (int)sizeof(TS)>(rp + x0, r.own.w);
    else load_chunk(rp, x0, r.own);
#pragma unroll
    for (int k = 0; k < NB; ++k)
#pragma unroll
      for (int w = 0; w < Chunk<TS>::NW; ++w) r.edge[k].w[w] = 0u;
    if (edge_lane) {
#pragma unroll
      for (int k = 1; k <= NB; ++k) {
        // a chunk outside the row is replaced by the edge fill: re-read the own
        const int jj = (lane == 0) ? (jx >= k ? jx - k : jx) : (jx + k < ng ? jx + k : jx);
        load_chunk(rp, 8 * jj, r.edge[k - 1]);
      }
    }
  };
  // x results of a loaded row into LDS row `dst` (8 floats at lane * 8)
  auto xrow = [&](const XgRow<TS, NB>& r, float* dst) {
    Chunk<TS> prev[NB], next[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k)
#pragma unroll
      for (int w = 0; w < Chunk<TS>::NW; ++w) {
        prev[k].w[w] = dpp_from_left(r.edge[k].w[w], k == 0 ? r.own.w[w] : prev[k - 1].w[w]);
        next[k].w[w] = dpp_from_right(r.edge[k].w[w], k == 0 ? r.own.w[w] : next[k - 1].w[w]);
      }
    if (jx < NB) {
      const TS first = word_px<TS>((jx == 0) ? r.own.w[0] : prev[0].w[0], 0);
#pragma unroll
      for (int k = 1; k <= NB; ++k)
        if (jx < k) fill_chunk<TS>(first, prev[k - 1]);
    }
    if (jx + NB >= ng) {
      constexpr int NW = Chunk<TS>::NW;
      const TS last = word_px<TS>((jx + 1 == ng) ? r.own.w[NW - 1] : next[0].w[NW - 1], 7);
#pragma unroll
      for (int k = 1; k <= NB; ++k)
        if (jx + k >= ng) fill_chunk<TS>(last, next[k - 1]);
    }
    const int tlo = CXT - cx, thi = CXT + cx;
    VX win[WIN];
#pragma unroll
    for (int i = 0; i < WIN; ++i) {
      if (i < tlo || i >= thi + 8) continue;  // wave-uniform: outside the job's window
      const int e = i - CXT;
      const int ch = (e >= 0) ? e / 8 : -((-e + 7) / 8);
      const int ei = e - 8 * ch;
      TS t;
      if (ch == 0) t = r.own.px(ei);
      else if (ch < 0) t = prev[-ch - 1].px(ei);
      else t = next[ch - 1].px(ei);
      win[i] = static_cast<VX>(t);
    }
    VX acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = 0;
#pragma unroll
    for (int t = 0; t <= 2 * CXT; ++t) {
      if (t < tlo || t > thi) continue;  // wave-uniform
      const float h = (OP == OP_GAUSS) ? p.wx[t] : 0.f;
#pragma unroll
      for (int q = 0; q < 8; ++q) acc[q] = tap_apply<OP, kFloatX, VX>(acc[q], win[q + t], h);
    }
    float o[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) o[q] = finish_dr<OP, TS, float, VX>(acc[q], p.divx, p.rdivx, p.fastx);
    float4* const d4 = reinterpret_cast<float4*>(dst + lane * 8);
    d4[0] = make_float4(o[0], o[1], o[2], o[3]);
    d4[1] = make_float4(o[4], o[5], o[6], o[7]);
  };

  // y phase: this thread's 2 columns (chunk t / 4, pixels 2 (t % 4) + {0, 1})
  const int col = 2 * (int)threadIdx.x;
  const int gy = (int)unit * 64 + (int)(threadIdx.x >> 2);
  const bool active = gy < p.ngt;
  const int gyc = active ? gy : p.ngt - 1;
  const int zy = gyc / ng;
  const int xy = 8 * (gyc - zy * ng) + (col & 7);
  float* const zd = static_cast<float*>(p.dst) + (int64_t)zy * H * W + xy;
  const bool ystore2 = xy + 2 <= W;
  const int tlo = CY - cy, thi = CY + cy;

  float ring[RS][COLS];
  auto take = [&](const float* row, float (&d)[COLS]) {
    const float2 v = *reinterpret_cast<const float2*>(row + col);
    d[0] = v.x;
    d[1] = v.y;
  };

  // The first step's window, ring rows tlo .. K + thi - 1 (row a0 - CY + i),
  // K rows (one per wave) at a time through LDS buffer 0.
  {
    const int n0 = K + 2 * cy;
    for (int bb = 0; bb < n0; bb += K) {
      if (bb > 0) __syncthreads();  // the previous batch's rows are taken
      if (bb + wave < n0) {
        XgRow<TS, NB> r;
        load_row(a0 - CY + tlo + bb + wave, r);
        xrow(r, &xres[0][wave][0]);
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < K + 2 * CY; ++i) {
        const int k = i - tlo - bb;
        if (i < tlo || i >= K + thi || k < 0 || k >= K) continue;  // uniform
        take(&xres[0][k][0], ring[i]);
      }
    }
  }

  float outs[K][COLS];
  auto compute = [&](const int u) {
    float acc[K][COLS];
#pragma unroll
    for (int q = 0; q < K; ++q)
#pragma unroll
      for (int j = 0; j < COLS; ++j) acc[q][j] = 0.f;
#pragma unroll
    for (int t = 0; t <= 2 * CY; ++t) {
      if (t < tlo || t > thi) continue;  // wave-uniform
      const float h = (OP == OP_GAUSS) ? p.wy[t] : 0.f;
#pragma unroll
      for (int q = 0; q < K; ++q)
#pragma unroll
        for (int j = 0; j < COLS; ++j) acc[q][j] = tap_apply<OP, false, float>(acc[q][j], ring[(u * K + q + t) % RS][j], h);
    }
#pragma unroll
    for (int q = 0; q < K; ++q)
#pragma unroll
      for (int j = 0; j < COLS; ++j) outs[q][j] = finish_dr<OP, float, float, float>(acc[q][j], p.divy, p.rdivy, p.fasty);
  };
  auto store_row = [&](const int y, const float (&o)[COLS]) {
    if (!active) return;
    float* const dp = zd + (int64_t)y * W;
    if (ystore2) {
      const u32x2 t = {__builtin_bit_cast(uint32_t, o[0]), __builtin_bit_cast(uint32_t, o[1])};
      __builtin_nontemporal_store(t, reinterpret_cast<CLK_GLOBAL u32x2_ua*>(gmem(dp)));
    } else if (xy < W) {
      gmem(dp)[0] = o[0];
    }
  };

  // Step n computes output rows p0 = a0 + n*K ..; its x phase filters ring
  // rows n*K + K + thi + wave (step n + 1's rows, loaded during step n - 1)
  // into LDS buffer (n + 1) % 2, taken into the ring after the barrier.
  XgRow<TS, NB> nx;
  if (a0 + K < a1) load_row(a0 - CY + K + thi + wave, nx);
  for (int s = 0;; s += U) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int n = s + u;
      const int p0 = a0 + n * K;
      if (p0 >= a1) return;  // block-uniform
      if (p0 + K < a1) {
        float* const buf = &xres[(n + 1) & 1][0][0];
        xrow(nx, buf + wave * 512);
        if (p0 + 2 * K < a1) load_row(a0 - CY + n * K + 2 * K + thi + wave, nx);
        __syncthreads();
#pragma unroll
        for (int jj = CY; jj < 2 * CY + K; ++jj) {
          const int k = jj - thi;
          if (k < 0 || k >= K) continue;  // uniform
          take(buf + k * 512, ring[(u * K + K + jj) % RS]);
        }
      }
      compute(u);
#pragma unroll
      for (int q = 0; q < K; ++q)
        if (p0 + q < a1) store_row(p0 + q, outs[q]);
    }
  }
}

template <int OP, class TS, int CXT, int CY>
__global__ __launch_bounds__(256) void batch_xyg_kernel(const XYParams* __restrict__ ps, const BatchGrid grid) {
  const BatchTile t = batch_tile(grid);
  sep_xyg_body<OP, TS, CXT, CY>(batch_param(ps, grid, t.job), t.vblk);
}

template <int OP, class TS>
inline bool xyg_go(hipStream_t st, int xtier, int ytier, const XYParams* ps, const BatchGrid grid, uint32_t n) {
#define CLK_XYG(XT, YT)                                                                                   \
  if (xtier == XT && ytier == YT) {                                                                       \
    hipLaunchKernelGGL((batch_xyg_kernel<OP, TS, XT, YT>), dim3(n), dim3(256), 0, st, ps, grid);          \
    return true;                                                                                          \
  }
  CLK_XYG(8, 5) CLK_XYG(8, 8) CLK_XYG(8, 16) CLK_XYG(16, 5) CLK_XYG(16, 8) CLK_XYG(16, 16)
#undef CLK_XYG
  return false;
}

// Launch one fused Gaussian/mean class of op OP; false when no kernel
// exists for it (sources uint8, uint16, int16, float).
template <int OP>
bool launch_xyg_op(hipStream_t st, int ts, int xtier, int ytier, const XYParams* ps, const BatchGrid grid,
                   uint32_t n) {
  switch (ts) {
    case CLK_UCHAR: return xyg_go<OP, uint8_t>(st, xtier, ytier, ps, grid, n);
    case CLK_SHORT: return xyg_go<OP, int16_t>(st, xtier, ytier, ps, grid, n);
    case CLK_USHORT: return xyg_go<OP, uint16_t>(st, xtier, ytier, ps, grid, n);
    case CLK_FLOAT: return xyg_go<OP, float>(st, xtier, ytier, ps, grid, n);
  }
  return false;
}

}  // namespace clk
