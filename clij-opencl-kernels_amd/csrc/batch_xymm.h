// Fused x and y passes of the batch engine for minimum/maximum of 16-bit
// integer pixels, in registers only (no LDS, no barriers).  One launch
// computes a job's y pass straight from its source; the x pass's temporary
// (the source type) never goes to memory.  One compilation unit instantiates
// the classes (batch_xymm.hip).
//
// A lane owns one 8-pixel chunk of a plane row (chunks numbered per row,
// ceil(W/8) per row, plane after plane, as the x tier pass numbers them) and
// a wave walks one segment of y with 62 chunks; its lanes 0 and 63 hold
// the chunks on either side and store nothing.  For every row a lane loads
// its chunk (one 16-byte vector for a whole chunk, aligned or not; pixel by
// pixel, clamped, for a row's partial last chunk), takes the neighbouring
// chunks from lanes l-1 and l+1 by DPP wave shifts (chunks outside the row
// are replaced by the row's edge pixel), and computes the 8 x results two at a time on 16-bit pairs
// (v_pk_min_u16 / v_pk_max_u16; pairs that start at an odd pixel are formed
// once per row by v_alignbyte).  The results enter a register ring of rows
// as in the y/z tier pass (sep_ydyn.h: static slots, taps outside the job's
// half width skipped by wave-uniform branches), whose y results are stored.
// Signed pixels are ranked as keys x ^ 0x8000.  Minimum and maximum of
// integers are exact in any order, so the results are those of the x and y
// single launches bit for bit.
//
// The same kernel runs a z pass (half width 0 along x, no x filter): the
// walk then goes over planes (row stride H*W) and its outer slices are the
// rows of a plane (stride W).
//
// Rows are loaded two steps ahead of their x filter: at step n the rows of
// step n + 1 (loaded at step n - 2) are filtered along x into the ring and
// the rows of step n + 3 are loaded, so two steps of compute cover their
// latency.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "batch.h"
#include "clk_device.h"
#include "clk_internal.h"
#include "sep_impl.h"

namespace clk {

typedef unsigned short mm_us2 __attribute__((ext_vector_type(2)));

template <int OP>
__device__ __forceinline__ uint32_t mm_op(uint32_t a, uint32_t b) {
  const mm_us2 x = __builtin_bit_cast(mm_us2, a), y = __builtin_bit_cast(mm_us2, b);
  if constexpr (OP == OP_MIN) return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(x, y));
  else return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(x, y));
}

template <int OP, class TS, int CY>
__device__ __forceinline__ void sep_xymm_body(const XYParams& p, const uint32_t vblk) {
  static_assert(sizeof(TS) == 2, "16-bit pixels");
  constexpr int K = 2;
  constexpr int U = ((2 * K + 2 * CY + K - 1) / K + 1) / 2 * 2;  // even: prefetch buffers alternate
  constexpr int RS = U * K;
  constexpr uint32_t FLIP = std::is_signed<TS>::value ? 0x80008000u : 0u;
  const int lane = threadIdx.x & 63;
  const uint32_t nunit = ((uint32_t)p.ngt + 61u) / 62u;
  const uint32_t w = __builtin_amdgcn_readfirstlane(vblk * 4u + (threadIdx.x >> 6));
  if (w >= nunit * (uint32_t)p.nseg) return;  // whole waves (no barriers in this kernel)
  const uint32_t sid = w / nunit;
  const uint32_t unit = w - sid * nunit;
  const int ng = p.ng, W = p.W, H = p.H;
  // lanes 1..62 own chunks unit*62 + lane - 1; lanes 0 and 63 hold the
  // chunks on either side (their neighbours' halo) and store nothing
  const int g = (int)unit * 62 + lane - 1;
  const bool active = lane >= 1 && lane <= 62 && g < p.ngt;
  const int gc = g < 0 ? 0 : (g < p.ngt ? g : p.ngt - 1);
  const int z = gc / ng;
  const int j = gc - z * ng;
  const int x0 = 8 * j;
  const TS* const zs = static_cast<const TS*>(p.src) + (int64_t)z * p.sz;
  TS* const zd = static_cast<TS*>(p.dst) + (int64_t)z * p.sz + x0;
  const int64_t sy = p.sy;
  const int cx = __builtin_amdgcn_readfirstlane(p.cx);
  const int cy = __builtin_amdgcn_readfirstlane(p.cy);
  const int tlo = CY - cy, thi = CY + cy;
  const int a0 = (int)sid * p.seg;
  const int a1 = (a0 + p.seg < H) ? a0 + p.seg : H;
  const bool full = x0 + 8 <= W;

  // the lane's chunk of row y (clamped) as raw 16-bit pixel pairs
  auto load_row = [&](int y, uint32_t (&c)[4]) {
    y = y < 0 ? 0 : (y > H - 1 ? H - 1 : y);
    const TS* const rp = zs + (int64_t)y * sy;
    if (full) {
      load_nt_words_ua<16>(rp + x0, c);
    } else {
      uint16_t v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = (uint16_t)gmem(rp)[(x0 + i < W) ? x0 + i : W - 1];
#pragma unroll
      for (int i = 0; i < 4; ++i) c[i] = (uint32_t)v[2 * i] | ((uint32_t)v[2 * i + 1] << 16);
    }
  };
  // the 8 x results of a loaded row as 4 key pairs
  auto xrow = [&](const uint32_t (&r)[4], uint32_t (&o)[4]) {
    if (cx == 0) {  // a z pass: rows as they are
#pragma unroll
      for (int i = 0; i < 4; ++i) o[i] = r[i] ^ FLIP;
      return;
    }
    uint32_t e[12];  // window pixels x0 - 8 .. x0 + 15 as pairs (2i, 2i + 1)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t own = r[i] ^ FLIP;
      e[4 + i] = own;
      e[i] = dpp_from_left(own, own);
      e[8 + i] = dpp_from_right(own, own);
    }
    // clamp to edge: the row's first chunk has no left neighbour, its last
    // chunk (pixels past the row end already hold the last pixel) no right one
    if (j == 0) {
      const uint32_t f = (e[4] & 0xffffu) * 0x10001u;
#pragma unroll
      for (int i = 0; i < 4; ++i) e[i] = f;
    }
    if (j + 1 == ng) {
      const uint32_t f = (e[7] >> 16) * 0x10001u;
#pragma unroll
      for (int i = 0; i < 4; ++i) e[8 + i] = f;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) o[q] = e[4 + q];
#pragma unroll
    for (int v = 1; v <= 8; ++v) {
      if (v > cx) break;  // wave-uniform
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        // pairs starting at window pixels 8 + 2q + v and 8 + 2q - v; an odd
        // start takes the high half of one pair and the low half of the next
        if (v % 2 == 0) {
          o[q] = mm_op<OP>(o[q], e[4 + q + v / 2]);
          o[q] = mm_op<OP>(o[q], e[4 + q - v / 2]);
        } else {
          const int a = 4 + q + (v - 1) / 2, b = 4 + q - (v + 1) / 2;
          o[q] = mm_op<OP>(o[q], __builtin_amdgcn_alignbyte(e[a + 1], e[a], 2));
          o[q] = mm_op<OP>(o[q], __builtin_amdgcn_alignbyte(e[b + 1], e[b], 2));
        }
      }
    }
  };

  uint32_t ring[RS][4];
  // the first step's window: ring rows tlo .. K + thi - 1 (row a0 - CY + i)
#pragma unroll
  for (int i = 0; i < K + 2 * CY; ++i)
    if (i >= tlo && i < K + thi) {
      uint32_t r[4];
      load_row(a0 - CY + i, r);
      xrow(r, ring[i]);
    }

  auto store_row = [&](const int y, const uint32_t (&o)[4]) {
    if (!active) return;
    TS* const dp = zd + (int64_t)y * sy;
    uint32_t v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = o[i] ^ FLIP;
    if (full) {
      const u32x4 t = {v[0], v[1], v[2], v[3]};
      __builtin_nontemporal_store(t, reinterpret_cast<CLK_GLOBAL u32x4_ua*>(gmem(dp)));
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i)
        if (x0 + i < W) gmem(dp)[i] = (TS)(uint16_t)(v[i / 2] >> (16 * (i % 2)));
    }
  };

  // Raw rows in flight, two steps of K: at step n buffer n % 2 holds step
  // n + 1's rows (ring rows n*K + K + thi + k, loaded at step n - 2); after
  // their x filter it takes step n + 3's rows.
  uint32_t nx[2][K][4];
  if (a0 + K < a1) {
#pragma unroll
    for (int k = 0; k < K; ++k) load_row(a0 - CY + K + thi + k, nx[0][k]);
  }
  if (a0 + 2 * K < a1) {
#pragma unroll
    for (int k = 0; k < K; ++k) load_row(a0 - CY + 2 * K + thi + k, nx[1][k]);
  }
  for (int s = 0;; s += U) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int n = s + u;
      const int p0 = a0 + n * K;
      if (p0 >= a1) return;  // wave-uniform
      if (p0 + K < a1) {
        uint32_t xk[K][4];
#pragma unroll
        for (int k = 0; k < K; ++k) xrow(nx[u % 2][k], xk[k]);
        // step n + 3's rows into the buffer just used
        if (p0 + 3 * K < a1) {
#pragma unroll
          for (int k = 0; k < K; ++k) load_row(a0 - CY + n * K + 3 * K + thi + k, nx[u % 2][k]);
        }
#pragma unroll
        for (int jj = CY; jj < 2 * CY + K; ++jj) {
#pragma unroll
          for (int k = 0; k < K; ++k)
            if (jj - thi == k) {
#pragma unroll
              for (int i = 0; i < 4; ++i) ring[(u * K + K + jj) % RS][i] = xk[k][i];
            }
        }
      }
      // y results of rows p0 .. p0 + K - 1
#pragma unroll
      for (int q = 0; q < K; ++q) {
        if (p0 + q >= a1) break;
        uint32_t o[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) o[i] = ring[(u * K + q + CY) % RS][i];
#pragma unroll
        for (int t = 0; t <= 2 * CY; ++t) {
          if (t == CY || t < tlo || t > thi) continue;  // wave-uniform
#pragma unroll
          for (int i = 0; i < 4; ++i) o[i] = mm_op<OP>(o[i], ring[(u * K + q + t) % RS][i]);
        }
        store_row(p0 + q, o);
      }
    }
  }
}

template <int OP, class TS, int CY>
__global__ __launch_bounds__(256) void batch_xymm_kernel(const XYParams* __restrict__ ps, const BatchGrid grid) {
  const BatchTile t = batch_tile(grid);
  sep_xymm_body<OP, TS, CY>(batch_param(ps, grid, t.job), t.vblk);
}

}  // namespace clk
