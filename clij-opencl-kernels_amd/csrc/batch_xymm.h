// Fused x and y passes of the batch engine for minimum/maximum of 16-bit
// integer pixels, in registers only (no LDS, no barriers).  One launch
// computes a job's y pass straight from its source; the x pass's temporary
// (the source type) never goes to memory.  One compilation unit instantiates
// the classes (batch_xymm.hip).
//
// A lane owns one 8-pixel chunk of a plane row (chunks numbered per row,
// ceil(W/8) per row, plane after plane, as the x tier pass numbers them) and
// a wave walks one segment of y with 61 chunks; its lanes 0, 1 and 63 hold
// the chunks on either side and store nothing.  For every row a lane loads
// its chunk (one 16-byte vector, aligned or not; chunk16_start and chunk16_fix), takes the
// neighbouring
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

// The 8 pixels x0 .. x0 + 7 of a row of W >= 8 16-bit pixels, clamped to
// the row (pixels past its end repeat its last one), as 4 pairs.  Every
// lane loads one 16-byte vector from chunk16_start(x0, W) = min(x0, W - 8),
// inside the row; a row's partial last chunk then moves its pixels down by
// d = x0 - (W - 8) and fills the top with the last pixel (chunk16_fix, in
// registers, where the row is used).  No load sits in a lane-dependent
// branch and no loaded value is used at once, so the compiler counts the
// loads in flight (a branch around or right after loads makes it wait for
// all of them).
__device__ __forceinline__ int chunk16_start(int x0, int W) { return x0 + 8 <= W ? x0 : W - 8; }

// sh = 16 * (x0 - chunk16_start(x0, W)) bits, 16 .. 112 for a partial chunk
__device__ __forceinline__ void chunk16_fix(uint32_t (&c)[4], int sh) {
  // the chunk as one 128-bit value, shifted down with the last pixel
  // shifted in (shifts, no indexed registers: a select among words by index
  // is turned into a private array in scratch memory)
  typedef unsigned __int128 u128;
  const uint32_t f = (c[3] >> 16) * 0x10001u;
  const u128 x = (u128)c[0] | ((u128)c[1] << 32) | ((u128)c[2] << 64) | ((u128)c[3] << 96);
  const u128 fill = (u128)f * ((((u128)1) << 96) | (((u128)1) << 64) | (((u128)1) << 32) | 1u);
  const u128 y = (x >> sh) | (fill << (128 - sh));
#pragma unroll
  for (int k = 0; k < 4; ++k) c[k] = (uint32_t)(y >> (32 * k));
}

template <int OP, class TS, int CY>
__device__ __forceinline__ void sep_xymm_body(const XYParams& p, const uint32_t vblk) {
  static_assert(sizeof(TS) == 2, "16-bit pixels");
  constexpr int K = 2;
  constexpr int U = ((2 * K + 2 * CY + K - 1) / K + 1) / 2 * 2;  // even: prefetch buffers alternate
  constexpr int RS = U * K;
  constexpr uint32_t FLIP = std::is_signed<TS>::value ? 0x80008000u : 0u;
  const int lane = threadIdx.x & 63;
  const uint32_t nunit = ((uint32_t)p.ngt + 60u) / 61u;
  const uint32_t w = __builtin_amdgcn_readfirstlane(vblk * 4u + (threadIdx.x >> 6));
  if (w >= nunit * (uint32_t)p.nseg) return;  // whole waves (no barriers in this kernel)
  const uint32_t sid = w / nunit;
  const uint32_t unit = w - sid * nunit;
  const int ng = p.ng, W = p.W, H = p.H;
  // lanes 2..62 own chunks unit*61 + lane - 2; lanes 0, 1 and 63 hold the
  // chunks on either side and store nothing (lane 1's results are valid, as
  // its neighbours' chunks are in the wave: a partial chunk at lane 2 takes
  // its overlap from them)
  const int g = (int)unit * 61 + lane - 2;
  const bool active = lane >= 2 && lane <= 62 && g < p.ngt;
  const int gc = g < 0 ? 0 : (g < p.ngt ? g : p.ngt - 1);
  const int z = gc / ng;
  const int j = gc - z * ng;
  const int x0 = 8 * j;
  const TS* const zs = static_cast<const TS*>(p.src) + (int64_t)z * p.sz;
  const int64_t sy = p.sy;
  const int cx = __builtin_amdgcn_readfirstlane(p.cx);
  const int cy = __builtin_amdgcn_readfirstlane(p.cy);
  const int tlo = CY - cy, thi = CY + cy;
  const int a0 = (int)sid * p.seg;
  const int a1 = (a0 + p.seg < H) ? a0 + p.seg : H;

  // the lane's chunk of row y (clamped) as raw 16-bit pixel pairs, from xs
  const int xs = chunk16_start(x0, W);
  const bool part = xs != x0;
  const int psh = 16 * (x0 - xs);
  auto load_row = [&](int y, uint32_t (&c)[4]) {
    y = y < 0 ? 0 : (y > H - 1 ? H - 1 : y);
    load_nt_words_ua<16>(zs + (int64_t)y * sy + xs, c);
  };
  // the 8 x results of a loaded row as 4 key pairs
  auto xrow = [&](const uint32_t (&rl)[4], uint32_t (&o)[4]) {
    uint32_t r[4] = {rl[0], rl[1], rl[2], rl[3]};
    if (part) chunk16_fix(r, psh);
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

  // Stores: every lane issues one 16-byte buffer store per output row, at
  // pixel xs of its chunk, whatever it holds, so no store sits in a branch
  // (the compiler cannot count stores issued in a lane-dependent branch and
  // then waits for every load and store in flight).  Offsets with bit 31 set
  // lie past the image (under 2^30 bytes, checked by the planner) and the
  // hardware drops those stores: lanes that own no chunk, rows past the
  // segment.  A row's partial last chunk stores pixels W - 8 .. W - 1: its
  // own results shifted up by d, below them the left lane's last d results
  // (lane 1 or an owner), which that lane also stores, with the same values.
  const __amdgpu_buffer_rsrc_t dr = byte_rsrc(p.dst, 2 * (int64_t)W * H * p.D);
  const uint32_t lane_off = (uint32_t)(2 * ((int64_t)z * p.sz + xs)) | (active ? 0u : 0x80000000u);
  auto store_row = [&](const int y, const bool live, const uint32_t (&o)[4]) {
    uint32_t v[4], l[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[i] = o[i] ^ FLIP;
      l[i] = dpp_from_left(v[i], v[i]);
    }
    if (part) {
      typedef unsigned __int128 u128;
      const u128 vv = (u128)v[0] | ((u128)v[1] << 32) | ((u128)v[2] << 64) | ((u128)v[3] << 96);
      const u128 ll = (u128)l[0] | ((u128)l[1] << 32) | ((u128)l[2] << 64) | ((u128)l[3] << 96);
      const u128 m = (ll >> (128 - psh)) | (vv << psh);
#pragma unroll
      for (int k = 0; k < 4; ++k) v[k] = (uint32_t)(m >> (32 * k));
    }
    const uint32_t off = (lane_off + (uint32_t)(2 * (int64_t)y * sy)) | (live ? 0u : 0x80000000u);
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    __builtin_amdgcn_raw_buffer_store_b128(v4u{v[0], v[1], v[2], v[3]}, dr, off, 0, 0);
  };

  // Raw rows in flight, two steps of K: at step n buffer n % 2 holds step
  // n + 1's rows (ring rows n*K + K + thi + k, loaded at step n - 2); after
  // their x filter it takes step n + 3's rows.  Every step loads and filters
  // (rows past the segment are clamped rows, never used), so no load sits in
  // a branch.
  uint32_t nx[2][K][4];
#pragma unroll
  for (int k = 0; k < K; ++k) load_row(a0 - CY + K + thi + k, nx[0][k]);
#pragma unroll
  for (int k = 0; k < K; ++k) load_row(a0 - CY + 2 * K + thi + k, nx[1][k]);
  for (int s = 0;; s += U) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int n = s + u;
      const int p0 = a0 + n * K;
      if (p0 >= a1) return;  // wave-uniform
      uint32_t xk[K][4];
#pragma unroll
      for (int k = 0; k < K; ++k) xrow(nx[u % 2][k], xk[k]);
      // step n + 3's rows into the buffer just used
#pragma unroll
      for (int k = 0; k < K; ++k) load_row(a0 - CY + n * K + 3 * K + thi + k, nx[u % 2][k]);
      // step n + 1's new ring rows (their slots depend on cy: uniform
      // branches over register moves); the slots they take hold rows older
      // than any row step n reads (the ring spans 2K + 2cy rows)
#pragma unroll
      for (int jj = CY; jj < 2 * CY + K; ++jj) {
#pragma unroll
        for (int k = 0; k < K; ++k)
          if (jj - thi == k) {
#pragma unroll
            for (int i = 0; i < 4; ++i) ring[(u * K + K + jj) % RS][i] = xk[k][i];
          }
      }
      // y results of rows p0 .. p0 + K - 1
#pragma unroll
      for (int q = 0; q < K; ++q) {
        uint32_t o[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) o[i] = ring[(u * K + q + CY) % RS][i];
#pragma unroll
        for (int t = 0; t <= 2 * CY; ++t) {
          if (t == CY || t < tlo || t > thi) continue;  // wave-uniform
#pragma unroll
          for (int i = 0; i < 4; ++i) o[i] = mm_op<OP>(o[i], ring[(u * K + q + t) % RS][i]);
        }
        store_row(p0 + q, p0 + q < a1, o);
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
