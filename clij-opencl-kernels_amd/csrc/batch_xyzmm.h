// A whole 3D minimum/maximum box filter of 16-bit integer pixels in one
// launch of the batch engine: no temporaries at all, 4 bytes of HBM traffic
// per voxel (the source read once, the result written once) against 8 for
// the fused x+y pass followed by a z pass.  Instantiated by batch_xyzmm.hip.
//
// A block (kXYZThreads threads) owns a tile of whole plane rows and walks z
// over one segment.  Thread t holds chunk j = t % ng (8 pixels) of tile row
// r = t / ng, R = kXYZThreads / ng rows a tile (its TY = R - 2cy middle rows
// are outputs, the cy rows on either side are their y halo).  Per plane:
//   1. every thread loads its chunk (one 16-byte vector, aligned or not,
//      from chunk16_start; chunk16_fix where it is used) and writes it to
//      LDS;
//   2. it reads the chunks on either side (or repeats the row's edge pixel)
//      and takes the 8 x results two at a time on 16-bit pairs
//      (v_pk_min_u16 / v_pk_max_u16, odd-offset pairs by v_alignbyte), into
//      a second LDS buffer;
//   3. output rows take the y results over the 2cy+1 rows above and below,
//      and push them into a register ring over z (the y/z tier pass's ring:
//      static slots, taps outside the job's half width skipped by uniform
//      branches), whose z results are stored.
// Two barriers a plane.  Chunks are loaded three planes ahead (a block is
// the only one on its CU, so nothing else covers a load's latency).  768
// threads (12 waves) let a wave keep up to 170 VGPRs (the ring of 2cz + 1
// planes of 8 pixels is 44 of them) with one block per CU.  Minimum and
// maximum of integers are exact in any order, so every output equals the
// chain of three single-axis launches; signed pixels rank as keys
// x ^ 0x8000.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "batch.h"
#include "batch_xymm.h"
#include "clk_device.h"
#include "clk_internal.h"
#include "sep_impl.h"

namespace clk {

template <int OP, class TS, int CZ>
__device__ __forceinline__ void sep_xyzmm_body(const MMParams& p, const uint32_t vblk) {
  static_assert(sizeof(TS) == 2, "16-bit pixels");
  constexpr uint32_t FLIP = std::is_signed<TS>::value ? 0x80008000u : 0u;
  constexpr int RS = 2 * CZ + 1;  // ring of xy planes, one new plane a step
  __shared__ u32x4 raw[kXYZThreads], xr[kXYZThreads];

  const int t = threadIdx.x;
  const int ng = p.ng, W = p.W, H = p.H, D = p.D;
  const int R = p.R;
  const uint32_t sid = vblk / (uint32_t)p.nyt;
  const int ty = (int)(vblk - sid * (uint32_t)p.nyt);
  const int cx = __builtin_amdgcn_readfirstlane(p.cx);
  const int cy = __builtin_amdgcn_readfirstlane(p.cy);
  const int cz = __builtin_amdgcn_readfirstlane(p.cz);
  const int r = t / ng;
  const int j = t - r * ng;
  const bool live = r < R;                       // thread holds a chunk
  const int yr = ty * p.TY - cy + r;             // its row (may lie outside: clamped loads)
  const int yc = yr < 0 ? 0 : (yr > H - 1 ? H - 1 : yr);
  const bool outrow = live && r >= cy && r < R - cy && yr < H;
  const int x0 = 8 * j;
  const bool full = x0 + 8 <= W;
  const int xs = chunk16_start(x0, W);  // the thread's loads start here
  const bool part = xs != x0;
  const int psh = 16 * (x0 - xs);
  const TS* const rs = static_cast<const TS*>(p.src) + (int64_t)yc * W;
  TS* const rd = static_cast<TS*>(p.dst) + (int64_t)yc * W + x0;
  const int64_t plane = (int64_t)H * W;
  const int a0 = (int)sid * p.zseg;
  const int a1 = (a0 + p.zseg < D) ? a0 + p.zseg : D;
  const int tlo = CZ - cz, thi = CZ + cz;

  auto load = [&](int z, uint32_t (&c)[4]) {
    z = z < 0 ? 0 : (z > D - 1 ? D - 1 : z);
    const TS* const rp = rs + (int64_t)z * plane;
    if (!live) {
#pragma unroll
      for (int i = 0; i < 4; ++i) c[i] = 0u;
    } else {
      load_nt_words_ua<16>(rp + xs, c);
    }
  };
  // xy results of the plane whose raw chunk is c[0]; c[1], c[2] hold the
  // next two planes' and plane zload is loaded behind them (two barriers;
  // every thread of the block calls it)
  auto xyplane = [&](uint32_t (&c)[3][4], const int zload, uint32_t (&o)[4]) {
    if (part) chunk16_fix(c[0], psh);
    raw[t] = u32x4{c[0][0] ^ FLIP, c[0][1] ^ FLIP, c[0][2] ^ FLIP, c[0][3] ^ FLIP};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      c[0][q] = c[1][q];
      c[1][q] = c[2][q];
    }
    load(zload, c[2]);
    __syncthreads();
    if (live) {
      uint32_t e[12];  // window pixels x0 - 8 .. x0 + 15 as pairs
      const u32x4 own = raw[t];
      e[4] = own.x; e[5] = own.y; e[6] = own.z; e[7] = own.w;
      if (j > 0) {
        const u32x4 l = raw[t - 1];
        e[0] = l.x; e[1] = l.y; e[2] = l.z; e[3] = l.w;
      } else {
        const uint32_t f = (e[4] & 0xffffu) * 0x10001u;  // the row's first pixel
        e[0] = e[1] = e[2] = e[3] = f;
      }
      if (j + 1 < ng) {
        const u32x4 rr = raw[t + 1];
        e[8] = rr.x; e[9] = rr.y; e[10] = rr.z; e[11] = rr.w;
      } else {
        const uint32_t f = (e[7] >> 16) * 0x10001u;  // its last (a partial chunk repeats it)
        e[8] = e[9] = e[10] = e[11] = f;
      }
      uint32_t x[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) x[q] = e[4 + q];
#pragma unroll
      for (int v = 1; v <= kXYMMMaxC; ++v) {
        if (v > cx) break;  // block-uniform
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (v % 2 == 0) {
            x[q] = mm_op<OP>(x[q], e[4 + q + v / 2]);
            x[q] = mm_op<OP>(x[q], e[4 + q - v / 2]);
          } else {
            const int a = 4 + q + (v - 1) / 2, b = 4 + q - (v + 1) / 2;
            x[q] = mm_op<OP>(x[q], __builtin_amdgcn_alignbyte(e[a + 1], e[a], 2));
            x[q] = mm_op<OP>(x[q], __builtin_amdgcn_alignbyte(e[b + 1], e[b], 2));
          }
        }
      }
      xr[t] = u32x4{x[0], x[1], x[2], x[3]};
    }
    __syncthreads();
    if (outrow) {
      const u32x4 m = xr[t];
      o[0] = m.x; o[1] = m.y; o[2] = m.z; o[3] = m.w;
#pragma unroll
      for (int dy = 1; dy <= kXYMMMaxC; ++dy) {
        if (dy > cy) break;  // block-uniform
        const u32x4 a = xr[t - dy * ng], b = xr[t + dy * ng];
        o[0] = mm_op<OP>(mm_op<OP>(o[0], a.x), b.x);
        o[1] = mm_op<OP>(mm_op<OP>(o[1], a.y), b.y);
        o[2] = mm_op<OP>(mm_op<OP>(o[2], a.z), b.z);
        o[3] = mm_op<OP>(mm_op<OP>(o[3], a.w), b.w);
      }
    }
  };
  auto store = [&](int z, const uint32_t (&o)[4]) {
    if (!outrow) return;
    TS* const dp = rd + (int64_t)z * plane;
    uint32_t v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = o[i] ^ FLIP;
    if (full) {
      const u32x4 s = {v[0], v[1], v[2], v[3]};
      __builtin_nontemporal_store(s, reinterpret_cast<CLK_GLOBAL u32x4_ua*>(gmem(dp)));
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i)
        if (x0 + i < W) gmem(dp)[i] = (TS)(uint16_t)(v[i / 2] >> (16 * (i % 2)));
    }
  };

  // ring row i holds the xy plane a0 - CZ + i; the first output needs rows
  // tlo .. thi; step n adds row n + thi + 1 (plane a0 + n + cz + 1, for
  // output n + 1) after output n
  uint32_t ring[RS][4];
  uint32_t cur[3][4];  // planes p, p + 1, p + 2 (p the next to filter)
  load(a0 - cz, cur[0]);
  load(a0 - cz + 1, cur[1]);
  load(a0 - cz + 2, cur[2]);
  for (int i = 0; i < 2 * CZ + 1; ++i) {
    // i is a run-time loop index; the ring slot is picked among static ones
    if (i > 2 * cz) break;  // block-uniform
    uint32_t o[4];
    xyplane(cur, a0 - cz + i + 3, o);
#pragma unroll
    for (int s = 0; s < 2 * CZ + 1; ++s)
      if (s == tlo + i) {
#pragma unroll
        for (int q = 0; q < 4; ++q) ring[s][q] = o[q];
      }
  }
  // cur[0] now holds plane a0 + cz + 1
  for (int s0 = 0;; s0 += RS) {
#pragma unroll
    for (int u = 0; u < RS; ++u) {
      const int n = s0 + u;
      const int z = a0 + n;
      if (z >= a1) return;  // block-uniform
      uint32_t o[4];
      if (outrow) {
#pragma unroll
        for (int q = 0; q < 4; ++q) o[q] = ring[(u + CZ) % RS][q];
#pragma unroll
        for (int tt = 0; tt <= 2 * CZ; ++tt) {
          if (tt == CZ || tt < tlo || tt > thi) continue;  // block-uniform
#pragma unroll
          for (int q = 0; q < 4; ++q) o[q] = mm_op<OP>(o[q], ring[(u + tt) % RS][q]);
        }
      }
      store(z, o);
      if (z + 1 < a1) {
        // plane z + cz + 1 into ring row n + thi + 1 (a static slot among
        // the candidates n + CZ + 1 .. n + 2 CZ + 1; it takes the slot of
        // row n + thi + 1 - RS, older than any row output n + 1 reads)
        uint32_t x[4];
        xyplane(cur, z + cz + 4, x);
#pragma unroll
        for (int jj = CZ + 1; jj <= 2 * CZ + 1; ++jj)
          if (jj == thi + 1) {
#pragma unroll
            for (int q = 0; q < 4; ++q) ring[(u + jj) % RS][q] = x[q];
          }
      }
    }
  }
}

template <int OP, class TS, int CZ>
__global__ __launch_bounds__(kXYZThreads) void batch_xyzmm_kernel(const MMParams* __restrict__ ps, const BatchGrid grid) {
  const BatchTile t = batch_tile(grid);
  sep_xyzmm_body<OP, TS, CZ>(batch_param(ps, grid, t.job), t.vblk);
}

}  // namespace clk
