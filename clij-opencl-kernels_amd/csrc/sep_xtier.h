// x pass of the batch engine for rows of any length and a half width known
// only at run time, up to a compile-time tier CMAX.  Included by
// batch_xtier.h.
//
// The register-window design of the single launches (sep_xreg.h): lane l
// owns one 8-pixel chunk, takes the chunks on either side from its
// neighbour lanes by DPP wave shifts, and computes its 8 outputs from
// static registers.  Two changes let it run every job of a batch:
//   - chunks are numbered per row (ceil(W/8) per row) instead of over the
//     flat volume, so a row's last chunk may be partial; the loader fills
//     its pixels beyond the row end with the row's last pixel, which is what
//     clamp-to-edge reads there, and the neighbour chunks beyond the row
//     end are filled with the same pixel (before the row start: its first
//     pixel), so no chunk straddles two rows whatever W is;
//   - the taps run over the tier's static window (t' = 0 .. 2*CMAX, weight
//     p.w[t'] as set_tier_weights stores it) and those outside
//     [CMAX - c, CMAX + c] are skipped by wave-uniform branches, in the
//     reference's order (v = -c .. c), so the results are those of the
//     single launches bit for bit.
// Whole chunks load and store with one vector access (an unaligned one
// where rows do not start 16-byte aligned); a row's partial last chunk one
// pixel at a time (clamped indices).
#pragma once

namespace clk {

// Chunk j of the row at rp (pixels 8j .. 8j+7, each clamped to the row).
template <class TS>
__device__ __forceinline__ void xt_load_chunk(const TS* __restrict__ rp, int x0, int W, bool vec, Chunk<TS>& c) {
  if (x0 + 8 <= W) {
    // a whole chunk: one vector load, aligned or not (inside the row, so
    // inside the volume)
    if (vec) load_nt_words<8 * (int)sizeof(TS)>(rp + x0, c.w);
    else load_nt_words_ua<8 * (int)sizeof(TS)>(rp + x0, c.w);
    return;
  }
  TS v[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = gmem(rp)[x0 + i < W ? x0 + i : W - 1];
  if constexpr (sizeof(TS) == 4) {
#pragma unroll
    for (int i = 0; i < 8; ++i) c.w[i] = __builtin_bit_cast(uint32_t, v[i]);
  } else if constexpr (sizeof(TS) == 2) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      c.w[i] = (uint32_t)(uint16_t)v[2 * i] | ((uint32_t)(uint16_t)v[2 * i + 1] << 16);
  } else {
#pragma unroll
    for (int i = 0; i < 2; ++i)
      c.w[i] = (uint32_t)(uint8_t)v[4 * i] | ((uint32_t)(uint8_t)v[4 * i + 1] << 8) |
               ((uint32_t)(uint8_t)v[4 * i + 2] << 16) | ((uint32_t)(uint8_t)v[4 * i + 3] << 24);
  }
}

template <class TS, int NB>
struct XtUnit {
  Chunk<TS> own, edge[NB];
  uint32_t row, j;
  bool active, live;
};

// p.L = W, p.outer = rows, p.nseg = chunks per row (with its division
// constants), p.ntasks = rows * chunks per row (< 2^31), p.vec_ld / p.vec_st:
// rows start 16-byte aligned for the source / destination.
template <class TS, int NB>
__device__ __forceinline__ void xt_load(const SepParams& p, const uint32_t unit, const int lane, XtUnit<TS, NB>& u) {
  const uint32_t per_row = (uint32_t)p.nseg;
  const uint32_t total = (uint32_t)p.ntasks;
  u.live = unit * 64u < total;  // wave-uniform
  if (!u.live) return;
  const uint32_t g = unit * 64u + (uint32_t)lane;
  u.active = g < total;
  const uint32_t gc = u.active ? g : total - 1;  // idle tail lanes mirror the last chunk
  u.row = div_nseg(p, gc);
  u.j = gc - u.row * per_row;
  const int W = (int)p.L;
  const TS* __restrict__ rp = static_cast<const TS*>(p.src) + (uint64_t)u.row * (uint64_t)W;
  const bool vec = p.vec_ld != 0;
  xt_load_chunk<TS>(rp, 8 * (int)u.j, W, vec, u.own);
#pragma unroll
  for (int k = 0; k < NB; ++k)
#pragma unroll
    for (int w = 0; w < Chunk<TS>::NW; ++w) u.edge[k].w[w] = 0u;
  // the wave's end lanes load the chunks beyond it; a chunk outside the row
  // is replaced by the clamp fill later, so it just re-reads the lane's own
  if (lane == 0 || lane == 63) {
#pragma unroll
    for (int k = 1; k <= NB; ++k) {
      const uint32_t jj = (lane == 0) ? (u.j >= (uint32_t)k ? u.j - k : u.j) : (u.j + k < per_row ? u.j + k : u.j);
      xt_load_chunk<TS>(rp, 8 * (int)jj, W, vec, u.edge[k - 1]);
    }
  }
}

template <int OP, class TS, class TD, int CMAX>
__device__ __forceinline__ void xt_compute(const SepParams& p, const XtUnit<TS, (CMAX + 7) / 8>& u) {
  using VT = xr_vt<OP, TS>;
  constexpr bool kFloatSrc = !Pix<TS>::kInt;
  constexpr int NB = (CMAX + 7) / 8;
  constexpr int WIN = 8 + 2 * CMAX;
  const uint32_t per_row = (uint32_t)p.nseg;
  const uint32_t j = u.j;
  const int c = __builtin_amdgcn_readfirstlane(p.c);
  const int tlo = CMAX - c, thi = CMAX + c;
  Chunk<TS> prev[NB], next[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k)
#pragma unroll
    for (int w = 0; w < Chunk<TS>::NW; ++w) {
      prev[k].w[w] = dpp_from_left(u.edge[k].w[w], k == 0 ? u.own.w[w] : prev[k - 1].w[w]);
      next[k].w[w] = dpp_from_right(u.edge[k].w[w], k == 0 ? u.own.w[w] : next[k - 1].w[w]);
    }
  // clamp to edge: chunks before the row start hold its first pixel, chunks
  // past its end its last pixel (pixel 7 of the row's last chunk, which the
  // loader clamped)
  if (j < (uint32_t)NB) {
    const TS first = word_px<TS>((j == 0) ? u.own.w[0] : prev[0].w[0], 0);
#pragma unroll
    for (int k = 1; k <= NB; ++k)
      if (j < (uint32_t)k) fill_chunk<TS>(first, prev[k - 1]);
  }
  if (j + NB >= per_row) {
    constexpr int NW = Chunk<TS>::NW;
    const TS last = word_px<TS>((j + 1 == per_row) ? u.own.w[NW - 1] : next[0].w[NW - 1], 7);
#pragma unroll
    for (int k = 1; k <= NB; ++k)
      if (j + k >= per_row) fill_chunk<TS>(last, next[k - 1]);
  }

  VT win[WIN];
#pragma unroll
  for (int i = 0; i < WIN; ++i) {
    const int e = i - CMAX;  // offset from the lane's chunk start
    const int ch = (e >= 0) ? e / 8 : -((-e + 7) / 8);
    const int ei = e - 8 * ch;
    TS t;
    if (ch == 0) t = u.own.px(ei);
    else if (ch < 0) t = prev[-ch - 1].px(ei);
    else t = next[ch - 1].px(ei);
    win[i] = static_cast<VT>(t);
  }
  VT acc[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    if constexpr (OP == OP_GAUSS || OP == OP_MEAN) acc[q] = 0;
    else acc[q] = win[q + CMAX];  // centre first
  }
#pragma unroll
  for (int t = 0; t <= 2 * CMAX; ++t) {
    if (t < tlo || t > thi) continue;  // wave-uniform
    const float h = (OP == OP_GAUSS) ? p.w[t] : 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if constexpr (OP == OP_GAUSS) acc[q] = __builtin_fmaf(h, win[q + t], acc[q]);
      else if constexpr (OP == OP_MEAN) acc[q] = acc[q] + win[q + t];
      else acc[q] = tap_apply<OP, kFloatSrc, VT>(acc[q], win[q + t], 0.f);
    }
  }
  TD out[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    if constexpr (OP == OP_GAUSS || OP == OP_MEAN) {
      out[q] = convert_sat<TD>(div_planned<OP, TS>(static_cast<float>(acc[q]), p));  // exact for integer sums
    } else if constexpr (std::is_same<VT, int32_t>::value) {
      out[q] = clamp_int<TD>(acc[q]);
    } else {
      out[q] = convert_sat<TD>(acc[q]);
    }
  }
  if (!u.active) return;
  const int W = (int)p.L;
  const int x0 = 8 * (int)j;
  TD* __restrict__ dp = static_cast<TD*>(p.dst) + (uint64_t)u.row * (uint64_t)W + x0;
  if (x0 + 8 <= W) {
    if (p.vec_st) store_nt<TD, 8>(dp, out);
    else store_nt_ua<TD, 8>(dp, out);
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (x0 + q < W) gmem(dp)[q] = out[q];
  }
}

template <int OP, class TS, class TD, int CMAX, int R>
__device__ __forceinline__ void sep_xtier_body(const SepParams& p, const uint32_t vblk) {
  constexpr int NB = (CMAX + 7) / 8;
  const int lane = threadIdx.x & 63;
  const uint32_t unit0 = (vblk * 4u + (threadIdx.x >> 6)) * (uint32_t)R;
  XtUnit<TS, NB> u[R];
  xt_load<TS, NB>(p, unit0, lane, u[0]);
#pragma unroll
  for (int k = 0; k < R; ++k) {
    if (k + 1 < R && u[k].live) xt_load<TS, NB>(p, unit0 + k + 1, lane, u[k + 1]);
    if (!u[k].live) return;  // whole waves only
    xt_compute<OP, TS, TD, CMAX>(p, u[k]);
  }
}

}  // namespace clk
