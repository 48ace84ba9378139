// x pass of the batch engine for rows of any length and a half width known
// only at run time, up to a compile-time tier CMAX.  Included by
// batch_xtier.h.
//
// The register-window design of the single launches (sep_xreg.h): lane l
// owns one 8-pixel chunk, takes the chunks on either side from its
// neighbour lanes by DPP wave shifts, and computes its 8 outputs from
// static registers.  Two changes let it run every job of a batch:
//   - chunks are numbered per row (ceil(W/8) per row) instead of over the
//     flat volume, so a row's last chunk may be partial; its pixels beyond
//     the row end hold the row's last pixel, which is what clamp-to-edge
//     reads there, and the neighbour chunks beyond the row
//     end are filled with the same pixel (before the row start: its first
//     pixel), so no chunk straddles two rows whatever W is;
//   - the taps run over the tier's static window (t' = 0 .. 2*CMAX, weight
//     p.w[t'] as set_tier_weights stores it) and those outside
//     [CMAX - c, CMAX + c] are skipped by wave-uniform branches, in the
//     reference's order (v = -c .. c), so the results are those of the
//     single launches bit for bit.
// A wave owns 64 - 2NB chunks (NB = chunks a window reaches on either side);
// its NB lanes at each end hold the chunks next to them and store nothing,
// so every lane loads exactly its own chunk: one vector from min(8j, W - 8)
// (rows of at least 8 pixels; inside the row, aligned or not), and a row's
// partial last chunk is moved into place in registers where it is used.
// No load sits in a lane-dependent branch (the compiler would wait for
// every load in flight after one).  Stores: whole chunks one vector, a
// partial last chunk pixel by pixel.
#pragma once

namespace clk {

// A chunk loaded from pixel s = min(8j, W - 8) of its row (W >= 8: inside
// the row) moved into place: a row's partial last chunk (d = 8j - s in
// 1..7) shifts its pixels down by d and fills the top with the row's last
// pixel, in registers (shifts, no indexed registers: a select among words
// by index is turned into a private array in scratch memory).
template <class TS>
__device__ __forceinline__ void xt_chunk_fix(Chunk<TS>& c, const int d) {
  typedef unsigned __int128 u128;
  if constexpr (sizeof(TS) == 1) {
    const uint64_t x = (uint64_t)c.w[0] | ((uint64_t)c.w[1] << 32);
    const uint64_t f = (uint64_t)(c.w[1] >> 24) * 0x0101010101010101ull;
    const int sh = 8 * d;  // 8 .. 56
    const uint64_t y = (x >> sh) | (f << (64 - sh));
    c.w[0] = (uint32_t)y;
    c.w[1] = (uint32_t)(y >> 32);
  } else if constexpr (sizeof(TS) == 2) {
    const uint32_t f = (c.w[3] >> 16) * 0x10001u;
    const u128 x = (u128)c.w[0] | ((u128)c.w[1] << 32) | ((u128)c.w[2] << 64) | ((u128)c.w[3] << 96);
    const u128 fill = (u128)f * ((((u128)1) << 96) | (((u128)1) << 64) | (((u128)1) << 32) | 1u);
    const int sh = 16 * d;  // 16 .. 112
    const u128 y = (x >> sh) | (fill << (128 - sh));
#pragma unroll
    for (int k = 0; k < 4; ++k) c.w[k] = (uint32_t)(y >> (32 * k));
  } else {
    const uint32_t f = c.w[7];
    const u128 lo = (u128)c.w[0] | ((u128)c.w[1] << 32) | ((u128)c.w[2] << 64) | ((u128)c.w[3] << 96);
    const u128 hi = (u128)c.w[4] | ((u128)c.w[5] << 32) | ((u128)c.w[6] << 64) | ((u128)c.w[7] << 96);
    const u128 fill = (u128)f * ((((u128)1) << 96) | (((u128)1) << 64) | (((u128)1) << 32) | 1u);
    const int sh = 32 * d;  // 32 .. 224
    u128 nlo, nhi;
    if (sh < 128) {
      nlo = (lo >> sh) | (hi << (128 - sh));
      nhi = (hi >> sh) | (fill << (128 - sh));
    } else if (sh == 128) {
      nlo = hi;
      nhi = fill;
    } else {
      nlo = (hi >> (sh - 128)) | (fill << (256 - sh));
      nhi = fill;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      c.w[k] = (uint32_t)(nlo >> (32 * k));
      c.w[4 + k] = (uint32_t)(nhi >> (32 * k));
    }
  }
}

// The chunks a wave owns: its lanes NB .. 63 - NB; the NB lanes on either
// side hold the chunks next to them (the DPP shifts' sources) and store
// nothing, so no lane loads more than its own chunk.
template <int CMAX>
constexpr int xt_owned() { return 64 - 2 * ((CMAX + 7) / 8); }

template <class TS>
struct XtUnit {
  Chunk<TS> own;
  uint32_t row, j;
  int fix;  // pixels to move the loaded chunk down (a partial last chunk)
  bool active, live;
};

// p.L = W (>= 8), p.outer = rows, p.nseg = chunks per row (with its division
// constants), p.ntasks = rows * chunks per row (< 2^31).  Every lane loads
// its chunk as one vector from min(8j, W - 8) (inside the row, aligned or
// not); the fix-up waits until the chunk is used, so the load stays in
// flight.
template <class TS, int NB>
__device__ __forceinline__ void xt_load(const SepParams& p, const uint32_t unit, const int lane, XtUnit<TS>& u) {
  constexpr int OWN = 64 - 2 * NB;
  const uint32_t per_row = (uint32_t)p.nseg;
  const int64_t total = p.ntasks;
  u.live = (int64_t)unit * OWN < total;  // wave-uniform
  const int64_t g = (int64_t)unit * OWN + lane - NB;
  u.active = u.live && lane >= NB && lane < 64 - NB && g < total;
  const uint32_t gc = (uint32_t)(g < 0 ? 0 : (g < total ? g : total - 1));
  u.row = div_nseg(p, gc);
  u.j = gc - u.row * per_row;
  const int W = (int)p.L;
  const int x0 = 8 * (int)u.j;
  const int s = x0 + 8 <= W ? x0 : W - 8;
  u.fix = x0 - s;
  const TS* __restrict__ rp = static_cast<const TS*>(p.src) + (uint64_t)u.row * (uint64_t)W;
  load_nt_words_ua<8 * (int)sizeof(TS)>(rp + s, u.own.w);
}

template <int OP, class TS, class TD, int CMAX>
__device__ __forceinline__ void xt_compute(const SepParams& p, const XtUnit<TS>& u) {
  using VT = xr_vt<OP, TS>;
  constexpr bool kFloatSrc = !Pix<TS>::kInt;
  constexpr int NB = (CMAX + 7) / 8;
  constexpr int WIN = 8 + 2 * CMAX;
  const uint32_t per_row = (uint32_t)p.nseg;
  const uint32_t j = u.j;
  const int c = __builtin_amdgcn_readfirstlane(p.c);
  const int tlo = CMAX - c, thi = CMAX + c;
  Chunk<TS> own = u.own;
  if (u.fix) xt_chunk_fix<TS>(own, u.fix);  // a row's partial last chunk
  Chunk<TS> prev[NB], next[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k)
#pragma unroll
    for (int w = 0; w < Chunk<TS>::NW; ++w) {
      const uint32_t v = k == 0 ? own.w[w] : prev[k - 1].w[w];
      const uint32_t x = k == 0 ? own.w[w] : next[k - 1].w[w];
      prev[k].w[w] = dpp_from_left(v, v);  // halo lanes at the wave's ends hold no output
      next[k].w[w] = dpp_from_right(x, x);
    }
  // clamp to edge: chunks before the row start hold its first pixel, chunks
  // past its end its last pixel (pixel 7 of the row's last chunk, which the
  // loader clamped)
  if (j < (uint32_t)NB) {
    const TS first = word_px<TS>((j == 0) ? own.w[0] : prev[0].w[0], 0);
#pragma unroll
    for (int k = 1; k <= NB; ++k)
      if (j < (uint32_t)k) fill_chunk<TS>(first, prev[k - 1]);
  }
  if (j + NB >= per_row) {
    constexpr int NW = Chunk<TS>::NW;
    const TS last = word_px<TS>((j + 1 == per_row) ? own.w[NW - 1] : next[0].w[NW - 1], 7);
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
    if (ch == 0) t = own.px(ei);
    else if (ch < 0) t = prev[-ch - 1].px(ei);
    else t = next[ch - 1].px(ei);
    win[i] = static_cast<VT>(t);
  }
  // the tier's weights read before the tap loop (the table entry holds
  // them all), so their scalar loads merge and are waited for once, not one
  // load and wait inside each tap's branch
  float wt[2 * CMAX + 1];
#pragma unroll
  for (int t = 0; t <= 2 * CMAX; ++t) wt[t] = (OP == OP_GAUSS) ? p.w[t] : 0.f;
  VT acc[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    if constexpr (OP == OP_GAUSS || OP == OP_MEAN) acc[q] = 0;
    else acc[q] = win[q + CMAX];  // centre first
  }
#pragma unroll
  for (int t = 0; t <= 2 * CMAX; ++t) {
    if (t < tlo || t > thi) continue;  // wave-uniform
    const float h = wt[t];
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
  // every unit's load is issued first (clamped indices past the end), so
  // none sits in a branch and unit k + 1's load overlaps unit k's filter
  XtUnit<TS> u[R];
#pragma unroll
  for (int k = 0; k < R; ++k) xt_load<TS, NB>(p, unit0 + k, lane, u[k]);
#pragma unroll
  for (int k = 0; k < R; ++k) {
    if (!u[k].live) return;  // whole waves only
    xt_compute<OP, TS, TD, CMAX>(p, u[k]);
  }
}

}  // namespace clk
