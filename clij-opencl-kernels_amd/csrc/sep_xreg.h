// x pass for small half widths: the whole window of a lane lives in
// registers.  Included by sep_impl.h.
//
// Lane l of a wave owns 8 consecutive outputs (one 8-pixel chunk) and loads
// exactly that chunk with one vector load.  The c pixels it needs on each side are the neighbouring lanes'
// chunks, moved across with DPP wave shifts (wave_shr:1 gives lane l the
// value of lane l-1, wave_shl:1 that of lane l+1); only lanes 0 and 63
// load halo chunks, which enter through the shift's "old" operand.  With
// the half width C a template parameter every tap is a static register, so
// there is no LDS, no barrier and no per-element address arithmetic.
// The fast path requires W % 8 == 0 and aligned rows, so chunks never
// straddle a row end.
#pragma once

namespace clk {

__device__ __forceinline__ uint32_t dpp_from_left(uint32_t old, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, 0x138, 0xf, 0xf, false);  // wave_shr:1
}
__device__ __forceinline__ uint32_t dpp_from_right(uint32_t old, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, 0x130, 0xf, 0xf, false);  // wave_shl:1
}

// Chunk of 8 pixels as 32-bit words.
template <class TS> struct Chunk {
  static constexpr int NW = 8 * (int)sizeof(TS) / 4;
  uint32_t w[NW];
  __device__ __forceinline__ TS px(int i) const {
    TS t;
    __builtin_memcpy(&t, reinterpret_cast<const char*>(w) + i * sizeof(TS), sizeof(TS));
    return t;
  }
};

// Load the chunk starting at x (a multiple of 8) of a row of length W.
template <class TS>
__device__ __forceinline__ void load_chunk(const TS* __restrict__ row, int64_t x, int64_t W, Chunk<TS>& c) {
  if (x >= 0 && x < W) {
    const PixVec<TS, 8> v = load_nt<TS, 8>(row + x);
    __builtin_memcpy(c.w, &v, sizeof(v));
  } else {
    const TS e = row[x < 0 ? 0 : W - 1];
    TS t[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) t[i] = e;
    __builtin_memcpy(c.w, t, sizeof(t));
  }
}

// Value type of the register window: int32 for integer pixels under
// min/max/mean (sums of at most 33 16-bit values are exact), else float.
template <int OP, class TS>
using xr_vt = std::conditional_t<(OP != OP_GAUSS && Pix<TS>::kInt), int32_t, float>;

template <class TS>
__device__ __forceinline__ void fill_chunk(TS v, Chunk<TS>& c) {
  TS t[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) t[i] = v;
  __builtin_memcpy(c.w, t, sizeof(t));
}

// Rows are packed: the wave's 64 lanes take 64 consecutive 8-pixel chunks
// of the row-major volume, whatever row each falls in, so rows shorter than
// 512 pixels keep every lane busy.  A lane whose chunk starts a row uses the
// row's first pixel repeated as its left neighbour (clamp to edge) instead
// of the lane to its left, and likewise at the row end.
template <int OP, class TS, class TD, int C>
__device__ __forceinline__ void sep_xreg_body(const SepParams& p, const uint32_t vblk) {
  static_assert(C >= 1 && C <= 16, "at most two neighbour chunks per side");
  using VT = xr_vt<OP, TS>;
  constexpr bool kFloatSrc = !Pix<TS>::kInt;
  constexpr int NB = (C + 7) / 8;  // neighbour chunks per side
  constexpr int WIN = 8 + 2 * C;
  const int lane = threadIdx.x & 63;
  const uint64_t wave = (uint64_t)vblk * 4 + (threadIdx.x >> 6);
  const uint64_t per_row = (uint64_t)p.L / 8;
  const uint64_t total = (uint64_t)p.outer * per_row;
  if (wave * 64 >= total) return;  // whole waves only
  const uint64_t g = wave * 64 + lane;
  const bool active = g < total;
  const uint64_t gc = active ? g : total - 1;  // idle tail lanes mirror the last chunk
  uint64_t row, j;
  if (total < (uint64_t(1) << 32)) {
    row = (uint32_t)gc / (uint32_t)per_row;
    j = (uint32_t)gc - (uint32_t)row * (uint32_t)per_row;
  } else {
    row = gc / per_row;
    j = gc - row * per_row;
  }
  const TS* __restrict__ rp = static_cast<const TS*>(p.src) + row * (uint64_t)p.L;
  const int64_t xl = 8 * (int64_t)j;
  auto load8 = [&](int64_t x, Chunk<TS>& c) {
    const PixVec<TS, 8> v = load_nt<TS, 8>(rp + x);
    __builtin_memcpy(c.w, &v, sizeof(v));
  };

  Chunk<TS> own;
  load8(xl, own);
  // chunks k = 1..NB away: from lane l-k / l+k by repeated wave shifts;
  // lanes 0 and 63 fetch theirs (they enter as the shifts' old values)
  Chunk<TS> lm[NB], rm[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k)
#pragma unroll
    for (int w = 0; w < Chunk<TS>::NW; ++w) lm[k].w[w] = rm[k].w[w] = 0u;
  if (lane == 0) {
#pragma unroll
    for (int k = 1; k <= NB; ++k)
      if (j >= (uint64_t)k) load8(xl - 8 * k, lm[k - 1]);
  }
  if (lane == 63) {
#pragma unroll
    for (int k = 1; k <= NB; ++k)
      if (j + k < per_row) load8(xl + 8 * k, rm[k - 1]);
  }
  Chunk<TS> prev[NB], next[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k)
#pragma unroll
    for (int w = 0; w < Chunk<TS>::NW; ++w) {
      prev[k].w[w] = dpp_from_left(lm[k].w[w], k == 0 ? own.w[w] : prev[k - 1].w[w]);
      next[k].w[w] = dpp_from_right(rm[k].w[w], k == 0 ? own.w[w] : next[k - 1].w[w]);
    }
  // clamp to edge: chunks before the row start are its first pixel, chunks
  // past its end the last pixel
  if (j < (uint64_t)NB) {
    const TS first = rp[0];
#pragma unroll
    for (int k = 1; k <= NB; ++k)
      if (j < (uint64_t)k) fill_chunk<TS>(first, prev[k - 1]);
  }
  if (j + NB >= per_row) {
    const TS last = rp[p.L - 1];
#pragma unroll
    for (int k = 1; k <= NB; ++k)
      if (j + k >= per_row) fill_chunk<TS>(last, next[k - 1]);
  }

  VT win[WIN];
#pragma unroll
  for (int i = 0; i < WIN; ++i) {
    const int e = i - C;  // offset from the lane's chunk start
    const int ch = (e >= 0) ? e / 8 : -((-e + 7) / 8);
    const int ei = e - 8 * ch;
    TS t;
    if (ch == 0) t = own.px(ei);
    else if (ch < 0) t = prev[-ch - 1].px(ei);
    else t = next[ch - 1].px(ei);
    win[i] = static_cast<VT>(t);
  }
  TD out[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    VT acc;
    if constexpr (OP == OP_GAUSS) {
      acc = 0.f;
#pragma unroll
      for (int t = 0; t <= 2 * C; ++t) acc = __builtin_fmaf(p.w[t], win[q + t], acc);
    } else if constexpr (OP == OP_MEAN) {
      acc = 0;
#pragma unroll
      for (int t = 0; t <= 2 * C; ++t) acc = acc + win[q + t];
    } else {
      acc = win[q + C];  // centre first
#pragma unroll
      for (int t = 0; t <= 2 * C; ++t) acc = tap_apply<OP, kFloatSrc, VT>(acc, win[q + t], 0.f);
    }
    if constexpr (OP == OP_GAUSS || OP == OP_MEAN) {
      const float a = static_cast<float>(acc);  // exact for integer sums
      out[q] = convert_sat<TD>(div_planned<OP, TS>(a, p));
    } else if constexpr (std::is_same<VT, int32_t>::value) {
      out[q] = clamp_int<TD>(acc);
    } else {
      out[q] = convert_sat<TD>(acc);
    }
  }
  if (active) store_nt<TD, 8>(static_cast<TD*>(p.dst) + row * (uint64_t)p.L + xl, out);
}

template <int OP, class TS, class TD, int C>
__global__ __launch_bounds__(256) void sep_xreg_kernel(const SepParams p) {
  sep_xreg_body<OP, TS, TD, C>(p, blockIdx.x);
}

// blocks of 4 waves for a packed x pass
inline int64_t xreg_blocks(const SepParams& p) {
  const int64_t chunks = p.outer * (p.L / 8);
  const int64_t waves = (chunks + 63) / 64;
  return (waves + 3) / 4;
}

}  // namespace clk
