// x pass for small half widths: the whole window of a lane lives in
// registers.  Included by sep_impl.h.
//
// A wave filters 512 consecutive outputs of one row; lane l owns outputs
// x0+8l .. x0+8l+7 and loads exactly that 8-pixel chunk with one vector
// load.  The c pixels it needs on each side are the neighbouring lanes'
// chunks, moved across with DPP wave shifts (wave_shr:1 gives lane l the
// value of lane l-1, wave_shl:1 that of lane l+1); only lanes 0 and 63
// load halo chunks, which enter through the shift's "old" operand.  With
// the half width C a template parameter every tap is a static register, so
// there is no LDS, no barrier and no per-element address arithmetic.
// Clamp-to-edge: a chunk that starts at or beyond the row end is the last
// pixel repeated, one that starts before 0 the first pixel; the fast path
// requires W % 8 == 0 and aligned rows, so chunks never straddle the end.
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

template <int OP, class TS, class TD, int C>
__global__ __launch_bounds__(256) void sep_xreg_kernel(const SepParams p) {
  using VT = xr_vt<OP, TS>;
  constexpr bool kFloatSrc = !Pix<TS>::kInt;
  constexpr int NB = (C + 7) / 8;   // neighbour chunks needed on each side
  constexpr int WIN = 8 + 2 * C;
  const int lane = threadIdx.x & 63;
  const uint32_t task = blockIdx.x * 4u + (threadIdx.x >> 6);
  if (task >= (uint32_t)p.ntasks) return;  // whole waves only (4 tasks per block)
  const uint32_t row = task / (uint32_t)p.nseg;
  const int64_t x0 = (int64_t)(task - row * (uint32_t)p.nseg) * kXSeg;
  const int64_t W = p.L;
  const TS* __restrict__ rp = static_cast<const TS*>(p.src) + (int64_t)row * W;
  const int64_t xl = x0 + 8 * lane;

  Chunk<TS> own;
  load_chunk<TS>(rp, xl, W, own);
  // halo chunks, loaded by the two edge lanes only: lane 0 holds chunk -k,
  // lane 63 chunk 64+k-1, in halo[k-1]
  Chunk<TS> halo[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k)
#pragma unroll
    for (int j = 0; j < Chunk<TS>::NW; ++j) halo[k].w[j] = 0u;
  if (lane == 0 || lane == 63) {
#pragma unroll
    for (int k = 1; k <= NB; ++k)
      load_chunk<TS>(rp, lane == 0 ? x0 - 8 * k : x0 + kXSeg + 8 * (k - 1), W, halo[k - 1]);
  }
  Chunk<TS> prev[NB], next[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) {
#pragma unroll
    for (int j = 0; j < Chunk<TS>::NW; ++j) {
      prev[k].w[j] = dpp_from_left(halo[k].w[j], k == 0 ? own.w[j] : prev[k - 1].w[j]);
      next[k].w[j] = dpp_from_right(halo[k].w[j], k == 0 ? own.w[j] : next[k - 1].w[j]);
    }
  }
  // window element i is the pixel at xl - C + i
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
      const float r = p.fastdiv ? div_uniform_fast(a, p.div, p.rdiv) : div_uniform(a, p.div, p.rdiv);
      out[q] = convert_sat<TD>(r);
    } else if constexpr (std::is_same<VT, int32_t>::value) {
      out[q] = clamp_int<TD>(acc);
    } else {
      out[q] = convert_sat<TD>(acc);
    }
  }
  if (xl < W) store_nt<TD, 8>(static_cast<TD*>(p.dst) + (int64_t)row * W + xl, out);
}

}  // namespace clk
