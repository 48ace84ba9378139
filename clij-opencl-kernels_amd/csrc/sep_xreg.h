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

// Pixel i (little-endian order) of a 32-bit word holding 4 / sizeof(TS)
// pixels, by shifts: no byte addressing, so the words stay in registers.
template <class TS>
__device__ __forceinline__ TS word_px(uint32_t word, int i) {
  if constexpr (sizeof(TS) == 4) {
    return __builtin_bit_cast(TS, word);
  } else {
    constexpr int per = 4 / (int)sizeof(TS);
    return static_cast<TS>(word >> (8 * (int)sizeof(TS) * (i % per)));
  }
}

// Chunk of 8 pixels as 32-bit words.
template <class TS> struct Chunk {
  static constexpr int NW = 8 * (int)sizeof(TS) / 4;
  static constexpr int PER = 4 / (int)sizeof(TS);
  uint32_t w[NW];
  __device__ __forceinline__ TS px(int i) const { return word_px<TS>(w[i / PER], i); }
};

// Value type of the register window: int32 for integer pixels under
// min/max/mean (sums of at most 33 16-bit values are exact), else float.
template <int OP, class TS>
using xr_vt = std::conditional_t<(OP != OP_GAUSS && Pix<TS>::kInt), int32_t, float>;

// Every pixel of the chunk set to v (words built by shifts, as word_px).
template <class TS>
__device__ __forceinline__ void fill_chunk(TS v, Chunk<TS>& c) {
  uint32_t word;
  if constexpr (sizeof(TS) == 4) {
    word = __builtin_bit_cast(uint32_t, v);
  } else if constexpr (sizeof(TS) == 2) {
    const uint32_t b = (uint32_t)(uint16_t)v;
    word = b | (b << 16);
  } else {
    word = (uint32_t)(uint8_t)v * 0x01010101u;
  }
#pragma unroll
  for (int i = 0; i < Chunk<TS>::NW; ++i) c.w[i] = word;
}

// Rows are packed: the wave's 64 lanes take 64 consecutive 8-pixel chunks
// of the row-major volume, whatever row each falls in, so rows shorter than
// 512 pixels keep every lane busy.  A lane whose chunk starts a row uses the
// row's first pixel repeated as its left neighbour (clamp to edge) instead
// of the lane to its left, and likewise at the row end.
//
// A wave filters R consecutive units of 64 chunks; the loads of unit u+1
// are issued before unit u is computed, so they are in flight meanwhile.

// One unit's inputs: the lane's chunk and, for the wave's two end lanes,
// the chunks beyond the wave (left of lane 0, right of lane 63).
template <class TS, int NB>
struct XUnit {
  Chunk<TS> own, edge[NB];
  uint64_t row, j;
  bool active, live;
};

template <class TS, int NB>
__device__ __forceinline__ void xreg_load(const SepParams& p, const uint64_t unit, const int lane,
                                          XUnit<TS, NB>& u) {
  const uint64_t per_row = (uint64_t)p.L / 8;
  const uint64_t total = (uint64_t)p.outer * per_row;
  u.live = unit * 64 < total;  // wave-uniform
  if (!u.live) return;
  const uint64_t g = unit * 64 + lane;
  u.active = g < total;
  const uint64_t gc = u.active ? g : total - 1;  // idle tail lanes mirror the last chunk
  if (total < (uint64_t(1) << 32)) {
    u.row = (uint32_t)gc / (uint32_t)per_row;
    u.j = (uint32_t)gc - (uint32_t)u.row * (uint32_t)per_row;
  } else {
    u.row = gc / per_row;
    u.j = gc - u.row * per_row;
  }
  const TS* __restrict__ rp = static_cast<const TS*>(p.src) + u.row * (uint64_t)p.L;
  const int64_t xl = 8 * (int64_t)u.j;
  auto load8 = [&](int64_t x, Chunk<TS>& c) { load_nt_words<8 * (int)sizeof(TS)>(rp + x, c.w); };
  load8(xl, u.own);
#pragma unroll
  for (int k = 0; k < NB; ++k)
#pragma unroll
    for (int w = 0; w < Chunk<TS>::NW; ++w) u.edge[k].w[w] = 0u;
  // One branch for both end lanes, and no per-chunk branch: a chunk outside
  // the row is replaced by the clamp fill later, so its load just re-reads
  // the lane's own chunk.
  if (lane == 0 || lane == 63) {
#pragma unroll
    for (int k = 1; k <= NB; ++k) {
      const int64_t x = (lane == 0) ? (u.j >= (uint64_t)k ? xl - 8 * k : xl)
                                    : (u.j + k < per_row ? xl + 8 * k : xl);
      load8(x, u.edge[k - 1]);
    }
  }
}

template <int OP, class TS, class TD, int C>
__device__ __forceinline__ void xreg_compute(const SepParams& p, const XUnit<TS, (C + 7) / 8>& u) {
  using VT = xr_vt<OP, TS>;
  constexpr bool kFloatSrc = !Pix<TS>::kInt;
  constexpr int NB = (C + 7) / 8;
  constexpr int WIN = 8 + 2 * C;
  const uint64_t per_row = (uint64_t)p.L / 8;
  const uint64_t j = u.j;
  // chunks k = 1..NB away: from lane l-k / l+k by repeated wave shifts;
  // lanes 0 and 63 loaded theirs, which enter as the shifts' old values
  // (a shift keeps the old value only in the lane it has no source for)
  Chunk<TS> prev[NB], next[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k)
#pragma unroll
    for (int w = 0; w < Chunk<TS>::NW; ++w) {
      prev[k].w[w] = dpp_from_left(u.edge[k].w[w], k == 0 ? u.own.w[w] : prev[k - 1].w[w]);
      next[k].w[w] = dpp_from_right(u.edge[k].w[w], k == 0 ? u.own.w[w] : next[k - 1].w[w]);
    }
  // clamp to edge: chunks before the row start are its first pixel, chunks
  // past its end its last pixel.  The row's first chunk is the lane's own
  // (j = 0) or the one to its left (j = 1), and likewise at the end, so no
  // extra load is needed.
  if (j < (uint64_t)NB) {
    const TS first = word_px<TS>((j == 0) ? u.own.w[0] : prev[0].w[0], 0);
#pragma unroll
    for (int k = 1; k <= NB; ++k)
      if (j < (uint64_t)k) fill_chunk<TS>(first, prev[k - 1]);
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
    const int e = i - C;  // offset from the lane's chunk start
    const int ch = (e >= 0) ? e / 8 : -((-e + 7) / 8);
    const int ei = e - 8 * ch;
    TS t;
    if (ch == 0) t = u.own.px(ei);
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
  if (u.active) store_nt<TD, 8>(static_cast<TD*>(p.dst) + u.row * (uint64_t)p.L + 8 * j, out);
}

template <int OP, class TS, class TD, int C, int R>
__device__ __forceinline__ void sep_xreg_body(const SepParams& p, const uint32_t vblk) {
  static_assert(C >= 1 && C <= 16, "at most two neighbour chunks per side");
  constexpr int NB = (C + 7) / 8;  // neighbour chunks per side
  const int lane = threadIdx.x & 63;
  const uint64_t unit0 = ((uint64_t)vblk * 4 + (threadIdx.x >> 6)) * R;
  XUnit<TS, NB> u[R];
  xreg_load<TS, NB>(p, unit0, lane, u[0]);
#pragma unroll
  for (int k = 0; k < R; ++k) {
    if (k + 1 < R && u[k].live) xreg_load<TS, NB>(p, unit0 + k + 1, lane, u[k + 1]);
    if (!u[k].live) return;  // whole waves only
    xreg_compute<OP, TS, TD, C>(p, u[k]);
  }
}

template <int OP, class TS, class TD, int C, int R>
__global__ __launch_bounds__(256) void sep_xreg_kernel(const SepParams p) {
  sep_xreg_body<OP, TS, TD, C, R>(p, blockIdx.x);
}

// blocks of 4 waves, R units of 64 chunks per wave, for a packed x pass
inline int64_t xreg_blocks(const SepParams& p, int R) {
  const int64_t chunks = p.outer * (p.L / 8);
  const int64_t units = (chunks + 63) / 64;
  const int64_t waves = (units + R - 1) / R;
  return (waves + 3) / 4;
}

}  // namespace clk
