// y/z pass for small half widths (C <= 8, Gaussians up to 12): the window
// width is a template parameter.  Included by sep_impl.h.
//
// Each thread walks a segment of the filtered axis for COLS adjacent
// columns, K = 8 outputs at a time.  It holds the K+N-1 inputs of the
// current block in registers and, before computing that block, issues the K
// loads of the next one, so a wave keeps K vector loads in flight while it
// computes; along the segment every input is loaded once.  Outputs are
// computed one at a time from static registers (the reference's tap order),
// so no accumulator array is live.  Addresses are a wave-uniform base (the
// block's first slab plus the axis position times the stride, computed on
// the scalar unit with 32-bit clamps) plus one 32-bit per-lane offset.
#pragma once

namespace clk {

template <int OP, class TS, class TD, int C, int COLS, int K = 8>
__global__ __launch_bounds__(256) void sep_ystatic_kernel(const SepParams p) {
  using VT = xr_vt<OP, TS>;
  constexpr bool kFloatSrc = !Pix<TS>::kInt;
  constexpr int N = 2 * C + 1;
  constexpr int WN = K + N - 1;
  uint32_t bid = blockIdx.x;
  if (p.order & 2) {
    // the hardware deals blocks to the 8 XCDs round robin; give each XCD one
    // contiguous range of logical blocks (a bijection for any grid size)
    const uint32_t G = gridDim.x, x = bid & 7u, q = G >> 3, r = G & 7u;
    bid = x * q + (x < r ? x : r) + (bid >> 3);
  }
  uint32_t sid, b;
  if (p.order & 1) {
    b = div_nseg(p, bid);
    sid = bid - b * (uint32_t)p.nseg;
  } else {
    sid = bid / p.nblk;
    b = bid - sid * p.nblk;
  }
  const int64_t g0 = (int64_t)b * 256;
  const int64_t g = g0 + threadIdx.x;
  if (g >= p.ncols_total) return;
  const int64_t o0 = g0 / p.ncolgroups;
  const int64_t o = g / p.ncolgroups;
  const int64_t cg = g - o * p.ncolgroups;
  const int64_t slab = p.L * p.inner;
  const uint32_t lso = (uint32_t)(((o - o0) * slab + cg * COLS) * (int64_t)sizeof(TS));
  const uint32_t ldo = (uint32_t)(((o - o0) * slab + cg * COLS) * (int64_t)sizeof(TD));
  const char* const ls0 = static_cast<const char*>(p.src) + o0 * slab * (int64_t)sizeof(TS);
  char* const ld0 = static_cast<char*>(p.dst) + o0 * slab * (int64_t)sizeof(TD);
  const int64_t ss = p.inner * (int64_t)sizeof(TS);
  const int64_t ds = p.inner * (int64_t)sizeof(TD);
  const int L = (int)p.L;
  const int a0 = (int)sid * p.seg;
  const int a1 = (a0 + p.seg < L) ? a0 + p.seg : L;

  auto ldp = [&](int pos, VT (&dst)[COLS]) {
    pos = pos < 0 ? 0 : (pos > L - 1 ? L - 1 : pos);
    load_vt_stream<TS, VT, COLS>(reinterpret_cast<const TS*>(ls0 + (int64_t)pos * ss + lso), dst);
  };

  VT win[WN][COLS];
#pragma unroll
  for (int i = 0; i < WN; ++i) ldp(a0 - C + i, win[i]);
  // Wait for the window here, once.  Otherwise the compiler's wait
  // placement merges this entry path into the loop and waits for every
  // load (the prefetch included) before each block's arithmetic, which
  // serialises loads and compute.
  __builtin_amdgcn_s_waitcnt(0);
  // one output row of the block: static q, so the window indexes stay static
  auto emit = [&](const int q, const int p0) {
    TD out[COLS];
#pragma unroll
    for (int j = 0; j < COLS; ++j) {
      VT acc;
      if constexpr (OP == OP_GAUSS) {
        acc = 0.f;
#pragma unroll
        for (int t = 0; t < N; ++t) acc = __builtin_fmaf(p.w[t], win[q + t][j], acc);
      } else if constexpr (OP == OP_MEAN) {
        acc = 0;
#pragma unroll
        for (int t = 0; t < N; ++t) acc = acc + win[q + t][j];
      } else {
        acc = win[q + C][j];  // centre first
#pragma unroll
        for (int t = 0; t < N; ++t) acc = tap_apply<OP, kFloatSrc, VT>(acc, win[q + t][j], 0.f);
      }
      if constexpr (OP == OP_GAUSS || OP == OP_MEAN) {
        const float a = static_cast<float>(acc);  // exact for integer sums
        out[j] = convert_sat<TD>(div_planned<OP, TS>(a, p));
      } else if constexpr (std::is_same<VT, int32_t>::value) {
        out[j] = clamp_int<TD>(acc);
      } else {
        out[j] = convert_sat<TD>(acc);
      }
    }
    store_nt<TD, COLS>(reinterpret_cast<TD*>(ld0 + (int64_t)(p0 + q) * ds + ldo), out);
  };

  // Blocks with a successor: prefetch it, emit all K rows (no per-row test),
  // shift.  The loop has one path, so the compiler's wait before the shift
  // covers only the prefetched rows, not the stores issued after them.
  int p0 = a0;
  for (; p0 + K < a1; p0 += K) {
    VT nxt[K][COLS];
#pragma unroll
    for (int i = 0; i < K; ++i) ldp(p0 + K + C + i, nxt[i]);
#pragma unroll
    for (int q = 0; q < K; ++q) emit(q, p0);
#pragma unroll
    for (int i = 0; i < N - 1; ++i)
#pragma unroll
      for (int j = 0; j < COLS; ++j) win[i][j] = win[i + K][j];
#pragma unroll
    for (int i = 0; i < K; ++i)
#pragma unroll
      for (int j = 0; j < COLS; ++j) win[N - 1 + i][j] = nxt[i][j];
  }
  // the segment's last block (a0 <= p0 < a1 <= p0 + K)
  if (p0 + K <= a1) {
#pragma unroll
    for (int q = 0; q < K; ++q) emit(q, p0);
  } else {
#pragma unroll
    for (int q = 0; q < K; ++q)
      if (p0 + q < a1) emit(q, p0);
  }
}

}  // namespace clk
