// y/z pass body for a half width known only at run time, up to a
// compile-time tier CMAX (the batch engine, whose jobs each bring their own
// c).  Included by sep_impl.h.
//
// The register ring of the static passes, with every slot a compile-time
// register: the ring holds input rows from a0 - CMAX on (slot = row % RS),
// output q of a step reads rows q + t' for t' = 0..2*CMAX, and the taps with
// t' outside [CMAX - c, CMAX + c] are skipped by wave-uniform branches (c is
// the job's, uniform over the block).  The host stores the Gaussian weight
// of tap v = t' - CMAX at w[t'] (set_tier_weights), so the weight index is
// static too.  The taps run in the reference's order (v = -c .. c), so the
// results are those of the single launches bit for bit.  A step loads only
// the K rows of the next step (from CMAX + K candidate slots, chosen by
// uniform branches); rows a job's c does not need are never loaded.  The
// rows in use span at most 2K + 2c < RS, so no slot is overwritten early.
#pragma once

namespace clk {

// Gaussian weights w[0..2c] moved to w[CMAX - c .. CMAX + c].
inline void set_tier_weights(SepParams& p, int cmax) {
  const int c = p.c;
  if (c == cmax) return;
  for (int t = 2 * c; t >= 0; --t) p.w[t + cmax - c] = p.w[t];
}

template <int OP, class TS, class TD, int CMAX, int COLS, int K>
__device__ __forceinline__ void sep_ydyn_body(const SepParams& p, const uint32_t vblk) {
  using VT = xr_vt<OP, TS>;
  constexpr bool kFloatSrc = !Pix<TS>::kInt;
  constexpr int NT = 2 * CMAX + 1;
  constexpr int U = (2 * K + 2 * CMAX + K - 1) / K;
  constexpr int RS = U * K;
  const uint32_t sid = vblk / p.nblk;
  const uint32_t b = vblk - sid * p.nblk;
  const int64_t g = (int64_t)b * 256 + threadIdx.x;
  if (g >= p.ncols_total) return;  // no barriers below
  const int64_t o = g / p.ncolgroups;
  const int64_t cg = g - o * p.ncolgroups;
  const int64_t eoff = o * (p.L * p.inner) + cg * COLS;
  const char* const ls = static_cast<const char*>(p.src) + eoff * (int64_t)sizeof(TS);
  char* const ld = static_cast<char*>(p.dst) + eoff * (int64_t)sizeof(TD);
  const int64_t ss = p.inner * (int64_t)sizeof(TS);
  const int64_t ds = p.inner * (int64_t)sizeof(TD);
  const int L = (int)p.L;
  const int c = __builtin_amdgcn_readfirstlane(p.c);
  const int tlo = CMAX - c, thi = CMAX + c;  // taps t' in [tlo, thi] are the job's
  const int a0 = (int)sid * p.seg;
  const int a1 = (a0 + p.seg < L) ? a0 + p.seg : L;

  VT ring[RS][COLS];
  auto ldrow = [&](int r, VT (&dst)[COLS]) {
    int pos = a0 - CMAX + r;
    pos = pos < 0 ? 0 : (pos > L - 1 ? L - 1 : pos);
    load_vt<TS, VT, COLS>(reinterpret_cast<const TS*>(ls + (int64_t)pos * ss), dst);
  };
#pragma unroll
  for (int i = 0; i < K + 2 * CMAX; ++i)
    if (i >= tlo && i < K + thi) ldrow(i, ring[i]);

  TD outs[K][COLS];
  // the K outputs of a step whose window starts at slot u*K (static),
  // tap-major: output q still sees its taps in order t' = tlo .. thi
  auto compute = [&](const int u) {
    VT acc[K][COLS];
#pragma unroll
    for (int q = 0; q < K; ++q)
#pragma unroll
      for (int j = 0; j < COLS; ++j) {
        if constexpr (OP == OP_GAUSS || OP == OP_MEAN) acc[q][j] = 0;
        else acc[q][j] = ring[(u * K + q + CMAX) % RS][j];  // centre first
      }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (t < tlo || t > thi) continue;  // wave-uniform
      const float h = (OP == OP_GAUSS) ? p.w[t] : 0.f;
#pragma unroll
      for (int q = 0; q < K; ++q)
#pragma unroll
        for (int j = 0; j < COLS; ++j) {
          const VT x = ring[(u * K + q + t) % RS][j];
          if constexpr (OP == OP_GAUSS) acc[q][j] = __builtin_fmaf(h, x, acc[q][j]);
          else if constexpr (OP == OP_MEAN) acc[q][j] = acc[q][j] + x;
          else acc[q][j] = tap_apply<OP, kFloatSrc, VT>(acc[q][j], x, 0.f);
        }
    }
#pragma unroll
    for (int q = 0; q < K; ++q)
#pragma unroll
      for (int j = 0; j < COLS; ++j) outs[q][j] = finish<OP, TS, TD, VT>(acc[q][j], p);
  };

  for (int s = 0;; s += U) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int p0 = a0 + (s + u) * K;
      if (p0 >= a1) return;
      if (p0 + K < a1) {
        // the next step's new rows are (s+u)*K + K + j for j in
        // [thi, thi + K - 1]; j runs over the static superset
        // [CMAX, 2*CMAX + K - 1] so that every slot is static
#pragma unroll
        for (int j = CMAX; j < 2 * CMAX + K; ++j)
          if (j >= thi && j < thi + K) ldrow((s + u) * K + K + j, ring[(u * K + K + j) % RS]);
        compute(u);
#pragma unroll
        for (int q = 0; q < K; ++q)
          store_nt<TD, COLS>(reinterpret_cast<TD*>(ld + (int64_t)(p0 + q) * ds), outs[q]);
      } else {
        compute(u);
#pragma unroll
        for (int q = 0; q < K; ++q)
          if (p0 + q < a1) store_nt<TD, COLS>(reinterpret_cast<TD*>(ld + (int64_t)(p0 + q) * ds), outs[q]);
      }
    }
  }
}

}  // namespace clk
