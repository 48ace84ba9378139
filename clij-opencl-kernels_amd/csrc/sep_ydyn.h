// y/z pass body for a half width known only at run time, up to a
// compile-time tier CMAX (the batch engine, whose jobs each bring their own
// c).  Included by sep_impl.h.
//
// The register ring of the static passes, with every slot a compile-time
// register.  The ring is anchored on the newest row of a step's window:
// ring row r holds input row a0 + c - 2*CMAX + r (clamped), so output q of
// step n reads ring rows n*K + q + t for taps t = 2(CMAX - c) .. 2*CMAX
// (v = t - 2*CMAX + c = -c .. c, the reference's order) and step n loads the
// K rows n*K + K + 2*CMAX + k of step n + 1, whatever c is.  Every load is
// then unconditional and goes to a static slot, so the compiler counts the
// loads in flight and waits only for the rows a step uses (a ring anchored on
// the window's start puts the new rows in slots that depend on c, chosen by
// uniform branches, across which every wait becomes a wait for all loads).
// Taps before 2(CMAX - c) are skipped by wave-uniform branches.  The host
// stores the Gaussian weight of tap t at w[t] (set_ydyn_weights), so the
// weight index is static too.  The rows in use span at most 2K + 2c < RS,
// so no slot is overwritten early.
//
// FIN: the Gaussian and the mean of data known to be finite (the float
// temporaries of chains from integer images) run every tap of the tier,
// the taps before 2(CMAX - c) with weight 0 and a mean's with weight 1, as
// fma(w, x, acc).  Those taps come first, while acc is still +0, and
// fma(0, x, +0) = +0 for every finite x, so the sums are those of the
// skipping body bit for bit; without the branches the compiler keeps the
// next step's loads in flight during a step (the skipping body's branches
// are structurized into blocks across which it waits for every load).
#pragma once

namespace clk {

// Gaussian weights w[0..2c] moved to w[CMAX - c .. CMAX + c] (the x tier
// pass's centred window).
inline void set_tier_weights(SepParams& p, int cmax) {
  const int c = p.c;
  if (c == cmax) return;
  for (int t = 2 * c; t >= 0; --t) p.w[t + cmax - c] = p.w[t];
}

// Gaussian weights w[0..2c] moved to w[2(CMAX - c) .. 2*CMAX] (the y/z tier
// pass's ring, anchored on the window's newest row).  fin (the branch-free
// body): the taps before 2(CMAX - c) get weight 0, and a mean gets weight 1
// on the job's taps.
inline void set_ydyn_weights(SepParams& p, int op, int cmax, bool fin) {
  const int c = p.c, lo = 2 * (cmax - c);
  if (op == OP_MEAN) {
    if (fin)
      for (int t = 0; t <= 2 * cmax; ++t) p.w[t] = t < lo ? 0.f : 1.f;
    return;
  }
  if (c != cmax)
    for (int t = 2 * c; t >= 0; --t) p.w[t + lo] = p.w[t];
  if (fin)
    for (int t = 0; t < lo; ++t) p.w[t] = 0.f;
}

template <int OP, class TS, class TD, int CMAX, int COLS, int K, bool FIN>
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
  // rows padded to whole waves (p.ntasks = the real column groups): the
  // padding lanes have nothing to do (no barriers below)
  if (p.ntasks > 0 && cg >= p.ntasks) return;
  const int64_t eoff = o * (p.L * p.inner) + cg * COLS;
  const char* const ls = static_cast<const char*>(p.src) + eoff * (int64_t)sizeof(TS);
  char* const ld = static_cast<char*>(p.dst) + eoff * (int64_t)sizeof(TD);
  const int64_t ss = p.inner * (int64_t)sizeof(TS);
  const int64_t ds = p.inner * (int64_t)sizeof(TD);
  const int L = (int)p.L;
  const int c = __builtin_amdgcn_readfirstlane(p.c);
  const int tlo = 2 * (CMAX - c);  // taps t in [tlo, 2*CMAX] are the job's
  const int a0 = (int)sid * p.seg;
  const int a1 = (a0 + p.seg < L) ? a0 + p.seg : L;
  const int r0 = a0 + c - 2 * CMAX;  // input row of ring row 0

  VT ring[RS][COLS];
  auto ldrow = [&](int r, VT (&dst)[COLS]) {
    int pos = r0 + r;
    pos = pos < 0 ? 0 : (pos > L - 1 ? L - 1 : pos);
    load_vt<TS, VT, COLS>(reinterpret_cast<const TS*>(ls + (int64_t)pos * ss), dst);
  };
  // the first step's window: ring rows tlo .. K + 2*CMAX - 1
#pragma unroll
  for (int i = 0; i < K + 2 * CMAX; ++i)
    if (FIN || i >= tlo) ldrow(i, ring[i]);

  TD outs[K][COLS];
  // the K outputs of a step whose window starts at slot u*K (static),
  // tap-major: output q still sees its taps in order t = tlo .. 2*CMAX
  auto compute = [&](const int u) {
    VT acc[K][COLS];
#pragma unroll
    for (int q = 0; q < K; ++q)
#pragma unroll
      for (int j = 0; j < COLS; ++j) {
        if constexpr (OP == OP_GAUSS || OP == OP_MEAN) {
          acc[q][j] = 0;
        } else if constexpr (!kFloatSrc) {
          // integer minimum/maximum is exact in any order: start from the
          // window's newest row (a static slot)
          acc[q][j] = ring[(u * K + q + 2 * CMAX) % RS][j];
        } else {
          // float: the centre first, as the reference (its slot depends
          // on c, picked by uniform branches)
#pragma unroll
          for (int cc = 0; cc <= CMAX; ++cc)
            if (cc == c) acc[q][j] = ring[(u * K + q + 2 * CMAX - cc) % RS][j];
        }
      }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if constexpr (!FIN) {
        if (t < tlo) continue;  // wave-uniform
      }
      const float h = (OP == OP_GAUSS || FIN) ? p.w[t] : 0.f;
#pragma unroll
      for (int q = 0; q < K; ++q)
#pragma unroll
        for (int j = 0; j < COLS; ++j) {
          const VT x = ring[(u * K + q + t) % RS][j];
          if constexpr (FIN || OP == OP_GAUSS) acc[q][j] = __builtin_fmaf(h, x, acc[q][j]);
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
        // the next step's K new rows, into static slots
#pragma unroll
        for (int k = 0; k < K; ++k) ldrow((s + u) * K + K + 2 * CMAX + k, ring[(u * K + K + 2 * CMAX + k) % RS]);
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
