// y/z pass for float sources and float outputs in blocks of 1024 threads
// that move in lockstep.  Included by sep_impl.h.
//
// Each thread walks a segment of the filtered axis for COLS adjacent
// columns, K outputs per step, like the static y/z pass, with two changes
// measured on float32 1024x1024x512 volumes:
//   - the window and the next step's rows share one register ring of
//     RS = U*K rows (U steps unrolled, so every slot is a static register
//     and no window shift moves are executed), and a step's K output rows
//     are stored together after they are computed;
//   - a block of 1024 threads (16 waves) covers 1024*COLS consecutive
//     columns and passes a barrier at the start of every step, so the
//     block's loads of one input row leave as one contiguous burst of 4-16
//     KB instead of 1 KB per wave at scattered times.  A copy with the z
//     pass's access pattern reached 74% of HBM this way against 60% for
//     256-thread blocks without the barrier; the y pass went from 0.83-0.86
//     to 0.79 ms.  The gain of the z pass depends on where the two buffers
//     lie in memory (0.77-0.89 ms against 0.85), so the autotuner decides
//     there.
// A step's loads are for the step D steps ahead (D = 1: the next one), so
// a lane keeps D*K rows of COLS columns in flight while the ring grows by
// only (D-1)*K rows; a smaller K with a larger D keeps more bytes in
// flight for the same registers.
// Lanes beyond the last column stay in the loop (the barrier needs them),
// load a valid column and store nothing.
#pragma once

namespace clk {

constexpr int kYRingThreads = 1024;

template <int OP, class TS, class TD, int C, int COLS, int K, int D>
__global__ __launch_bounds__(kYRingThreads) void sep_yring_kernel(const SepParams p) {
  using VT = xr_vt<OP, TS>;
  constexpr bool kFloatSrc = !Pix<TS>::kInt;
  constexpr int N = 2 * C + 1;
  constexpr int U = ((D + 1) * K + 2 * C + K - 1) / K;  // window (K + 2C rows) + D steps' new rows
  constexpr int RS = U * K;
  uint32_t bid = blockIdx.x;
  if (p.order & 2) {
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
  const int64_t g0 = (int64_t)b * kYRingThreads;
  const bool live = g0 + threadIdx.x < p.ncols_total;
  const int64_t g = live ? g0 + threadIdx.x : g0;
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

  // relative row r (input position a0 - C + r, clamped) lives in slot r % RS
  VT ring[RS][COLS];
  auto ldrow = [&](int r, VT (&dst)[COLS]) {
    int pos = a0 - C + r;
    pos = pos < 0 ? 0 : (pos > L - 1 ? L - 1 : pos);
    load_vt_stream<TS, VT, COLS>(reinterpret_cast<const TS*>(ls0 + (int64_t)pos * ss + lso), dst);
  };
#pragma unroll
  for (int i = 0; i < D * K + 2 * C; ++i) ldrow(i, ring[i]);

  TD outs[K][COLS];
  // output q of a step whose window starts at slot u*K (static)
  auto compute = [&](const int u, const int q) {
#pragma unroll
    for (int j = 0; j < COLS; ++j) {
      VT acc;
      if constexpr (OP == OP_GAUSS) {
        acc = 0.f;
#pragma unroll
        for (int t = 0; t < N; ++t) acc = __builtin_fmaf(p.w[t], ring[(u * K + q + t) % RS][j], acc);
      } else if constexpr (OP == OP_MEAN) {
        acc = 0;
#pragma unroll
        for (int t = 0; t < N; ++t) acc = acc + ring[(u * K + q + t) % RS][j];
      } else {
        acc = ring[(u * K + q + C) % RS][j];  // centre first
#pragma unroll
        for (int t = 0; t < N; ++t) acc = tap_apply<OP, kFloatSrc, VT>(acc, ring[(u * K + q + t) % RS][j], 0.f);
      }
      outs[q][j] = finish<OP, TS, TD, VT>(acc, p);
    }
  };
  auto store = [&](const int q, const int p0) {
    if (live) store_nt<TD, COLS>(reinterpret_cast<TD*>(ld0 + (int64_t)(p0 + q) * ds + ldo), outs[q]);
  };

  // step s covers outputs a0 + s*K ..; s advances U steps per pass of the
  // unrolled loop, so step s + u reads its window from slots u*K ..
  for (int s = 0;; s += U) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int p0 = a0 + (s + u) * K;
      if (p0 >= a1) return;  // block-uniform: the segment is the block's
      __syncthreads();
      if (p0 + D * K < a1) {
        // the new rows of step s+u+D (relative (s+u+D)*K + 2C + i) go to
        // slots neither this step nor the loads in flight use
#pragma unroll
        for (int i = 0; i < K; ++i)
          ldrow((s + u + D) * K + 2 * C + i, ring[((u + D) * K + 2 * C + i) % RS]);
#pragma unroll
        for (int q = 0; q < K; ++q) compute(u, q);
#pragma unroll
        for (int q = 0; q < K; ++q) store(q, p0);
      } else {
#pragma unroll
        for (int q = 0; q < K; ++q)
          if (p0 + q < a1) {
            compute(u, q);
            store(q, p0);
          }
      }
    }
  }
}

}  // namespace clk
