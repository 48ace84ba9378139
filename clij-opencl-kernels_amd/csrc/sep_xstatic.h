// x pass through LDS with the half width fixed at compile time (Gaussians
// of half width 9..16 on long rows, where the register window would need
// more than two neighbour chunks per side).  Included by sep_impl.h.
//
// A task is SUB x 512 consecutive outputs of one row.  Window element
// e (e = x - x0 + C) is kept at LDS word P(e) = e + e/8 (one pad word per 8,
// so the lanes' stride-9 accesses hit distinct banks).  A lane's chunk
// starts at e = 8(64m + l) + C and its 8 + 2C window at e = 8(64m + l), so
// both sit at 9(64m + l) plus compile-time offsets: one address per lane and
// immediate offsets for every LDS access.  The row index, the segment start
// and all base addresses are wave-uniform (scalar registers).
#pragma once

namespace clk {

// One task's inputs in registers: the lane's SUB chunks and its halo pixel.
template <class TS, int SUB>
struct XsTask {
  PixVec<TS, kXK> v[SUB];
  TS halo;
  uint32_t row;
  int64_t x0;
  bool full, live;
};

template <class TS, int C, int SUB>
__device__ __forceinline__ void xs_load(const SepParams& p, const uint32_t task, const int lane, XsTask<TS, SUB>& t) {
  constexpr int K = kXK;
  constexpr int TSEG = kXSeg * SUB;
  const int64_t W = p.L;
  t.live = task < (uint32_t)p.ntasks;  // wave-uniform
  if (!t.live) return;
  t.row = div_nseg(p, task);
  t.x0 = (int64_t)(task - t.row * (uint32_t)p.nseg) * TSEG;
  t.full = p.vec_ld && t.x0 + TSEG <= W;
  const TS* __restrict__ rp = static_cast<const TS*>(p.src) + (int64_t)t.row * W;
#pragma unroll
  for (int m = 0; m < SUB; ++m) {
    const int64_t xi = t.x0 + kXSeg * m + K * lane;
    if (t.full) {
      t.v[m] = load_nt<TS, K>(rp + xi);
    } else {
#pragma unroll
      for (int j = 0; j < K; ++j) t.v[m].v[j] = rp[xi + j < W ? xi + j : W - 1];
    }
  }
  // halo: lanes [0, C) the C pixels left of the segment, lanes [C, 2C)
  // the C right of it, clamped to the row (other lanes load nothing)
  t.halo = TS(0);
  if (lane < 2 * C) {
    const int64_t xh = lane < C ? t.x0 - C + lane : t.x0 + TSEG + (lane - C);
    t.halo = rp[clamp64(xh, 0, W - 1)];
  }
}

template <int OP, class TS, class TD, int C, int SUB>
__device__ __forceinline__ void xs_compute(const SepParams& p, const int lane, const XsTask<TS, SUB>& t,
                                           vt_t<OP, TS>* const buf) {
  using VT = vt_t<OP, TS>;
  constexpr bool kFloatSrc = !Pix<TS>::kInt;
  constexpr int K = kXK;
  constexpr int TSEG = kXSeg * SUB;
  constexpr int WIN = K + 2 * C;
  const int64_t W = p.L;
  if (lane < 2 * C) buf[lds_pad(lane < C ? lane : TSEG + lane)] = static_cast<VT>(t.halo);
  VT* const ob = buf + 9 * lane;
#pragma unroll
  for (int m = 0; m < SUB; ++m)
#pragma unroll
    for (int j = 0; j < K; ++j) ob[9 * 64 * m + lds_pad(C + j)] = static_cast<VT>(t.v[m].v[j]);
  // LDS operations of one wave complete in order; the fences keep the
  // compiler from moving the reads above the writes of other lanes
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int m = 0; m < SUB; ++m) {
    // chunk m > 0 lies beyond the reach of ds_read2's 8-bit word offsets
    // from the m = 0 address: form its base once (the empty asm keeps the
    // compiler from folding it back into one add per read)
    uint32_t off = 9u * (64u * m + (uint32_t)lane);
    if (m > 0) asm volatile("" : "+v"(off));
    const VT* const lb = buf + off;
    VT win[WIN];
#pragma unroll
    for (int i = 0; i < WIN; ++i) win[i] = lb[lds_pad(i)];
    TD out[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      VT acc;
      if constexpr (OP == OP_GAUSS) {
        acc = 0.f;
#pragma unroll
        for (int u = 0; u <= 2 * C; ++u) acc = __builtin_fmaf(p.w[u], win[q + u], acc);
      } else if constexpr (OP == OP_MEAN) {
        acc = 0;
#pragma unroll
        for (int u = 0; u <= 2 * C; ++u) acc = acc + win[q + u];
      } else {
        acc = win[q + C];  // centre first
#pragma unroll
        for (int u = 0; u <= 2 * C; ++u) acc = tap_apply<OP, kFloatSrc, VT>(acc, win[q + u], 0.f);
      }
      out[q] = finish<OP, TS, TD, VT>(acc, p);
    }
    const int64_t xs = t.x0 + kXSeg * m + (int64_t)K * lane;
    TD* __restrict__ dp = static_cast<TD*>(p.dst) + (int64_t)t.row * W + xs;
    if (t.full && p.vec_st) {
      store_nt<TD, K>(dp, out);
    } else if (xs < W) {
#pragma unroll
      for (int q = 0; q < K; ++q)
        if (xs + q < W) dp[q] = out[q];
    }
  }
  // the next task's staging must not overtake this task's reads
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A wave filters R consecutive tasks, loading task r+1 into registers
// before task r is staged and computed.  Each wave stages into and reads
// from its own LDS rows only, so there is no block barrier: a wave without
// work just leaves.
template <int OP, class TS, class TD, int C, int SUB, int R>
__device__ __forceinline__ void sep_xs_body(const SepParams& p, const uint32_t vblk, uint32_t* lds_raw) {
  static_assert(C >= 1 && C <= 32, "halo lanes: 2C <= 64");
  using VT = vt_t<OP, TS>;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  VT* const buf = reinterpret_cast<VT*>(lds_raw) + wv * p.lds_stride;
  const uint32_t task0 = (vblk * 4u + (uint32_t)wv) * R;
  XsTask<TS, SUB> t[R];
  xs_load<TS, C, SUB>(p, task0, lane, t[0]);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (!t[r].live) return;
    if (r + 1 < R) xs_load<TS, C, SUB>(p, task0 + r + 1, lane, t[r + 1]);
    xs_compute<OP, TS, TD, C, SUB>(p, lane, t[r], buf);
  }
}

template <int OP, class TS, class TD, int C, int SUB, int R>
__global__ __launch_bounds__(256) void sep_xs_kernel(const SepParams p) {
  extern __shared__ uint32_t lds_raw[];
  sep_xs_body<OP, TS, TD, C, SUB, R>(p, blockIdx.x, lds_raw);
}

}  // namespace clk
