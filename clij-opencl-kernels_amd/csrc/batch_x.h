// x pass for short rows (at most 512 pixels) in the batch engine: a wave
// takes k whole consecutive rows (k (W + 2c) <= 512 + 2c), i.e. one
// contiguous span of the volume, and
//   1. loads the span with aligned 16-byte vector loads (the elements before
//      the first and after the last aligned address one by one) into LDS as
//      padded rows (c edge pixels, the row, c edge pixels);
//   2. computes the 512 padded positions c + 8*lane + q with the register
//      ring of the LDS x pass (x_ring), dropping the halo positions;
//   3. writes the outputs densely to a second LDS buffer and stores the span
//      with aligned 16-byte vector stores (ends one by one).
// Each wave uses its own LDS, so waves synchronise only with themselves.
#pragma once

namespace clk {

// LDS bytes per wave: the padded rows (p.lds_stride 4-byte words, as
// plan_xpack sets it) and the dense outputs (512 pixels plus alignment).
template <class TD>
constexpr size_t xrows_obuf_bytes() { return 512 * sizeof(TD) + 16; }

template <int OP, class TS, class TD>
__device__ __forceinline__ void sep_xrows_body(const SepParams& p, const uint32_t vblk, uint32_t* lds_raw) {
  using VT = vt_t<OP, TS>;
  constexpr int K = kXK;
  constexpr int EI = 16 / (int)sizeof(TS);  // pixels per 16-byte load
  constexpr int EO = 16 / (int)sizeof(TD);  // pixels per 16-byte store
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const uint32_t task = vblk * 4u + (uint32_t)wv;
  if (task >= (uint32_t)p.ntasks) return;
  VT* const buf = reinterpret_cast<VT*>(lds_raw) + wv * p.lds_stride;
  TD* const obuf = reinterpret_cast<TD*>(reinterpret_cast<char*>(lds_raw) + (size_t)4 * 4 * p.lds_stride +
                                         (size_t)wv * xrows_obuf_bytes<TD>());
  const int W = (int)p.L;
  const int c = p.c;
  const int PW = W + 2 * c;
  const int64_t row0 = (int64_t)task * p.nseg;
  const int k = (int)(p.outer - row0 < p.nseg ? p.outer - row0 : p.nseg);
  const int n = k * W;
  const CLK_GLOBAL TS* __restrict__ s0 = gmem(static_cast<const TS*>(p.src)) + row0 * W;
  CLK_GLOBAL TD* __restrict__ d0 = gmem(static_cast<TD*>(p.dst)) + row0 * W;
  const float rW = 1.0f / (float)W;
  // (row, x) of span element f < 512 + W: a float quotient, then corrected
  auto rowx = [&](int f, int& r, int& x) {
    r = (int)((float)f * rW);
    x = f - r * W;
    if (x < 0) { --r; x += W; }
    if (x >= W) { ++r; x -= W; }
  };

  // 1. the span into padded rows
  // elements before the first 16-byte boundary (all of them one by one if
  // the buffer is not even aligned to its pixel size)
  const int hi = ((uintptr_t)s0 % sizeof(TS)) ? n : (int)(((16 - ((uintptr_t)s0 & 15)) & 15) / sizeof(TS));
  const int head = hi < n ? hi : n;
  const int nch = (n - head) / EI;
  for (int f = lane; f < head; f += 64) {
    int r, x;
    rowx(f, r, x);
    buf[lds_pad(r * PW + c + x)] = static_cast<VT>(s0[f]);
  }
  for (int ci = lane; ci < nch; ci += 64) {
    const int f0 = head + ci * EI;
    const PixVec<TS, EI> v = load_nt<TS, EI>(gen(s0 + f0));
    int r, x;
    rowx(f0, r, x);
#pragma unroll
    for (int j = 0; j < EI; ++j) {
      buf[lds_pad(r * PW + c + x)] = static_cast<VT>(v.v[j]);
      if (++x == W) {
        x = 0;
        ++r;
      }
    }
  }
  for (int f = head + nch * EI + lane; f < n; f += 64) {
    int r, x;
    rowx(f, r, x);
    buf[lds_pad(r * PW + c + x)] = static_cast<VT>(s0[f]);
  }
  // edge pixels of the halos
  for (int i = lane; i < 2 * c * k; i += 64) {
    const int r = i / (2 * c);
    const int e = i - r * 2 * c;
    if (e < c) buf[lds_pad(r * PW + e)] = static_cast<VT>(s0[(int64_t)r * W]);
    else buf[lds_pad(r * PW + W + e)] = static_cast<VT>(s0[(int64_t)r * W + W - 1]);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  // 2. 8 padded positions per lane
  VT acc[K];
  x_ring<OP, TS, VT>(buf + (K + 1) * lane, c, p.ntaps, p, acc);
  {
    // position i = c + 8 lane + q -> padded row i / PW, x = i mod PW - c
    int pr = (c + K * lane) / PW;
    int px = c + K * lane - pr * PW - c;
#pragma unroll
    for (int q = 0; q < K; ++q) {
      if (pr < k && px >= 0 && px < W) obuf[pr * W + px] = finish<OP, TS, TD, VT>(acc[q], p);
      if (++px == W + c) {
        px = -c;
        ++pr;
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  // 3. the dense outputs to the span
  const int ho = ((uintptr_t)d0 % sizeof(TD)) ? n : (int)(((16 - ((uintptr_t)d0 & 15)) & 15) / sizeof(TD));
  const int headd = ho < n ? ho : n;
  const int nco = (n - headd) / EO;
  for (int f = lane; f < headd; f += 64) d0[f] = obuf[f];
  for (int ci = lane; ci < nco; ci += 64) {
    const int f0 = headd + ci * EO;
    TD o[EO];
#pragma unroll
    for (int j = 0; j < EO; ++j) o[j] = obuf[f0 + j];
    store_nt<TD, EO>(gen(d0 + f0), o);
  }
  for (int f = headd + nco * EO + lane; f < n; f += 64) d0[f] = obuf[f];
}

}  // namespace clk
