// median_box_3d for gfx950.
//
// Reference: kernels/median_box_3d_x.cl:23-65.  Each output is the element
// of rank n/2 (0-based, ascending) of the n = (2ex+1)(2ey+1)(2ez+1)
// clamp-to-edge box values, each value taken as (dst_t)(float)src
// (median_box_3d_x.cl:54-55), written without CONVERT (:64).  The reference
// finds it with an insertion sort of a private array (:3-20); the value of
// that element does not depend on how it is found, so the kernels here use
// selection methods that fit registers and the VALU:
//
//  * 3x3x3 boxes: the 27 values sit in registers as order-preserving keys
//    and Batcher's merge sort, pruned to its middle output, leaves the
//    median.  For 8- and 16-bit dst two voxels share each instruction
//    (16-bit pairs) and single launches slide along z so that a step stages
//    one new plane of rows.
//  * 8-bit dst and up to 255 values: each lane walks a z-column of outputs
//    with a 256-bin histogram of its box in LDS (one byte per bin); a step
//    in z removes the leaving plane, adds the entering one and moves the
//    tracked median a few bins (Huang's running median, along z).  Counter
//    dword d of lane l sits at d*64 + l, so every lane stays in its own
//    LDS bank.
//  * any other box: a bisection between the window's smallest and largest
//    key counts, per probe, how many window values are <= the probe (at most
//    8, 16 or 32 probes; a count of exactly n/2 + 1 ends it, the result being
//    the largest value <= that probe).  No array, so no MAX_ARRAY_SIZE limit.
//    Boxes within 7x7x1 or 5x5x5 (16- and 32-bit dst) keep the window's keys
//    in registers, read once; a box that fills its register tier exactly
//    takes the pruned network instead when its key range is wide.
//
// Integer pixel types are exact.  For float dst the value of the rank-n/2
// element is reproduced; a zero median is returned as +0 and NaN inputs have
// no defined order (the reference's insertion sort is order-dependent there).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>

#include "clk_device.h"
#include "clk_internal.h"
#include "clk_tuning.h"

namespace clk {

inline bool hist8_enabled() { return tuning(TUNE_NO_HIST8) == 0; }

// A source value as the reference stores it in the median's dst_t array.
// median_box_3d (wrap = 0): (dst_t)(float)READ(src), a C conversion of the
// float; in range it truncates toward zero, out of range C leaves it
// undefined and this one saturates (NaN -> 0).  median_box_2d (wrap = 1):
// (dst_t)READ(src).x, a C cast of the widened int/uint, which for an integer
// dst keeps the low bits (two's complement); float sources convert as above.
template <class TS, class TD>
__device__ __forceinline__ float gather_value(TS x, int32_t wrap) {
  if constexpr (Pix<TS>::kInt && Pix<TD>::kInt) {
    if (wrap) return static_cast<float>(static_cast<TD>(static_cast<int32_t>(x)));
  }
  return static_cast<float>(convert_sat<TD>(static_cast<float>(x)));
}

// Order-preserving integer key of a dst_t value (held in a float).
template <class TD>
__device__ __forceinline__ uint32_t value_key(float v) {
  if constexpr (Pix<TD>::kInt) {
    return static_cast<uint32_t>(static_cast<int32_t>(v) - static_cast<int32_t>(Pix<TD>::lo));
  } else {
    if (v == 0.f) v = 0.f;  // -0 and +0 are one value
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  }
}

template <class TD>
__device__ __forceinline__ float key_value(uint32_t k) {
  if constexpr (Pix<TD>::kInt) {
    return static_cast<float>(static_cast<int32_t>(k) + static_cast<int32_t>(Pix<TD>::lo));
  } else {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
  }
}

template <class TD> struct KeyBits { static constexpr int v = sizeof(TD) * 8; };
template <> struct KeyBits<float> { static constexpr int v = 32; };

}  // namespace clk
#include "clk_median.h"
namespace clk {

template <class TS, class TD>
__device__ __forceinline__ void median_generic_body(const MedParams& p, uint32_t bx, uint32_t by, uint32_t bz) {
  const int64_t x = (int64_t)bx * 256 + threadIdx.x;
  const int64_t y = by;
  const int64_t z = bz;
  if (x >= p.W) return;
  const CLK_GLOBAL TS* __restrict__ src = gmem(static_cast<const TS*>(p.src));
  // f(key) for every value of the window
  auto each = [&](auto&& f) {
    for (int dz = -p.ez; dz <= p.ez; ++dz) {
      const int64_t zz = clamp64(z + dz, 0, p.D - 1);
      for (int dy = -p.ey; dy <= p.ey; ++dy) {
        const int64_t yy = clamp64(y + dy, 0, p.H - 1);
        const CLK_GLOBAL TS* rowp = src + (zz * p.H + yy) * p.W;
        for (int dx = -p.ex; dx <= p.ex; ++dx) f(value_key<TD>(gather_value<TS, TD>(rowp[clamp64(x + dx, 0, p.W - 1)], p.wrap)));
      }
    }
  };
  // The result r (rank k) lies between the window's smallest and largest
  // key.  A probe whose count of keys <= it is exactly k + 1 ends the search:
  // r is the largest of those keys; otherwise the search ends at lo == hi == r.
  const int64_t k = p.n / 2;
  uint32_t lo = 0xffffffffu, hi = 0u;
  each([&](uint32_t v) {
    lo = v < lo ? v : lo;
    hi = v > hi ? v : hi;
  });
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    int64_t cnt = 0;
    each([&](uint32_t v) { cnt += v <= mid; });
    if (cnt > k) {
      hi = mid;
      if (cnt == k + 1) break;
    } else {
      lo = mid + 1;
    }
  }
  uint32_t r = 0;
  each([&](uint32_t v) { r = (v <= hi && v > r) ? v : r; });
  gmem(static_cast<TD*>(p.dst))[(z * p.H + y) * p.W + x] = static_cast<TD>(key_value<TD>(r));
}

// Median (rank N/2) of N = 2k+1 values held in registers by forgetful
// selection: keep a buffer of k+2 values; its minimum and maximum cannot be
// the median of the values seen so far plus the rest, so move them to the
// buffer ends with compare-exchanges, drop both, and put the next unseen
// value into the freed slot.  After k-1 rounds the median of the last 3 is
// the answer.
typedef unsigned short us2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float vmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float vmax(float a, float b) { return fmaxf(a, b); }
// two 16-bit lanes per instruction (v_pk_min_u16 / v_pk_max_u16)
__device__ __forceinline__ us2 vmin(us2 a, us2 b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ us2 vmax(us2 a, us2 b) { return __builtin_elementwise_max(a, b); }

template <int N, class T>
__device__ __forceinline__ T forgetful_median(T (&a)[N]) {
  constexpr int k = N / 2;
  constexpr int m = k + 2;
  constexpr int hi = m - 1;
#pragma unroll
  for (int step = 0; step < N - m; ++step) {
    const int lo = step;
#pragma unroll
    for (int i = lo + 1; i <= hi; ++i) {
      const T u = a[lo], w = a[i];
      a[lo] = vmin(u, w);
      a[i] = vmax(u, w);
    }
#pragma unroll
    for (int i = lo + 1; i < hi; ++i) {
      const T u = a[i], w = a[hi];
      a[i] = vmin(u, w);
      a[hi] = vmax(u, w);
    }
    a[hi] = a[m + step];
  }
  const T x0 = a[N - m], x1 = a[N - m + 1], x2 = a[hi];  // the last 3 (N - m + 2 == hi)
  return vmax(vmin(x0, x1), vmin(vmax(x0, x1), x2));
}

__device__ __forceinline__ uint32_t vmin(uint32_t a, uint32_t b) { return a < b ? a : b; }
__device__ __forceinline__ uint32_t vmax(uint32_t a, uint32_t b) { return a > b ? a : b; }

// Rank N/2 of N keys (16-bit key pairs, or 32-bit keys) by Batcher's
// odd-even merge sort (Knuth, TAOCP vol. 3, 5.3.4, algorithm M) on
// P = 2^ceil(log2 N) wires, wires N..P-1 held at the largest key.  Only wire
// N/2 is used: the compiler folds every comparator that meets a padding wire
// (min(x, top) = x) and drops those whose results cannot reach wire N/2,
// leaving 224 min/max instructions for N = 27 where forgetful selection
// needs 364, and 2244 for N = 125.
template <int N, class T>
__device__ __forceinline__ T batcher_median(const T (&a)[N], const T top) {
  constexpr int P = N <= 8 ? 8 : N <= 16 ? 16 : N <= 32 ? 32 : N <= 64 ? 64 : 128;
  static_assert(N <= P, "at most 128 wires");
  T w[P];
#pragma unroll
  for (int i = 0; i < P; ++i) w[i] = i < N ? a[i] : top;
#pragma unroll
  for (int p = 1; p < P; p *= 2) {
#pragma unroll
    for (int k = p; k >= 1; k /= 2) {
#pragma unroll
      for (int j = k % p; j < P - k; j += 2 * k) {
#pragma unroll
        for (int i = 0; i < k; ++i) {
          if (i < P - j - k && (i + j) / (2 * p) == (i + j + k) / (2 * p)) {
            const T u = w[i + j], v = w[i + j + k];
            w[i + j] = vmin(u, v);
            w[i + j + k] = vmax(u, v);
          }
        }
      }
    }
  }
  return w[N / 2];
}

template <int N>
__device__ __forceinline__ us2 batcher_median(const us2 (&a)[N]) {
  return batcher_median<N, us2>(a, us2{0xffff, 0xffff});
}

// Register tiers of the bisection median, by size: half extents (MX, MY, MZ)
// of a box that holds the requested one.
constexpr int kMedTiers = 6;
constexpr int8_t kMedTier[kMedTiers][3] = {{1, 1, 0}, {2, 2, 0}, {1, 1, 1}, {3, 3, 0}, {2, 2, 1}, {2, 2, 2}};
// min/max instructions of the pruned Batcher network on each tier's slots
constexpr int16_t kMedTierNetOps[kMedTiers] = {40, 202, 226, 590, 1198, 2244};

// The bisection of median_generic_body with the window read once: each
// offset of tier I's box has its own register holding the value's key, and
// every step counts over registers.  Offsets outside the requested box hold
// 0xffffffff, which no step counts (a midpoint is always below the upper
// bound, itself at most 0xffffffff).  The search starts from the window's
// smallest and largest key, between which the result lies, and may end early
// (below), so the result is the generic body's key for key.
template <class TS, class TD, int I>
__device__ __forceinline__ void median_regs_body(const MedParams& p, uint32_t bx, uint32_t by, uint32_t bz) {
  constexpr int MX = kMedTier[I][0], MY = kMedTier[I][1], MZ = kMedTier[I][2];
  constexpr int NX = 2 * MX + 1, NY = 2 * MY + 1, NS = NX * NY * (2 * MZ + 1);
  const int64_t x = (int64_t)bx * 256 + threadIdx.x;
  const int64_t y = by;
  const int64_t z = bz;
  if (x >= p.W) return;
  const CLK_GLOBAL TS* __restrict__ src = gmem(static_cast<const TS*>(p.src));
  int64_t xs[NX];
#pragma unroll
  for (int dx = -MX; dx <= MX; ++dx) xs[dx + MX] = clamp64(x + dx, 0, p.W - 1);
  uint32_t key[NS];
  uint32_t lo = 0xffffffffu, hi = 0u;
  if (p.ex == MX && p.ey == MY && p.ez == MZ && p.W * p.H * p.D * (int64_t)sizeof(TS) < (int64_t(1) << 31)) {
    // The box fills the tier: every slot is read, with no branch between
    // the loads, through a buffer resource whose row offset (one row y of
    // plane z per block) is a scalar operand and whose lane offsets are
    // the NX clamped columns: no address arithmetic per slot.
    const __amdgpu_buffer_rsrc_t r = byte_rsrc(gen(src), p.W * p.H * p.D * (int64_t)sizeof(TS));
    uint32_t xo[NX];
#pragma unroll
    for (int i = 0; i < NX; ++i) xo[i] = (uint32_t)xs[i] * (uint32_t)sizeof(TS);
#pragma unroll
    for (int dz = -MZ; dz <= MZ; ++dz) {
      const int64_t zz = clamp64(z + dz, 0, p.D - 1);
#pragma unroll
      for (int dy = -MY; dy <= MY; ++dy) {
        const int64_t yy = clamp64(y + dy, 0, p.H - 1);
        const uint32_t ro = (uint32_t)((zz * p.H + yy) * p.W * (int64_t)sizeof(TS));
#pragma unroll
        for (int dx = -MX; dx <= MX; ++dx) {
          TS v;
          if constexpr (sizeof(TS) == 1)
            v = __builtin_bit_cast(TS, __builtin_amdgcn_raw_buffer_load_b8(r, xo[dx + MX], ro, 0));
          else if constexpr (sizeof(TS) == 2)
            v = __builtin_bit_cast(TS, __builtin_amdgcn_raw_buffer_load_b16(r, xo[dx + MX], ro, 0));
          else
            v = __builtin_bit_cast(TS, __builtin_amdgcn_raw_buffer_load_b32(r, xo[dx + MX], ro, 0));
          const uint32_t k = value_key<TD>(gather_value<TS, TD>(v, p.wrap));
          lo = k < lo ? k : lo;
          hi = k > hi ? k : hi;
          key[((dz + MZ) * NY + dy + MY) * NX + dx + MX] = k;
        }
      }
    }
  } else {
#pragma unroll
  for (int dz = -MZ; dz <= MZ; ++dz) {
    const int64_t zz = clamp64(z + dz, 0, p.D - 1);
#pragma unroll
    for (int dy = -MY; dy <= MY; ++dy) {
      const int64_t yy = clamp64(y + dy, 0, p.H - 1);
      const CLK_GLOBAL TS* rowp = src + (zz * p.H + yy) * p.W;
      const bool row_in = dz >= -p.ez && dz <= p.ez && dy >= -p.ey && dy <= p.ey;  // wave-uniform
#pragma unroll
      for (int dx = -MX; dx <= MX; ++dx) {
        uint32_t k = 0xffffffffu;
        if (row_in && dx >= -p.ex && dx <= p.ex) {
          k = value_key<TD>(gather_value<TS, TD>(rowp[xs[dx + MX]], p.wrap));
          lo = k < lo ? k : lo;
          hi = k > hi ? k : hi;
        }
        key[((dz + MZ) * NY + dy + MY) * NX + dx + MX] = k;
      }
    }
  }
  }
  CLK_GLOBAL TD* const out = gmem(static_cast<TD*>(p.dst)) + (z * p.H + y) * p.W + x;
  // A box that fills the tier has its rank at a compile-time position, so
  // the pruned network applies; it is taken when the bisection would need
  // more instructions (about two per slot for each bit of the key range:
  // 8-bit data stays on the bisection, float data takes the network).
  const int range_bits = 32 - __builtin_clz((hi - lo) | 1u);
  if (p.ex == MX && p.ey == MY && p.ez == MZ && 2 * NS * range_bits > kMedTierNetOps[I]) {
    *out = static_cast<TD>(key_value<TD>(batcher_median<NS, uint32_t>(key, 0xffffffffu)));
    return;
  }
  // The result r (rank kth) stays in [lo, hi].  A step whose count of keys
  // <= hi is exactly kth + 1 ends the search: r is then the largest of those
  // keys.  Otherwise the search ends at lo == hi == r, the largest key <= r.
  const int32_t kth = p.n / 2;
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    int32_t cnt = 0;
#pragma unroll
    for (int i = 0; i < NS; ++i) cnt += key[i] <= mid;
    if (cnt > kth) {
      hi = mid;
      if (cnt == kth + 1) break;
    } else {
      lo = mid + 1;
    }
  }
  uint32_t r = 0;
#pragma unroll
  for (int i = 0; i < NS; ++i) r = (key[i] <= hi && key[i] > r) ? key[i] : r;
  *out = static_cast<TD>(key_value<TD>(r));
}

// 3x3x3 median with an 8- or 16-bit dst: a block filters 512 consecutive x
// of one row y for a run of kPk8ZRun planes.  Source values, converted to the
// dst type and biased to unsigned 16-bit keys (signed types shift by -lo,
// which keeps their order), are staged in LDS as 16-bit value pairs twice,
// pairs starting at even and at odd positions, so the three pairs a thread
// needs per row (its two outputs shifted by -1, 0, +1) are three aligned
// 32-bit reads; the selection network then runs on two voxels per
// instruction (v_pk_min_u16 / v_pk_max_u16).  LDS holds a ring of four
// planes (rows y-1, y, y+1); a z step reads the slots of planes z-1, z, z+1,
// while the loads of plane z+2 are in flight in registers, then writes them
// to the fourth slot (that of plane z-2, not read in this step) before one
// block barrier.
constexpr int kPk8ZRun = 16;

template <class TS, class TD>
__device__ __forceinline__ void median_27_pk8_slide_body(const MedParams& p, uint32_t bx, uint32_t by, uint32_t bz) {
  __shared__ uint32_t pe[4][3][257];
  __shared__ uint32_t po[4][3][256];
  const int t = threadIdx.x;
  const int64_t xt = (int64_t)bx * 512;
  const int64_t y = by;
  const int64_t z0 = (int64_t)bz * kPk8ZRun;
  const int64_t z1 = (z0 + kPk8ZRun < p.D) ? z0 + kPk8ZRun : p.D;
  const int W = (int)p.W;
  const CLK_GLOBAL TS* __restrict__ src = gmem(static_cast<const TS*>(p.src));
  constexpr int kBias = -static_cast<int>(Pix<TD>::lo);  // signed values shift to 0..2^bits-1
  // this thread's columns x = xt - 1 + 2t + i, i = 0..3 (i = 3 for t = 255 only), clamped
  int64_t xc[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int64_t xx = xt - 1 + 2 * t + i;
    xc[i] = xx < 0 ? 0 : (xx >= W ? W - 1 : xx);
  }
  int64_t rows[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) rows[r] = clamp64(y + r - 1, 0, p.H - 1) * p.W;
  const int64_t PS = p.H * p.W;
  TS raw[3][4];
  auto fetch = [&](int64_t zz) {
    const CLK_GLOBAL TS* pl = src + clamp64(zz, 0, p.D - 1) * PS;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
      for (int i = 0; i < 3; ++i) raw[r][i] = pl[rows[r] + xc[i]];
      raw[r][3] = t == 255 ? pl[rows[r] + xc[3]] : TS(0);
    }
  };
  auto put = [&](int slot) {
    // the value as the reference stores it, biased to an unsigned key; for
    // an integer source of the dst type that value is the pixel itself
    // (both of gather_value's casts are the identity): no float round trip
    auto v = [&](TS x) -> uint32_t {
      if constexpr (std::is_same<TS, TD>::value && std::is_integral<TS>::value) return (uint32_t)((int)x + kBias);
      else return (uint32_t)((int)gather_value<TS, TD>(x, p.wrap) + kBias);
    };
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const uint32_t v0 = v(raw[r][0]), v1 = v(raw[r][1]), v2 = v(raw[r][2]);
      pe[slot][r][t] = v0 | (v1 << 16);
      po[slot][r][t] = v1 | (v2 << 16);
      if (t == 255) pe[slot][r][256] = v2 | (v(raw[r][3]) << 16);
    }
  };
  // plane zz sits in slot (zz - z0 + 1) & 3
  fetch(z0 - 1);
  put(0);
  fetch(z0);
  put(1);
  fetch(z0 + 1);
  put(2);
  __syncthreads();
  const int64_t x = xt + 2 * t;
  CLK_GLOBAL TD* __restrict__ dst = gmem(static_cast<TD*>(p.dst));
  for (int64_t z = z0; z < z1; ++z) {
    const int zi = (int)(z - z0);
    const bool more = z + 1 < z1;
    if (more) fetch(z + 2);
    us2 a[27];
#pragma unroll
    for (int dz = 0; dz < 3; ++dz) {
      const int sl = (zi + dz) & 3;  // plane z - 1 + dz
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const uint32_t w0 = pe[sl][r][t], w1 = po[sl][r][t], w2 = pe[sl][r][t + 1];
        __builtin_memcpy(&a[9 * dz + 3 * r + 0], &w0, 4);
        __builtin_memcpy(&a[9 * dz + 3 * r + 1], &w1, 4);
        __builtin_memcpy(&a[9 * dz + 3 * r + 2], &w2, 4);
      }
    }
    const us2 m = batcher_median<27>(a);
    CLK_GLOBAL TD* const row = dst + (z * p.H + y) * p.W;
    if (x < W) row[x] = static_cast<TD>((int)m.x - kBias);
    if (x + 1 < W) row[x + 1] = static_cast<TD>((int)m.y - kBias);
    if (more) put((zi + 3) & 3);
    __syncthreads();
  }
}

// 3x3 in one plane (median_box_2d, median_slice_by_slice_box_3d) with an 8-
// or 16-bit dst: the same pairs and network over a run of kPk9YRun rows per
// block, sliding along y over a ring of four staged rows (rows y-1, y, y+1
// read while row y+2 is loaded into registers, then written to the fourth
// slot before one block barrier).
constexpr int kPk9YRun = 16;

template <class TS, class TD>
__device__ __forceinline__ void median_9_pk8_slide_body(const MedParams& p, uint32_t bx, uint32_t by, uint32_t bz) {
  __shared__ uint32_t pe[4][257];
  __shared__ uint32_t po[4][256];
  const int t = threadIdx.x;
  const int64_t xt = (int64_t)bx * 512;
  const int64_t y0 = (int64_t)by * kPk9YRun;
  const int64_t y1 = (y0 + kPk9YRun < p.H) ? y0 + kPk9YRun : p.H;
  const int64_t z = bz;
  const int W = (int)p.W;
  const CLK_GLOBAL TS* __restrict__ plane = gmem(static_cast<const TS*>(p.src)) + z * p.H * p.W;
  constexpr int kBias = -static_cast<int>(Pix<TD>::lo);
  int64_t xc[4];  // x = xt - 1 + 2t + i, clamped (i = 3 for t = 255 only)
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int64_t xx = xt - 1 + 2 * t + i;
    xc[i] = xx < 0 ? 0 : (xx >= W ? W - 1 : xx);
  }
  TS raw[4];
  auto fetch = [&](int64_t yy) {
    const CLK_GLOBAL TS* row = plane + clamp64(yy, 0, p.H - 1) * p.W;
#pragma unroll
    for (int i = 0; i < 3; ++i) raw[i] = row[xc[i]];
    raw[3] = t == 255 ? row[xc[3]] : TS(0);
  };
  auto put = [&](int slot) {
    // the value as the reference stores it, biased to an unsigned key; for
    // an integer source of the dst type that value is the pixel itself
    // (both of gather_value's casts are the identity): no float round trip
    auto v = [&](TS x) -> uint32_t {
      if constexpr (std::is_same<TS, TD>::value && std::is_integral<TS>::value) return (uint32_t)((int)x + kBias);
      else return (uint32_t)((int)gather_value<TS, TD>(x, p.wrap) + kBias);
    };
    const uint32_t v0 = v(raw[0]), v1 = v(raw[1]), v2 = v(raw[2]);
    pe[slot][t] = v0 | (v1 << 16);
    po[slot][t] = v1 | (v2 << 16);
    if (t == 255) pe[slot][256] = v2 | (v(raw[3]) << 16);
  };
  // row yy sits in slot (yy - y0 + 1) & 3
  fetch(y0 - 1);
  put(0);
  fetch(y0);
  put(1);
  fetch(y0 + 1);
  put(2);
  __syncthreads();
  const int64_t x = xt + 2 * t;
  CLK_GLOBAL TD* __restrict__ dst = gmem(static_cast<TD*>(p.dst)) + z * p.H * p.W;
  for (int64_t y = y0; y < y1; ++y) {
    const int yi = (int)(y - y0);
    const bool more = y + 1 < y1;
    if (more) fetch(y + 2);
    us2 a[9];
#pragma unroll
    for (int dr = 0; dr < 3; ++dr) {
      const int sl = (yi + dr) & 3;  // row y - 1 + dr
      const uint32_t w0 = pe[sl][t], w1 = po[sl][t], w2 = pe[sl][t + 1];
      __builtin_memcpy(&a[3 * dr + 0], &w0, 4);
      __builtin_memcpy(&a[3 * dr + 1], &w1, 4);
      __builtin_memcpy(&a[3 * dr + 2], &w2, 4);
    }
    const us2 m = forgetful_median<9>(a);
    CLK_GLOBAL TD* const row = dst + y * p.W;
    if (x < W) row[x] = static_cast<TD>((int)m.x - kBias);
    if (x + 1 < W) row[x + 1] = static_cast<TD>((int)m.y - kBias);
    if (more) put((yi + 3) & 3);
    __syncthreads();
  }
}

template <class TS, class TD>
__device__ __forceinline__ void median_27_body(const MedParams& p, uint32_t bx, uint32_t by, uint32_t bz) {
  const int64_t x = (int64_t)bx * 256 + threadIdx.x;
  const int64_t y = by;
  const int64_t z = bz;
  if (x >= p.W) return;
  const CLK_GLOBAL TS* __restrict__ src = gmem(static_cast<const TS*>(p.src));
  // order-preserving keys (-0 and +0 are one key), one voxel per instruction
  uint32_t a[27];
  int i = 0;
#pragma unroll
  for (int dz = -1; dz <= 1; ++dz) {
    const int64_t zz = clamp64(z + dz, 0, p.D - 1);
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
      const int64_t yy = clamp64(y + dy, 0, p.H - 1);
      const CLK_GLOBAL TS* rowp = src + (zz * p.H + yy) * p.W;
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        const int64_t xx = clamp64(x + dx, 0, p.W - 1);
        a[i++] = value_key<TD>(gather_value<TS, TD>(rowp[xx], p.wrap));
      }
    }
  }
  const uint32_t m = batcher_median<27, uint32_t>(a, 0xffffffffu);
  gmem(static_cast<TD*>(p.dst))[(z * p.H + y) * p.W + x] = static_cast<TD>(key_value<TD>(m));
}

// 3x3x3 median with a float dst in single launches: a block filters 256
// consecutive x of one row y for a run of kF27ZRun planes.  The 32-bit keys
// of the source values are staged in LDS, three rows of 258 per plane, in a
// ring of four planes; a z step reads planes z-1, z, z+1 while the loads of
// plane z+2 are in flight in registers, then writes them to the fourth slot
// (that of plane z-2, not read in this step) before one block barrier.
constexpr int kF27ZRun = 16;

template <class TS, class TD>
__device__ __forceinline__ void median_27_key_slide_body(const MedParams& p, uint32_t bx, uint32_t by, uint32_t bz) {
  __shared__ uint32_t ring[4][3][258];
  const int t = threadIdx.x;
  const int64_t xt = (int64_t)bx * 256;
  const int64_t y = by;
  const int64_t z0 = (int64_t)bz * kF27ZRun;
  const int64_t z1 = (z0 + kF27ZRun < p.D) ? z0 + kF27ZRun : p.D;
  const CLK_GLOBAL TS* __restrict__ src = gmem(static_cast<const TS*>(p.src));
  // staged columns x = xt - 1 + c: c = t, and c = 256 + t for t < 2; clamped
  const int64_t xa = clamp64(xt - 1 + t, 0, p.W - 1);
  const int64_t xb = clamp64(xt + 255 + t, 0, p.W - 1);
  const bool two = t < 2;
  int64_t rows[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) rows[r] = clamp64(y + r - 1, 0, p.H - 1) * p.W;
  const int64_t PS = p.H * p.W;
  TS va[3], vb[3];
  auto fetch = [&](int64_t zz) {
    const CLK_GLOBAL TS* pl = src + clamp64(zz, 0, p.D - 1) * PS;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      va[r] = pl[rows[r] + xa];
      vb[r] = two ? pl[rows[r] + xb] : TS(0);
    }
  };
  auto put = [&](int slot) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      ring[slot][r][t] = value_key<TD>(gather_value<TS, TD>(va[r], p.wrap));
      if (two) ring[slot][r][256 + t] = value_key<TD>(gather_value<TS, TD>(vb[r], p.wrap));
    }
  };
  // plane zz sits in slot (zz - z0 + 1) & 3
  fetch(z0 - 1);
  put(0);
  fetch(z0);
  put(1);
  fetch(z0 + 1);
  put(2);
  __syncthreads();
  const int64_t x = xt + t;
  CLK_GLOBAL TD* __restrict__ dst = gmem(static_cast<TD*>(p.dst));
  for (int64_t z = z0; z < z1; ++z) {
    const int zi = (int)(z - z0);
    const bool more = z + 1 < z1;
    if (more) fetch(z + 2);
    uint32_t a[27];
#pragma unroll
    for (int dz = 0; dz < 3; ++dz) {
      const int sl = (zi + dz) & 3;  // plane z - 1 + dz
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int i = 0; i < 3; ++i) a[9 * dz + 3 * r + i] = ring[sl][r][t + i];
    }
    const uint32_t m = batcher_median<27, uint32_t>(a, 0xffffffffu);
    if (x < p.W) dst[(z * p.H + y) * p.W + x] = static_cast<TD>(key_value<TD>(m));
    if (more) put((zi + 3) & 3);
    __syncthreads();
  }
}

// 8-bit dst: histogram bin of a gathered value.
template <class TD>
__device__ __forceinline__ int bin8(float v) {
  return static_cast<int>(v) - static_cast<int>(Pix<TD>::lo);
}

constexpr int kHistZSeg = 64;  // outputs per lane along z

// Shared by the histogram median bodies: 256 byte counters per lane packed
// four to a dword; dword d of lane l is stored at d*64 + l, so every lane
// always uses its own bank.  Counters change by LDS atomic adds whose result
// is not returned (no dependency chain between updates); -1 on a byte is
// +(2^32 - 2^(8j)), which cannot borrow because a count never goes below 0.
struct Hist8Lane {
  uint32_t* h;  // this lane's column of the block's counters
  __device__ __forceinline__ void clear() const {
#pragma unroll 8
    for (int i = 0; i < 64; ++i) h[i * 64] = 0u;
  }
  __device__ __forceinline__ uint32_t* word(int b) const { return h + (b >> 2) * 64; }
  __device__ __forceinline__ void add(int b, bool inc) const {
    const uint32_t one = 1u << ((b & 3) * 8);
    __hip_atomic_fetch_add(word(b), inc ? one : 0u - one, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ __forceinline__ int count(int b) const { return (int)((*word(b) >> ((b & 3) * 8)) & 255u); }
  // the same update for a key b << 6: its counter dword sits key & 0x3f00
  // bytes from h, the byte shift inside it is (key >> 3) & 31
  __device__ __forceinline__ void add_key(int key, bool inc) const {
    const uint32_t one = 1u << ((key >> 3) & 31);
    __hip_atomic_fetch_add(reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(h) + (key & 0x3f00)),
                           inc ? one : 0u - one, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  // The first median of a segment: whole counter dwords (four bins, summed
  // by v_sad_u8 against 0) are skipped while they stay at or below rank k,
  // then the bins of the dword that crosses it are resolved in registers.
  __device__ __forceinline__ void walk_from_zero(int& med, int& lt, int k) const {
    int d = 0;
    lt = 0;
    uint32_t w = h[0];
    for (int s = (int)__builtin_amdgcn_sad_u8(w, 0u, 0u); lt + s <= k;
         w = h[++d * 64], s = (int)__builtin_amdgcn_sad_u8(w, 0u, 0u))
      lt += s;
    med = d * 4;
    for (int c = (int)(w & 255u); lt + c <= k; c = (int)((w >> (8 * (med & 3))) & 255u)) {
      lt += c;
      ++med;
    }
  }
  // median = the bin m with lt(m) <= k < lt(m) + count(m); lt = number of
  // values below med on entry and on return
  __device__ __forceinline__ void walk(int& med, int& lt, int k) const {
    while (lt > k) {
      --med;
      lt -= count(med);
    }
    for (int c = count(med); lt + c <= k; c = count(med)) {
      lt += c;
      ++med;
    }
  }
};

// bin of a source value for an 8-bit dst: for the same type the value itself
template <class TS, class TD>
__device__ __forceinline__ int hist8_bin(TS v, int wrap) {
  if constexpr (std::is_same<TS, TD>::value) return (int)v - static_cast<int>(Pix<TD>::lo);
  else return bin8<TD>(gather_value<TS, TD>(v, wrap));
}

// Any window (half widths at run time): each lane loads every window
// position itself.
template <class TS, class TD>
__device__ __forceinline__ void median_hist8_body(const MedParams& p, uint32_t bx, uint32_t by, uint32_t bz) {
  __shared__ uint32_t hist[64 * 64];
  const int lane = threadIdx.x;
  const int64_t x = (int64_t)bx * 64 + lane;
  const int64_t y = by;
  const int64_t z0 = (int64_t)bz * kHistZSeg;
  const int64_t z1 = (z0 + kHistZSeg < p.D) ? z0 + kHistZSeg : p.D;
  if (x >= p.W) return;  // lanes never share data: no barrier follows
  const Hist8Lane hl{hist + lane};
  hl.clear();
  const CLK_GLOBAL TS* __restrict__ src = gmem(static_cast<const TS*>(p.src));
  const int ex = p.ex, ey = p.ey, ez = p.ez;
  const int W = (int)p.W;
  const int xi = (int)x;
  // the whole block's x window inside the row: no clamp, offsets from x
  const bool interior = (int64_t)bx * 64 - ex >= 0 && (int64_t)bx * 64 + 63 + ex < p.W;
  // row pointer of a (z, y) row: uniform across the wave
  auto rowp = [&](int64_t zz, int64_t yy) -> const CLK_GLOBAL TS* {
    zz = clamp64(zz, 0, p.D - 1);
    yy = clamp64(yy, 0, p.H - 1);
    return src + (zz * p.H + yy) * p.W;
  };
  auto bin_at = [&](const CLK_GLOBAL TS* row, int dx) -> int {
    int xx = xi + dx;
    if (!interior) xx = xx < 0 ? 0 : (xx >= W ? W - 1 : xx);
    return hist8_bin<TS, TD>(row[xx], p.wrap);
  };
  const int k = p.n / 2;
  for (int dz = -ez; dz <= ez; ++dz)
    for (int dy = -ey; dy <= ey; ++dy) {
      const CLK_GLOBAL TS* row = rowp(z0 + dz, y + dy);
      for (int dx = -ex; dx <= ex; ++dx) hl.add(bin_at(row, dx), true);
    }
  int med = 0, lt = 0;
  hl.walk_from_zero(med, lt, k);
  CLK_GLOBAL TD* __restrict__ dst = gmem(static_cast<TD*>(p.dst));
  const int64_t plane = p.H * p.W;
  dst[z0 * plane + y * p.W + x] = static_cast<TD>(med + static_cast<int>(Pix<TD>::lo));
  for (int64_t z = z0 + 1; z < z1; ++z) {
    const int64_t zo = z - 1 - ez, zi = z + ez;
    for (int dy = -ey; dy <= ey; ++dy) {
      const CLK_GLOBAL TS* ro = rowp(zo, y + dy);
      const CLK_GLOBAL TS* ri = rowp(zi, y + dy);
      for (int dx = -ex; dx <= ex; ++dx) {
        const int a = bin_at(ro, dx);
        const int b = bin_at(ri, dx);
        hl.add(a, false);
        hl.add(b, true);
        lt += (b < med) - (a < med);
      }
    }
    hl.walk(med, lt, k);
    dst[z * plane + y * p.W + x] = static_cast<TD>(med + static_cast<int>(Pix<TD>::lo));
  }
}

// x and y half widths fixed at compile time (z at run time).  Each lane
// loads one value per window row; the bins of its x neighbours come from the
// adjacent lanes by DPP wave shifts, lanes 0 and 63 loading the clamped halo
// (x-1-k, resp. x+1+k).  The rows a z step needs (the plane leaving and the
// plane entering the window) are loaded one step ahead into the other of two
// register buffers, so their latency overlaps the previous step's counter
// updates and median walk.  Lanes past the row end act as its last pixel
// (the clamp) and store nothing.  Bins travel as keys b << 6: the byte
// offset of b's counter dword is then key & 0x3f00, the byte shift inside it
// (key >> 3) & 31, and keys order like bins.
template <class TS, class TD, int EX, int EY>
__device__ __forceinline__ void median_hist8_fixed_body(const MedParams& p, uint32_t bx, uint32_t by, uint32_t bz) {
  constexpr int NX = 2 * EX + 1, NY = 2 * EY + 1;
  // integer sources stay in registers extended to int by the load
  using Raw = typename std::conditional<std::is_integral<TS>::value, int32_t, TS>::type;
  __shared__ uint32_t hist[64 * 64];
  const int lane = threadIdx.x;
  const int64_t x = (int64_t)bx * 64 + lane;
  const int64_t y = by;
  const int64_t z0 = (int64_t)bz * kHistZSeg;
  const int64_t z1 = (z0 + kHistZSeg < p.D) ? z0 + kHistZSeg : p.D;
  const bool valid = x < p.W;
  const Hist8Lane hl{hist + lane};
  hl.clear();
  auto add_key = [&](int key, bool inc) { hl.add_key(key, inc); };
  auto key_of = [&](Raw r) -> int {
    if constexpr (std::is_same<TS, TD>::value && std::is_integral<TS>::value)
      return (r - static_cast<int>(Pix<TD>::lo)) << 6;
    else return hist8_bin<TS, TD>(static_cast<TS>(r), p.wrap) << 6;
  };
  const CLK_GLOBAL TS* __restrict__ src = gmem(static_cast<const TS*>(p.src));
  const int W = (int)p.W;
  const int ez = p.ez;
  const int xi = valid ? (int)x : W - 1;
  const bool edge_lane = lane == 0 || lane == 63;
  int hx[EX];  // halo columns of lanes 0 and 63, clamped into the row
#pragma unroll
  for (int k = 0; k < EX; ++k) {
    const int xx = xi + (lane == 0 ? -(k + 1) : k + 1);
    hx[k] = xx < 0 ? 0 : (xx >= W ? W - 1 : xx);
  }
  int64_t rows[NY];  // element offsets of the window rows within a plane
#pragma unroll
  for (int j = 0; j < NY; ++j) rows[j] = clamp64(y + j - EY, 0, p.H - 1) * p.W;
  struct Plane {
    Raw own[NY];
    Raw halo[NY][EX];
  };
  auto load = [&](int64_t zz, Plane& r) {
    const CLK_GLOBAL TS* pl = src + clamp64(zz, 0, p.D - 1) * p.H * p.W;
#pragma unroll
    for (int j = 0; j < NY; ++j) r.own[j] = static_cast<Raw>(pl[rows[j] + xi]);
#pragma unroll
    for (int j = 0; j < NY; ++j)
#pragma unroll
      for (int k = 0; k < EX; ++k) r.halo[j][k] = Raw(0);
    if (edge_lane) {
#pragma unroll
      for (int j = 0; j < NY; ++j)
#pragma unroll
        for (int k = 0; k < EX; ++k) r.halo[j][k] = static_cast<Raw>(pl[rows[j] + hx[k]]);
    }
  };
  // keys of window row j at x-EX .. x+EX
  auto keys = [&](const Plane& r, int j, int (&b)[NX]) {
    b[EX] = key_of(r.own[j]);
#pragma unroll
    for (int k = 0; k < EX; ++k) {
      const uint32_t e = (uint32_t)key_of(r.halo[j][k]);
      b[EX - 1 - k] = (int)dpp_from_left(e, (uint32_t)b[EX - k]);
      b[EX + 1 + k] = (int)dpp_from_right(e, (uint32_t)b[EX + k]);
    }
  };
  const int k = p.n / 2;
  for (int dz = -ez; dz <= ez; ++dz) {
    Plane r;
    load(z0 + dz, r);
#pragma unroll
    for (int j = 0; j < NY; ++j) {
      int b[NX];
      keys(r, j, b);
#pragma unroll
      for (int i = 0; i < NX; ++i) add_key(b[i], true);
    }
  }
  int med = 0, lt = 0;
  hl.walk_from_zero(med, lt, k);
  CLK_GLOBAL TD* __restrict__ dst = gmem(static_cast<TD*>(p.dst));
  const int64_t plane = p.H * p.W;
  if (valid) dst[z0 * plane + y * p.W + x] = static_cast<TD>(med + static_cast<int>(Pix<TD>::lo));
  // step z removes plane z-1-ez and adds plane z+ez, and loads the planes of
  // step z+1; loads past the volume clamp, so the prefetch after the last
  // step stays in bounds
  auto step = [&](int64_t z, const Plane& co, const Plane& ci, Plane& no, Plane& ni) {
    load(z - ez, no);
    load(z + 1 + ez, ni);
    const int mkey = med << 6;
#pragma unroll
    for (int j = 0; j < NY; ++j) {
      int a[NX], b[NX];
      keys(co, j, a);
      keys(ci, j, b);
#pragma unroll
      for (int i = 0; i < NX; ++i) {
        add_key(a[i], false);
        add_key(b[i], true);
        lt += (b[i] < mkey) - (a[i] < mkey);
      }
    }
    hl.walk(med, lt, k);
    if (valid) dst[z * plane + y * p.W + x] = static_cast<TD>(med + static_cast<int>(Pix<TD>::lo));
  };
  Plane ao, ai, bo, bi;
  load(z0 - ez, ao);
  load(z0 + 1 + ez, ai);
  for (int64_t z = z0 + 1; z < z1; z += 2) {
    step(z, ao, ai, bo, bi);
    if (z + 1 < z1) step(z + 1, bo, bi, ao, ai);
  }
}

// Order LDS accesses of the lanes of one wave (the block of the column
// histogram kernel): a wave's LDS instructions execute in issue order, so
// only the compiler must not move them across this point; a workgroup
// barrier would also wait for the loads in flight for the next step.
__device__ __forceinline__ void wave_lds_order() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// Column histograms: the 5x5x5 (and any EX <= 3, EY <= 3) window at x is the
// union of the 2EX+1 "columns" at x-EX .. x+EX, each column being the
// (2EY+1) x (2ez+1) window of the y/z plane at one x.  Lane l keeps the
// histogram of its own column (x = bx*(64-2EX) + l - EX, clamped into the
// row, so lanes past the row ends hold the edge columns the clamp reads), so
// a z step changes 2(2EY+1) counters per lane instead of 2(2EX+1)(2EY+1);
// lanes EX .. 63-EX store outputs.  A column keeps 256 byte counters (64
// dwords, dword d of lane l at d*64 + l) and 16 coarse byte counters of 16
// bins each (4 dwords after them).  The window's counts are the byte-wise
// sums of 2EX+1 lanes' dwords (no byte overflows: a count is at most
// n <= 255), and its median is found without a walk: the coarse bin holding
// rank n/2 from the 4 coarse window dwords (dword sums by v_sad_u8, then
// bytes, compared branch-free), then the bin inside it from the 4 fine
// dwords it spans.  Every lane does the same work each step.
template <class TS, class TD, int EX, int EY>
__device__ __forceinline__ void median_hist8_cols_body(const MedParams& p, uint32_t bx, uint32_t by, uint32_t bz) {
  constexpr int NY = 2 * EY + 1;
  constexpr int OUTW = 64 - 2 * EX;
  using Raw = typename std::conditional<std::is_integral<TS>::value, int32_t, TS>::type;
  __shared__ uint32_t hist[68 * 64];
  const int lane = threadIdx.x;
  const int W = (int)p.W;
  const int64_t xo = (int64_t)bx * OUTW + lane - EX;
  const int xi = xo < 0 ? 0 : (xo >= W ? W - 1 : (int)xo);
  const bool out = lane >= EX && lane < 64 - EX && xo < W;
  const int64_t y = by;
  // z in 32 bits (D < 2^31, checked by the host): uniform 32-bit compares
  // stay on the scalar unit
  const int D = (int)p.D;
  const int z0 = (int)bz * kHistZSeg;
  const int z1 = (z0 + kHistZSeg < D) ? z0 + kHistZSeg : D;
#pragma unroll 4
  for (int i = 0; i < 68; ++i) hist[i * 64 + lane] = 0u;
  auto key_of = [&](Raw r) -> int {
    if constexpr (std::is_same<TS, TD>::value && std::is_integral<TS>::value)
      return (r - static_cast<int>(Pix<TD>::lo)) << 6;
    else return hist8_bin<TS, TD>(static_cast<TS>(r), p.wrap) << 6;
  };
  // a key b << 6 (b the bin): fine counter dword at byte (key & 0x3f00) from
  // the lane's column, byte shift (key >> 3) & 31; coarse bin b >> 4 in
  // coarse dword key >> 12, byte shift (key >> 7) & 24
  char* const col = reinterpret_cast<char*>(hist + lane);
  auto update = [&](int key, bool inc) {
    const uint32_t one = 1u << ((key >> 3) & 31);
    const uint32_t cone = 1u << ((key >> 7) & 24);
    __hip_atomic_fetch_add(reinterpret_cast<uint32_t*>(col + (key & 0x3f00)), inc ? one : 0u - one,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(reinterpret_cast<uint32_t*>(col + 64 * 64 * 4 + ((key >> 12) << 8)),
                           inc ? cone : 0u - cone, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  };
  const CLK_GLOBAL TS* __restrict__ src = gmem(static_cast<const TS*>(p.src));
  const int ez = p.ez;
  // byte offsets of the lane's pixels in a plane (a plane is under 2^31
  // bytes, checked by the host): each load is a buffer load from the
  // plane's resource at a 32-bit lane offset
  uint32_t roff[NY];
#pragma unroll
  for (int j = 0; j < NY; ++j)
    roff[j] = (uint32_t)((clamp64(y + j - EY, 0, p.H - 1) * p.W + xi) * (int64_t)sizeof(TS));
  const int64_t pbytes = p.H * p.W * (int64_t)sizeof(TS);
  // the raw pixels of plane zz (clamped) at the lane's NY rows; their keys
  // are formed where they are used, so a load in flight is not waited for
  auto load = [&](int zz, Raw (&v)[NY]) {
    zz = zz < 0 ? 0 : (zz > D - 1 ? D - 1 : zz);
    const __amdgpu_buffer_rsrc_t r = byte_rsrc(reinterpret_cast<const char*>(gen(src)) + (int64_t)zz * pbytes, pbytes);
#pragma unroll
    for (int j = 0; j < NY; ++j) v[j] = static_cast<Raw>(rsrc_px<TS>(r, roff[j]));
  };
  // the window's packed counts of dword i (fine 0..63, coarse 64..67):
  // lanes l-EX .. l+EX (halo lanes read neighbours' words that exist, and
  // their results are not stored)
  const int wl = lane < EX ? EX : (lane > 63 - EX ? 63 - EX : lane);
  const uint32_t* const hw = hist + wl - EX;
  auto kword = [&](int i) -> uint32_t {
    uint32_t s = 0;
#pragma unroll
    for (int dx = 0; dx <= 2 * EX; ++dx) s += hw[i * 64 + dx];
    return s;
  };
  // among 16 packed byte counts (4 dwords), the byte holding rank r and the
  // count below it
  // (each choice is a chain of selects on the monotone prefix tests: a
  // chain on the index d itself is turned into a private array indexed at
  // run time, i.e. a scratch store and load per output)
  auto select16 = [](const uint32_t (&w)[4], int r, int& below) -> int {
    const int s0 = (int)__builtin_amdgcn_sad_u8(w[0], 0u, 0u);
    const int s1 = s0 + (int)__builtin_amdgcn_sad_u8(w[1], 0u, 0u);
    const int s2 = s1 + (int)__builtin_amdgcn_sad_u8(w[2], 0u, 0u);
    const bool g0 = r >= s0, g1 = r >= s1, g2 = r >= s2;  // g2 => g1 => g0
    int base = 0, d = 0;
    uint32_t x = w[0];
    if (g0) { base = s0; x = w[1]; d = 1; }
    if (g1) { base = s1; x = w[2]; d = 2; }
    if (g2) { base = s2; x = w[3]; d = 3; }
    const int b0 = (int)(x & 255u), b1 = b0 + (int)((x >> 8) & 255u), b2 = b1 + (int)((x >> 16) & 255u);
    const int rr = r - base;
    const bool h0 = rr >= b0, h1 = rr >= b1, h2 = rr >= b2;
    int lo = 0, b = 0;
    if (h0) { lo = b0; b = 1; }
    if (h1) { lo = b1; b = 2; }
    if (h2) { lo = b2; b = 3; }
    below = base + lo;
    return 4 * d + b;
  };
  const int k = p.n / 2;
  CLK_GLOBAL TD* __restrict__ dst = gmem(static_cast<TD*>(p.dst));
  const int64_t plane = p.H * p.W;
  auto coarse_reads = [&](uint32_t (&cw)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) cw[q] = kword(64 + q);
  };
  for (int dz = -ez; dz <= ez; ++dz) {
    Raw v[NY];
    load(z0 + dz, v);
#pragma unroll
    for (int j = 0; j < NY; ++j) update(key_of(v[j]), true);
  }
  wave_lds_order();  // every lane's counter updates before the reads
  uint32_t cw[4];
  coarse_reads(cw);
  // The planes that step s (the window moving from s - 1 to s) removes and
  // adds, in buffer (s - z0 - 1) % 3: three steps are in flight, so their
  // loads are covered, and the buffers' roles are static (no copies).
  Raw pb[3][2][NY];
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    load(z0 + u - ez, pb[u][0]);
    load(z0 + u + 1 + ez, pb[u][1]);
  }
  // Step z: the coarse bin from the coarse reads issued in step z - 1, the
  // fine reads of z, then step z + 1's counter updates and coarse reads
  // (after the last output too, so the body has no branch), and only then
  // the fine bin of z.  A wave's LDS instructions run in issue order, so
  // the fine reads see the counts of step z, and their latency is covered
  // by z + 1's updates and reads.
  for (int zb = z0;; zb += 3) {
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      const int z = zb + u;
      if (z >= z1) return;  // wave-uniform
      int lt = 0;
      const int c = select16(cw, k, lt);
      uint32_t fw[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) fw[q] = kword(4 * c + q);
      wave_lds_order();  // step z's reads before these updates
#pragma unroll
      for (int j = 0; j < NY; ++j) {
        update(key_of(pb[u][0][j]), false);
        update(key_of(pb[u][1][j]), true);
      }
      // step z + 4's planes into the buffer just used
      load(z + 3 - ez, pb[u][0]);
      load(z + 4 + ez, pb[u][1]);
      wave_lds_order();  // counter updates before the reads
      coarse_reads(cw);
      int lt2 = 0;
      const int f = select16(fw, k - lt, lt2);
      CLK_GLOBAL TD* const drow = dst + (int64_t)z * plane + y * p.W;  // uniform
      if (out) drow[(uint32_t)xi] = static_cast<TD>(16 * c + f + static_cast<int>(Pix<TD>::lo));
    }
  }
}

template <class TS, class TD, int EX, int EY>
__global__ __launch_bounds__(64) void median_hist8_cols_kernel(const MedParams p) {
  median_hist8_cols_body<TS, TD, EX, EY>(p, blockIdx.x, blockIdx.y, blockIdx.z);
}

#define CLK_MEDIAN_KERNEL(NAME, THREADS)                                              \
  template <class TS, class TD>                                                      \
  __global__ __launch_bounds__(THREADS) void NAME##_kernel(const MedParams p) {      \
    NAME##_body<TS, TD>(p, blockIdx.x, blockIdx.y, blockIdx.z);                      \
  }
CLK_MEDIAN_KERNEL(median_generic, 256)
CLK_MEDIAN_KERNEL(median_27, 256)
#undef CLK_MEDIAN_KERNEL

template <class TS, class TD>
__global__ __launch_bounds__(64) void median_hist8_kernel(const MedParams p) {
  median_hist8_body<TS, TD>(p, blockIdx.x, blockIdx.y, blockIdx.z);
}

template <class TS, class TD, int EX, int EY>
__global__ __launch_bounds__(64) void median_hist8_fixed_kernel(const MedParams p) {
  median_hist8_fixed_body<TS, TD, EX, EY>(p, blockIdx.x, blockIdx.y, blockIdx.z);
}

template <class TS, class TD, int I>
__global__ __launch_bounds__(256) void median_regs_kernel(const MedParams p) {
  median_regs_body<TS, TD, I>(p, blockIdx.x, blockIdx.y, blockIdx.z);
}

template <class TS, class TD>
__global__ __launch_bounds__(256) void median_27_key_slide_kernel(const MedParams p) {
  median_27_key_slide_body<TS, TD>(p, blockIdx.x, blockIdx.y, blockIdx.z);
}

template <class TS, class TD>
__global__ __launch_bounds__(256) void median_27_pk8_slide_kernel(const MedParams p) {
  median_27_pk8_slide_body<TS, TD>(p, blockIdx.x, blockIdx.y, blockIdx.z);
}

template <class TS, class TD>
__global__ __launch_bounds__(256) void median_9_pk8_slide_kernel(const MedParams p) {
  median_9_pk8_slide_body<TS, TD>(p, blockIdx.x, blockIdx.y, blockIdx.z);
}


// MED_REGS + I: register tier I of the bisection median
enum MedKind : int { MED_PK8 = 0, MED_HIST8 = 1, MED_27 = 2, MED_GENERIC = 3, MED_PK9 = 4, MED_REGS = 5 };

// The smallest register tier whose box holds the requested one, when the
// requested box fills at least a quarter of it (else the generic body, whose
// steps re-read only the requested window, does less work); -1 if none.
inline int median_tier(const MedParams& p) {
  for (int i = 0; i < kMedTiers; ++i) {
    if (p.ex > kMedTier[i][0] || p.ey > kMedTier[i][1] || p.ez > kMedTier[i][2]) continue;
    const int ns = (2 * kMedTier[i][0] + 1) * (2 * kMedTier[i][1] + 1) * (2 * kMedTier[i][2] + 1);
    return 4 * p.n >= ns ? i : -1;
  }
  return -1;
}

// Kernel and grid for one median launch; sets p.gx, p.gy.
inline MedKind plan_median(MedParams& p, int td_size, dim3* grid, int* threads) {
  MedKind k;
  if ((td_size == 1 || td_size == 2) && p.ex == 1 && p.ey == 1 && (p.ez == 1 || p.ez == 0)) {
    // integer dst of 8 or 16 bits (float dst is 4 bytes); 3x3x3 or 3x3x1
    k = p.ez == 1 ? MED_PK8 : MED_PK9;
    *grid = dim3((uint32_t)((p.W + 511) / 512), (uint32_t)p.H, (uint32_t)p.D);
    *threads = 256;
  } else if (td_size == 1 && p.n <= 255 && hist8_enabled()) {
    k = MED_HIST8;
    *grid = dim3((uint32_t)((p.W + 63) / 64), (uint32_t)p.H, (uint32_t)((p.D + kHistZSeg - 1) / kHistZSeg));
    *threads = 64;
  } else {
    const int tier = td_size >= 2 ? median_tier(p) : -1;  // 8-bit dst: generic when the histogram is off
    if (p.ex == 1 && p.ey == 1 && p.ez == 1) k = MED_27;
    else k = tier >= 0 ? (MedKind)(MED_REGS + tier) : MED_GENERIC;
    *grid = dim3((uint32_t)((p.W + 255) / 256), (uint32_t)p.H, (uint32_t)p.D);
    *threads = 256;
  }
  p.gx = grid->x;
  p.gy = grid->y;
  return k;
}

}  // namespace clk
