// Ellipsoid-neighbourhood filters: minimum/maximum/mean/median_sphere_3d
// (kernels/minimum_sphere_3d_x.cl, maximum_sphere_3d_x.cl,
// mean_sphere_3d_x.cl, median_sphere_3d_x.cl).
//
// The neighbourhood is the box of half widths e = (N-1)/2 restricted to
// x^2/a^2 + y^2/b^2 + z^2/c^2 <= 1 (a^2 = e.x^2, or FLT_MIN when e.x = 0,
// minimum_sphere_3d_x.cl:16-28,37).  The per-axis terms x^2/a^2 are
// computed on the host with the same float divisions and passed in, so the
// kernel only adds three floats per offset in the reference's order.  One
// thread per voxel; offsets in the reference's loop order x, y, z.
#pragma once

#include "clk_device.h"
#include "clk_median.h"
#include "median_impl.h"

namespace clk {

constexpr int kSphereMaxE = 31;  // half width per axis held in the parameter block

enum SphereOp : int { SPH_MIN = 0, SPH_MAX = 1, SPH_MEAN = 2, SPH_MEDIAN = 3 };

struct SphereParams {
  const void* src;
  void* dst;
  int64_t W, H, D;
  int32_t ex, ey, ez;
  int32_t n;  // offsets inside the ellipsoid
  float tx[2 * kSphereMaxE + 1], ty[2 * kSphereMaxE + 1], tz[2 * kSphereMaxE + 1];
};

__device__ __forceinline__ bool sphere_in(const SphereParams& p, int x, int y, int z) {
  return p.tx[x + p.ex] + p.ty[y + p.ey] + p.tz[z + p.ez] <= 1.0f;  // :37, sum in x, y, z order
}

template <class TS, class TD, int OP>
__global__ __launch_bounds__(256) void sphere_kernel(const SphereParams p) {
  const int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t y = blockIdx.y, z = blockIdx.z;
  if (x >= p.W) return;
  const TS* __restrict__ src = static_cast<const TS*>(p.src);
  auto rd = [&](int dx, int dy, int dz) -> float {
    const int64_t xx = clamp64(x + dx, 0, p.W - 1), yy = clamp64(y + dy, 0, p.H - 1),
                  zz = clamp64(z + dz, 0, p.D - 1);
    return static_cast<float>(src[(zz * p.H + yy) * p.W + xx]);
  };
  TD* const out = static_cast<TD*>(p.dst) + (z * p.H + y) * p.W + x;
  if constexpr (OP == SPH_MEDIAN) {
    // rank count/2 of the dst-typed values (median_sphere_3d_x.cl:54-68), by
    // bisection over order-preserving keys as median_box_3d
    const int64_t k = p.n / 2;
    uint32_t lo = 0;
    uint32_t hi = (KeyBits<TD>::v == 32) ? 0xffffffffu : ((1u << KeyBits<TD>::v) - 1u);
    while (lo < hi) {
      const uint32_t mid = lo + ((hi - lo) >> 1);
      int64_t cnt = 0;
      for (int dx = -p.ex; dx <= p.ex; ++dx)
        for (int dy = -p.ey; dy <= p.ey; ++dy)
          for (int dz = -p.ez; dz <= p.ez; ++dz)
            if (sphere_in(p, dx, dy, dz))
              cnt += value_key<TD>(gather_value<TS, TD>(static_cast<TS>(rd(dx, dy, dz)), 0)) <= mid;
      if (cnt >= k + 1) hi = mid; else lo = mid + 1;
    }
    *out = static_cast<TD>(key_value<TD>(lo));
  } else if constexpr (OP == SPH_MEAN) {
    float sum = 0.f;  // mean_sphere_3d_x.cl:17, sequential in loop order
    int count = 0;
    for (int dx = -p.ex; dx <= p.ex; ++dx)
      for (int dy = -p.ey; dy <= p.ey; ++dy)
        for (int dz = -p.ez; dz <= p.ez; ++dz)
          if (sphere_in(p, dx, dy, dz)) {
            sum = sum + rd(dx, dy, dz);  // :44
            ++count;
          }
    *out = convert_sat<TD>(sum / static_cast<float>(count));  // :50, IEEE divide
  } else {
    float m = rd(0, 0, 0);  // minimum_sphere_3d_x.cl:15: the centre first
    for (int dx = -p.ex; dx <= p.ex; ++dx)
      for (int dy = -p.ey; dy <= p.ey; ++dy)
        for (int dz = -p.ez; dz <= p.ez; ++dz)
          if (sphere_in(p, dx, dy, dz)) {
            const float v = rd(dx, dy, dz);
            if (OP == SPH_MIN ? (v < m) : (v > m)) m = v;  // :44-46
          }
    *out = convert_sat<TD>(m);
  }
}

// LDS-tiled form for small neighbourhoods: a block filters 256 consecutive x
// of one (y, z).  It stages every source row (y+dy, z+dz) that holds a
// neighbourhood offset, x0-ex .. x0+255+ex clamped, once in LDS, then each
// thread walks the offset list (host-built, in the reference's x, y, z loop
// order) reading LDS.  Values are staged as floats (min/max/mean) or as
// order-preserving keys of the dst-typed value (median, by bisection).
constexpr int kSphereTileW = 256;
constexpr int kSphereTileMaxRows = 49;
constexpr int kSphereTileMaxOff = 400;

struct SphereTileParams {
  const void* src;
  void* dst;
  int64_t W, H, D;
  int32_t ex, nrows, noff, n;
  int8_t rdy[kSphereTileMaxRows], rdz[kSphereTileMaxRows];
  uint32_t off[kSphereTileMaxOff];  // (dx + 128) | row << 8
};

template <class TS, class TD, int OP>
__global__ __launch_bounds__(256) void sphere_tile_kernel(const SphereTileParams p) {
  extern __shared__ uint32_t tile[];
  const int t = threadIdx.x;
  const int64_t x0 = (int64_t)blockIdx.x * kSphereTileW;
  const int64_t y = blockIdx.y, z = blockIdx.z;
  const int rowlen = kSphereTileW + 2 * p.ex;
  const TS* __restrict__ src = static_cast<const TS*>(p.src);
  for (int r = 0; r < p.nrows; ++r) {
    const int64_t yy = clamp64(y + p.rdy[r], 0, p.H - 1), zz = clamp64(z + p.rdz[r], 0, p.D - 1);
    const TS* row = src + (zz * p.H + yy) * p.W;
    for (int e = t; e < rowlen; e += kSphereTileW) {
      const TS v = row[clamp64(x0 - p.ex + e, 0, p.W - 1)];
      uint32_t w;
      if constexpr (OP == SPH_MEDIAN) w = value_key<TD>(gather_value<TS, TD>(v, 0));
      else w = __float_as_uint(static_cast<float>(v));
      tile[r * rowlen + e] = w;
    }
  }
  __syncthreads();
  const int64_t x = x0 + t;
  if (x >= p.W) return;
  auto at = [&](uint32_t o) -> uint32_t {
    return tile[(int)(o >> 8) * rowlen + t + p.ex + ((int)(o & 0xff) - 128)];
  };
  TD* const out = static_cast<TD*>(p.dst) + (z * p.H + y) * p.W + x;
  if constexpr (OP == SPH_MEDIAN) {
    const int k = p.n / 2;
    uint32_t lo = 0;
    uint32_t hi = (KeyBits<TD>::v == 32) ? 0xffffffffu : ((1u << KeyBits<TD>::v) - 1u);
    while (lo < hi) {
      const uint32_t mid = lo + ((hi - lo) >> 1);
      int cnt = 0;
      for (int i = 0; i < p.noff; ++i) cnt += at(p.off[i]) <= mid;
      if (cnt >= k + 1) hi = mid; else lo = mid + 1;
    }
    *out = static_cast<TD>(key_value<TD>(lo));
  } else if constexpr (OP == SPH_MEAN) {
    float sum = 0.f;
    for (int i = 0; i < p.noff; ++i) sum = sum + __uint_as_float(at(p.off[i]));
    *out = convert_sat<TD>(sum / static_cast<float>(p.noff));
  } else {
    // the centre is an offset of every neighbourhood that is not empty; an
    // empty one (negative N) never takes this kernel
    float m = static_cast<float>(src[(z * p.H + y) * p.W + x]);
    for (int i = 0; i < p.noff; ++i) {
      const float v = __uint_as_float(at(p.off[i]));
      if (OP == SPH_MIN ? (v < m) : (v > m)) m = v;
    }
    *out = convert_sat<TD>(m);
  }
}

// Sliding form for neighbourhoods up to 9 wide in x and 7 in y: a block
// filters 256 consecutive x of TY rows (y0 .. y0+TY-1) for a run of zs
// planes.  LDS holds a ring of 2ez+2 source planes, each the rows
// y0-ey .. y0+TY-1+ey by columns x0-ex .. x0+255+ex (clamped into the
// volume) at a fixed row pitch of kSlideRL words.  Plane zz sits in slot
// (zz - z0 + ez) mod (2ez+2).  A z step reads planes z-ez .. z+ez and loads
// only plane z+ez+1: its loads are issued before the step's arithmetic and
// written to LDS after it, into the one slot the step does not read (the
// plane z-ez-1 it replaces); a block barrier then closes the step.  The
// offset list (host-built, the reference's x, y, z loop order) gives each
// offset one wave-uniform LDS base, which the TY outputs of a thread read at
// immediate offsets.  Values are staged as floats (min/max/mean) or as
// order-preserving keys of the dst-typed value (median, by bisection).
// Launches a sliding form when the neighbourhood fits one (and the
// sphere_direct knob is 0): the z-sliding byte histogram for medians with an
// 8-bit dst, else the z-sliding LDS ring; false otherwise.  Host code,
// sphere_slide.hip.
struct SphereParams;
bool launch_sphere_slide(int op, hipStream_t st, const SphereParams& p, int src_t, int dst_t);

constexpr int kSlideRL = 264;       // row pitch in words: 256 + 2*ex for ex <= 4
constexpr int kSlideMaxEY = 3;
// LDS per block: up to 80 KB (two blocks per CU) buys 8 output rows per
// block for 5x5x5-sized neighbourhoods, which beats 4 rows at three blocks
// per CU while a step waits on its plane's loads; the median network, whose
// registers already limit it to three waves per SIMD, keeps 64 KB
constexpr int kSlideMaxLDS = 80 * 1024;
constexpr int kSlideMaxLDSNet = 64 * 1024;

struct SphereSlideParams {
  const void* src;
  void* dst;
  int64_t W, H, D;
  int32_t ex, ey, ez, noff;
  int32_t zs;  // output planes per block
  int32_t reserved;
  // per offset: (dy + ey) * kSlideRL + dx + ex (its word within a ring slot,
  // output row 0) | (dz + ez) << 16; padded with the centre to a multiple of 8
  uint32_t off[kSphereTileMaxOff];
};
static_assert(kSphereTileMaxOff % 8 == 0, "offset chunks of 8");

// PK = 1: values travel as 16-bit keys and the ring word of row r holds
// rows r and r+1 (low and high half), so one LDS read serves a pair of
// output rows and v_pk_{min,max,add}_u16 handle both: min/max of 8/16-bit
// integer sources (keys v - lo keep the order), the mean of byte sources
// over at most 257 values (the key sum fits 16 bits; the reference's float
// sum of integers is exact, so the integer sum gives its result), and the
// median network.  NET > 0 (median, exactly NET offsets): the keys of each
// output row (of each pair of rows for a dst of 8 or 16 bits) are gathered
// once and the pruned Batcher network of the box median runs on them,
// instead of bisection rounds over LDS.
template <class TS, class TD, int OP, int TY, int NET = 0, int PK = 0>
__global__ __launch_bounds__(256) void sphere_slide_kernel(const SphereSlideParams p) {
  static_assert(NET == 0 || OP == SPH_MEDIAN, "the network is the median's");
  static_assert(PK == 0 || TY % 2 == 0, "rows in pairs");
  constexpr int NMAX = TY + 2 * kSlideMaxEY;
  constexpr int RL = kSlideRL;
  extern __shared__ uint32_t ring[];
  const int t = threadIdx.x;
  const int64_t x0 = (int64_t)blockIdx.x * 256;
  const int64_t y0 = (int64_t)blockIdx.y * TY;
  const int64_t z0 = (int64_t)blockIdx.z * p.zs;
  const int64_t z1 = (z0 + p.zs < p.D) ? z0 + p.zs : p.D;
  const int ex = p.ex, ey = p.ey, ez = p.ez;
  const int nrow = TY + 2 * ey;
  const int S = 2 * ez + 2;
  const int slot_words = nrow * RL;
  const TS* __restrict__ src = static_cast<const TS*>(p.src);
  // staging columns of this thread: c = t, and c = 256 + t for t < 2ex
  const int64_t xa = clamp64(x0 - ex + t, 0, p.W - 1);
  const bool two = t < 2 * ex;
  const int64_t xb = clamp64(x0 - ex + 256 + t, 0, p.W - 1);
  auto word = [&](TS v) -> uint32_t {
    if constexpr (OP == SPH_MEDIAN) return value_key<TD>(gather_value<TS, TD>(v, 0));
    else if constexpr (PK) return (uint32_t)((int)v - (int)Pix<TS>::lo);
    else return __float_as_uint(static_cast<float>(v));
  };
  // raw values of one plane's rows, converted when written to LDS so that
  // the loads stay in flight across a step's arithmetic
  TS va[NMAX], vb[NMAX];
  auto fetch = [&](int64_t zz) {
    const TS* pl = src + clamp64(zz, 0, p.D - 1) * p.H * p.W;
#pragma unroll
    for (int r = 0; r < NMAX; ++r) {
      if (r < nrow) {
        const TS* row = pl + clamp64(y0 - ey + r, 0, p.H - 1) * p.W;
        va[r] = row[xa];
        vb[r] = two ? row[xb] : TS(0);
      }
    }
  };
  auto put = [&](int slot) {
    uint32_t* s = ring + slot * slot_words + t;
#pragma unroll
    for (int r = 0; r < NMAX; ++r) {
      if (r < nrow) {
        if constexpr (PK) {
          const int r1 = r + 1 < NMAX ? r + 1 : r;  // row r + 1 (the last row's high half is never read)
          const bool has1 = r + 1 < nrow;
          s[r * RL] = word(va[r]) | (has1 ? word(va[r1]) << 16 : 0u);
          if (two) s[r * RL + 256] = word(vb[r]) | (has1 ? word(vb[r1]) << 16 : 0u);
        } else {
          s[r * RL] = word(va[r]);
          if (two) s[r * RL + 256] = word(vb[r]);
        }
      }
    }
  };
  auto pair_at = [](const uint32_t* b, int jp) -> us2 {
    const uint32_t w = b[(2 * jp) * RL];
    us2 u;
    __builtin_memcpy(&u, &w, 4);
    return u;
  };
  for (int i = 0; i <= 2 * ez; ++i) {
    fetch(z0 - ez + i);
    put(i);
  }
  __syncthreads();
  const int64_t x = x0 + t;
  const uint32_t* const lane = ring + t;  // this thread's staging column
  TD* const dst = static_cast<TD*>(p.dst);
  int cs = 0;          // slot of plane z - ez
  int ws = 2 * ez + 1; // slot the next plane goes to
  for (int64_t z = z0; z < z1; ++z) {
    const bool more = z + 1 < z1;
    if (more) fetch(z + ez + 1);
    // LDS word of a window offset for output row 0: its ring slot follows
    // from dz and the slot of plane z - ez
    auto base = [&](uint32_t o) -> const uint32_t* {
      int sl = cs + (int)(o >> 16);
      if (sl >= S) sl -= S;
      return lane + sl * slot_words + (int)(o & 0xffffu);
    };
    auto store = [&](int j, TD v) {
      if (x < p.W && y0 + j < p.H) dst[(z * p.H + y0 + j) * p.W + x] = v;
    };
    // f(LDS word of row 0) for every offset in list order, eight offsets
    // per scalar load so that a chunk's LDS reads are issued together.  The
    // host pads the list to a multiple of 8 with the centre offset: min/max
    // run the padded chunks whole (a second look at the centre changes
    // nothing), mean/median skip the padding.
    auto walk = [&](auto&& f) {
      constexpr bool kPadOk = OP == SPH_MIN || OP == SPH_MAX;
      const int n8 = kPadOk ? (p.noff + 7) & ~7 : p.noff & ~7;
      for (int i = 0; i < n8; i += 8) {
        uint32_t o[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) o[u] = p.off[i + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) f(base(o[u]));
      }
      if (!kPadOk && n8 < p.noff) {
        uint32_t o[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) o[u] = p.off[n8 + u];
        const int rem = p.noff - n8;
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (u < rem) f(base(o[u]));
      }
    };
    if constexpr (OP == SPH_MEDIAN && NET > 0 && PK) {
      // rank NET/2 of the keys (median_sphere_3d_x.cl:54-68); keys of an
      // 8- or 16-bit dst fit 16 bits, rows j and j+1 share a register
      static_assert(TY % 2 == 0 && (NET & 1) && KeyBits<TD>::v <= 16, "row pairs, odd count, 16-bit keys");
      // one pair of rows at a time keeps a single NET-entry array live
#pragma unroll
      for (int jp = 0; jp < TY / 2; ++jp) {
        us2 a[NET];
#pragma unroll
        for (int i = 0; i < NET; ++i) a[i] = pair_at(base(p.off[i]), jp);
        const us2 m = batcher_median<NET>(a);
        store(2 * jp, static_cast<TD>(key_value<TD>(m.x)));
        store(2 * jp + 1, static_cast<TD>(key_value<TD>(m.y)));
      }
    } else if constexpr (OP == SPH_MEDIAN && NET > 0) {
      // float dst: 32-bit keys, one output row at a time through the same
      // pruned network (the keys are the bisection's, so is the result)
#pragma unroll
      for (int j = 0; j < TY; ++j) {
        uint32_t a[NET];
#pragma unroll
        for (int i = 0; i < NET; ++i) a[i] = base(p.off[i])[j * RL];
        store(j, static_cast<TD>(key_value<TD>(batcher_median<NET, uint32_t>(a, 0xffffffffu))));
      }
    } else if constexpr (OP == SPH_MEDIAN) {
      // rank noff/2 of the keys (median_sphere_3d_x.cl:54-68) by bisection:
      // each round halves [lo, hi], so KeyBits rounds leave lo == hi
      const int k = p.noff / 2;
      uint32_t lo[TY], hi[TY];
#pragma unroll
      for (int j = 0; j < TY; ++j) {
        lo[j] = 0;
        hi[j] = (KeyBits<TD>::v == 32) ? 0xffffffffu : ((1u << KeyBits<TD>::v) - 1u);
      }
      for (int round = 0; round < KeyBits<TD>::v; ++round) {
        uint32_t mid[TY];
        int cnt[TY];
#pragma unroll
        for (int j = 0; j < TY; ++j) {
          mid[j] = lo[j] + ((hi[j] - lo[j]) >> 1);
          cnt[j] = 0;
        }
        walk([&](const uint32_t* b) {
#pragma unroll
          for (int j = 0; j < TY; ++j) cnt[j] += b[j * RL] <= mid[j];
        });
#pragma unroll
        for (int j = 0; j < TY; ++j) {
          if (cnt[j] >= k + 1) hi[j] = mid[j];
          else lo[j] = mid[j] + 1;
        }
      }
#pragma unroll
      for (int j = 0; j < TY; ++j) store(j, static_cast<TD>(key_value<TD>(lo[j])));
    } else if constexpr (OP == SPH_MEAN && PK) {
      us2 sum[TY / 2];
#pragma unroll
      for (int jp = 0; jp < TY / 2; ++jp) sum[jp] = us2{0, 0};
      walk([&](const uint32_t* b) {
#pragma unroll
        for (int jp = 0; jp < TY / 2; ++jp) sum[jp] += pair_at(b, jp);
      });
      const int bias = p.noff * (int)Pix<TS>::lo;
#pragma unroll
      for (int jp = 0; jp < TY / 2; ++jp) {
        store(2 * jp, convert_sat<TD>(static_cast<float>((int)sum[jp].x + bias) / static_cast<float>(p.noff)));
        store(2 * jp + 1, convert_sat<TD>(static_cast<float>((int)sum[jp].y + bias) / static_cast<float>(p.noff)));
      }
    } else if constexpr (PK) {
      // min / max of the keys, the centre first
      us2 m[TY / 2];
      const uint32_t* c = base((uint32_t)(ey * RL + ex) | ((uint32_t)ez << 16));
#pragma unroll
      for (int jp = 0; jp < TY / 2; ++jp) m[jp] = pair_at(c, jp);
      walk([&](const uint32_t* b) {
#pragma unroll
        for (int jp = 0; jp < TY / 2; ++jp) m[jp] = OP == SPH_MIN ? vmin(m[jp], pair_at(b, jp)) : vmax(m[jp], pair_at(b, jp));
      });
      constexpr int lo = (int)Pix<TS>::lo;
#pragma unroll
      for (int jp = 0; jp < TY / 2; ++jp) {
        store(2 * jp, convert_sat<TD>(static_cast<float>((int)m[jp].x + lo)));
        store(2 * jp + 1, convert_sat<TD>(static_cast<float>((int)m[jp].y + lo)));
      }
    } else if constexpr (OP == SPH_MEAN) {
      float sum[TY];
#pragma unroll
      for (int j = 0; j < TY; ++j) sum[j] = 0.f;  // mean_sphere_3d_x.cl:17
      walk([&](const uint32_t* b) {
#pragma unroll
        for (int j = 0; j < TY; ++j) sum[j] = sum[j] + __uint_as_float(b[j * RL]);  // :44, loop order
      });
#pragma unroll
      for (int j = 0; j < TY; ++j) store(j, convert_sat<TD>(sum[j] / static_cast<float>(p.noff)));
    } else {
      // the centre first (minimum_sphere_3d_x.cl:15), then every offset
      float m[TY];
      const uint32_t* c = base((uint32_t)(ey * RL + ex) | ((uint32_t)ez << 16));
#pragma unroll
      for (int j = 0; j < TY; ++j) m[j] = __uint_as_float(c[j * RL]);
      walk([&](const uint32_t* b) {
#pragma unroll
        for (int j = 0; j < TY; ++j) {
          const float v = __uint_as_float(b[j * RL]);
          if constexpr (Pix<TS>::kInt) {
            // integer values: no NaN or -0, so the IEEE min/max is the same
            m[j] = OP == SPH_MIN ? __builtin_fminf(v, m[j]) : __builtin_fmaxf(v, m[j]);
          } else if (OP == SPH_MIN ? (v < m[j]) : (v > m[j])) {
            m[j] = v;  // :44-46
          }
        }
      });
#pragma unroll
      for (int j = 0; j < TY; ++j) store(j, convert_sat<TD>(m[j]));
    }
    if (more) {
      put(ws);
      ws = ws + 1 == S ? 0 : ws + 1;
    }
    cs = cs + 1 == S ? 0 : cs + 1;
    __syncthreads();
  }
}

// median_sphere_3d with an 8-bit dst and at most 255 values: the per-lane
// byte histogram of the box median slid along z.  The ellipsoid is a set of
// columns (dx, dy), each holding the offsets dz in [-h, h] (x^2/a^2 + y^2/b^2
// + z^2/c^2 grows with |z| and float addition is monotone, so the set is
// contiguous and symmetric); stepping from z to z+1 removes plane z-h and
// adds plane z+1+h of every column.  A lane keeps the clamped in-plane
// offsets of its columns in registers and loads each step's values one step
// ahead.  The median is the rank n/2 of the dst-typed values, as the
// reference's sort (median_sphere_3d_x.cl:54-68).
constexpr int kSphereHistMaxCols = 32;

struct SphereHistParams {
  const void* src;
  void* dst;
  int64_t W, H, D;
  int32_t ncol, n;
  uint32_t col[kSphereHistMaxCols];  // (dx + 128) | (dy + 128) << 8 | h << 16
};

template <class TS, class TD, int NC>
__global__ __launch_bounds__(64) void sphere_median_hist8_kernel(const SphereHistParams p) {
  using Raw = typename std::conditional<std::is_integral<TS>::value, int32_t, TS>::type;
  __shared__ uint32_t hist[64 * 64];
  const int lane = threadIdx.x;
  const int64_t x = (int64_t)blockIdx.x * 64 + lane;
  const int64_t y = blockIdx.y;
  const int64_t z0 = (int64_t)blockIdx.z * kHistZSeg;
  const int64_t z1 = (z0 + kHistZSeg < p.D) ? z0 + kHistZSeg : p.D;
  if (x >= p.W) return;  // lanes never share data: no barrier follows
  const Hist8Lane hl{hist + lane};
  hl.clear();
  const TS* __restrict__ src = static_cast<const TS*>(p.src);
  const int64_t PS = p.H * p.W;
  int32_t off[NC];  // in-plane offset of column c at this lane (H*W < 2^31)
  int hh[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    off[c] = 0;
    hh[c] = 0;
    if (c < p.ncol) {
      const uint32_t o = p.col[c];
      const int64_t xx = clamp64(x + (int)(o & 255u) - 128, 0, p.W - 1);
      const int64_t yy = clamp64(y + (int)((o >> 8) & 255u) - 128, 0, p.H - 1);
      off[c] = (int32_t)(yy * p.W + xx);
      hh[c] = (int)(o >> 16);
    }
  }
  auto at = [&](int c, int64_t zz) -> Raw { return static_cast<Raw>(src[clamp64(zz, 0, p.D - 1) * PS + off[c]]); };
  auto key_of = [&](Raw r) -> int {
    if constexpr (std::is_same<TS, TD>::value && std::is_integral<TS>::value)
      return (r - static_cast<int>(Pix<TD>::lo)) << 6;
    else return hist8_bin<TS, TD>(static_cast<TS>(r), 0) << 6;
  };
  const int k = p.n / 2;
#pragma unroll
  for (int c = 0; c < NC; ++c)
    if (c < p.ncol)
      for (int dz = -hh[c]; dz <= hh[c]; ++dz) hl.add_key(key_of(at(c, z0 + dz)), true);
  int med = 0, lt = 0;
  hl.walk_from_zero(med, lt, k);
  TD* __restrict__ dst = static_cast<TD*>(p.dst);
  dst[z0 * PS + y * p.W + x] = static_cast<TD>(med + static_cast<int>(Pix<TD>::lo));
  struct Step {
    Raw out[NC], in[NC];
  };
  // step z removes plane z-1-h and adds plane z+h of each column; loads
  // past the volume clamp, so the prefetch after the last step stays in
  // bounds
  auto load = [&](int64_t z, Step& s) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (c < p.ncol) {
        s.out[c] = at(c, z - 1 - hh[c]);
        s.in[c] = at(c, z + hh[c]);
      }
    }
  };
  auto step = [&](int64_t z, const Step& cur, Step& nxt) {
    load(z + 1, nxt);
    const int mkey = med << 6;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (c < p.ncol) {
        const int a = key_of(cur.out[c]), b = key_of(cur.in[c]);
        hl.add_key(a, false);
        hl.add_key(b, true);
        lt += (b < mkey) - (a < mkey);
      }
    }
    hl.walk(med, lt, k);
    dst[z * PS + y * p.W + x] = static_cast<TD>(med + static_cast<int>(Pix<TD>::lo));
  };
  Step sa, sb;
  load(z0 + 1, sa);
  for (int64_t z = z0 + 1; z < z1; z += 2) {
    step(z, sa, sb);
    if (z + 1 < z1) step(z + 1, sb, sa);
  }
}

}  // namespace clk
