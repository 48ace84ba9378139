// Float medians over windows that hold NaN or -0: a brick scan and a
// recompute of the output bricks whose windows reach a flagged brick, with
// the reference's insertion-sort result (see median_exact.h).
#include <float.h>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <string>

#include "median_exact.h"

namespace clk {

namespace {

constexpr int kBrick = 8;
constexpr int kNzMaxE = 31;  // largest ellipsoid half width of median_sphere_3d

struct NzParams {
  const float* src;
  float* dst;
  int64_t W, H, D;
  int ex, ey, ez;
  int n;       // neighbourhood size
  int sphere;  // 1: only offsets with tx + ty + tz <= 1 belong to it
  uint8_t* flags;
  int bw, bh, bd;  // bricks per axis
  float tx[2 * kNzMaxE + 1], ty[2 * kNzMaxE + 1], tz[2 * kNzMaxE + 1];
};

// One block per 8x8x8 brick (two voxels per thread): 1 when the brick holds
// a NaN or a -0, else 0.
__global__ __launch_bounds__(256) void nanzero_scan_kernel(const NzParams p) {
  const int t = threadIdx.x;
  const int64_t x = (int64_t)blockIdx.x * kBrick + (t & 7);
  const int64_t y = (int64_t)blockIdx.y * kBrick + ((t >> 3) & 7);
  int hit = 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t z = (int64_t)blockIdx.z * kBrick + (t >> 6) + 4 * h;
    if (x < p.W && y < p.H && z < p.D) {
      const float v = p.src[x + y * p.W + z * p.W * p.H];
      hit |= (v != v) || (__float_as_uint(v) == 0x80000000u);
    }
  }
  hit = __syncthreads_or(hit);
  if (t == 0) p.flags[((int64_t)blockIdx.z * p.bh + blockIdx.y) * p.bw + blockIdx.x] = hit ? 1 : 0;
}

__device__ __forceinline__ bool nz_member(const NzParams& p, int dx, int dy, int dz) {
  return !p.sphere || (p.tx[dx + p.ex] + p.ty[dy + p.ey] + p.tz[dz + p.ez] <= 1.0f);
}

__device__ __forceinline__ float nz_read(const NzParams& p, int64_t x, int64_t y, int64_t z) {
  x = x < 0 ? 0 : (x >= p.W ? p.W - 1 : x);
  y = y < 0 ? 0 : (y >= p.H ? p.H - 1 : y);
  z = z < 0 ? 0 : (z >= p.D ? p.D - 1 : z);
  return p.src[x + y * p.W + z * p.W * p.H];
}

// Gather-order walk over the members of the window (x outer, z inner): f
// gets the member index and the value; it returns false to stop the walk.
template <class F>
__device__ __forceinline__ void nz_walk(const NzParams& p, int64_t x, int64_t y, int64_t z, F&& f) {
  int idx = 0;
  for (int dx = -p.ex; dx <= p.ex; ++dx)
    for (int dy = -p.ey; dy <= p.ey; ++dy)
      for (int dz = -p.ez; dz <= p.ez; ++dz)
        if (nz_member(p, dx, dy, dz)) {
          if (!f(idx, nz_read(p, x + dx, y + dy, z + dz))) return;
          ++idx;
        }
}

// Element n/2 of the window sorted by the reference's insertion sort: the
// window in gather order split at its NaNs, each run sorted stably (ties,
// i.e. -0 and +0, keep their order), NaNs in place.  Returns false when the
// value the median kernel stored is already that element:
//  - no NaN and no -0 in the window: the same sort;
//  - no NaN: the sorted values agree up to the sign of zeros, so only a
//    zero at position n/2 needs the sign of the (n/2 - #negatives)-th zero
//    in gather order.
// With a NaN, the member of stable rank n/2 - s within the run [s, e) that
// holds position n/2 (a quadratic count, as the reference's sort).
__device__ bool nz_exact(const NzParams& p, int64_t x, int64_t y, int64_t z, float* out) {
  const int t = p.n / 2;
  int s = 0, e = p.n, nneg = 0, nzero = 0;
  bool nan_seen = false, negzero = false, nan_at_t = false;
  float vt = 0.f;
  nz_walk(p, x, y, z, [&](int idx, float v) {
    if (v != v) {
      nan_seen = true;
      if (idx == t) {  // the NaN at position t stays there
        nan_at_t = true;
        vt = v;
        return false;
      }
      if (idx > t) {
        e = idx;
        return false;
      }
      s = idx + 1;
    } else {
      negzero |= __float_as_uint(v) == 0x80000000u;
      nneg += v < 0.f ? 1 : 0;
      nzero += v == 0.f ? 1 : 0;
    }
    return true;
  });
  if (nan_at_t) {
    *out = vt;
    return true;
  }
  if (!nan_seen) {
    if (!negzero || t < nneg || t >= nneg + nzero) return false;
    int k = t - nneg;  // the k-th zero in gather order
    nz_walk(p, x, y, z, [&](int, float v) {
      if (v == 0.f && k-- == 0) {
        *out = v;
        return false;
      }
      return true;
    });
    return true;
  }
  const int r = t - s;
  bool found = false;
  nz_walk(p, x, y, z, [&](int i, float vi) {
    if (i >= e) return false;
    if (i < s) return true;
    int rank = 0;
    nz_walk(p, x, y, z, [&](int j, float vj) {
      if (j >= e) return false;
      if (j >= s) rank += (vj < vi || (vj == vi && j < i)) ? 1 : 0;
      return true;
    });
    if (rank == r) {
      *out = vi;
      found = true;
      return false;
    }
    return true;
  });
  return found;  // stable ranks within a run are a permutation: always found
}

// One block per output brick: nothing to do unless a brick the windows of
// this brick read (clamped reads stay inside the volume) is flagged.
__global__ __launch_bounds__(256) void nanzero_fix_kernel(const NzParams p) {
  const int64_t bx0 = (int64_t)blockIdx.x * kBrick, by0 = (int64_t)blockIdx.y * kBrick,
                bz0 = (int64_t)blockIdx.z * kBrick;
  auto lo = [](int64_t a, int e) { return a - e < 0 ? (int64_t)0 : a - e; };
  auto hi = [](int64_t a, int e, int64_t n) { return a + kBrick - 1 + e > n - 1 ? n - 1 : a + kBrick - 1 + e; };
  const int64_t x0 = lo(bx0, p.ex) / kBrick, x1 = hi(bx0, p.ex, p.W) / kBrick;
  const int64_t y0 = lo(by0, p.ey) / kBrick, y1 = hi(by0, p.ey, p.H) / kBrick;
  const int64_t z0 = lo(bz0, p.ez) / kBrick, z1 = hi(bz0, p.ez, p.D) / kBrick;
  bool any = false;
  for (int64_t cz = z0; cz <= z1 && !any; ++cz)
    for (int64_t cy = y0; cy <= y1 && !any; ++cy)
      for (int64_t cx = x0; cx <= x1 && !any; ++cx) any = p.flags[(cz * p.bh + cy) * p.bw + cx] != 0;
  if (!any) return;
  const int t = threadIdx.x;
  const int64_t x = bx0 + (t & 7), y = by0 + ((t >> 3) & 7);
#pragma unroll 1
  for (int h = 0; h < 2; ++h) {
    const int64_t z = bz0 + (t >> 6) + 4 * h;
    float v;
    if (x < p.W && y < p.H && z < p.D && nz_exact(p, x, y, z, &v)) p.dst[x + y * p.W + z * p.W * p.H] = v;
  }
}

// x^2 / a^2 per offset with a^2 = e^2 as a float, FLT_MIN for e = 0
// (median_sphere_3d_x.cl:43-61; x*x is an int converted to float).
void nz_axis_terms(int e, float* t) {
  float a2 = (float)(e * e);
  if (a2 == 0) a2 = FLT_MIN;
  for (int x = -e; x <= e; ++x) t[x + e] = (float)(x * x) / a2;
}

}  // namespace

int64_t nanzero_flag_bytes(const clk_image& img) {
  return ((img.width + kBrick - 1) / kBrick) * ((img.height + kBrick - 1) / kBrick) *
         ((img.depth + kBrick - 1) / kBrick);
}

int32_t launch_median_nanzero(hipStream_t st, const clk_image& dst, const clk_image& src, int ex, int ey, int ez,
                              bool sphere, void* flags, std::string* err) {
  if (src.dtype != CLK_FLOAT || dst.dtype != CLK_FLOAT) return 0;
  if (dst.width * dst.height * dst.depth == 0) return 0;
  if (ex < 0 || ey < 0 || ez < 0 || (sphere && (ex > kNzMaxE || ey > kNzMaxE || ez > kNzMaxE))) {
    *err = "median: half widths out of range for the NaN/-0 pass";
    return CLK_INVALID_VALUE;
  }
  NzParams p;
  memset(&p, 0, sizeof(p));
  p.src = static_cast<const float*>(src.ptr);
  p.dst = static_cast<float*>(dst.ptr);
  p.W = dst.width;
  p.H = dst.height;
  p.D = dst.depth;
  p.ex = ex;
  p.ey = ey;
  p.ez = ez;
  p.sphere = sphere ? 1 : 0;
  p.flags = static_cast<uint8_t*>(flags);
  const int64_t bw = (p.W + kBrick - 1) / kBrick, bh = (p.H + kBrick - 1) / kBrick,
                bd = (p.D + kBrick - 1) / kBrick;
  if (bw > 0x7fffffffLL || bh > 65535 || bd > 65535) {
    *err = "median: volume too large for the NaN/-0 pass grid";
    return CLK_INVALID_VALUE;
  }
  p.bw = (int)bw;
  p.bh = (int)bh;
  p.bd = (int)bd;
  int64_t n = 0;
  if (sphere) {
    nz_axis_terms(ex, p.tx);
    nz_axis_terms(ey, p.ty);
    nz_axis_terms(ez, p.tz);
    for (int x = -ex; x <= ex; ++x)
      for (int y = -ey; y <= ey; ++y)
        for (int z = -ez; z <= ez; ++z)
          if (p.tx[x + ex] + p.ty[y + ey] + p.tz[z + ez] <= 1.0f) ++n;
  } else {
    n = (int64_t)(2 * ex + 1) * (2 * ey + 1) * (2 * ez + 1);
  }
  if (n <= 0 || n > 0x7fffffff) {
    *err = "median: neighbourhood size out of range for the NaN/-0 pass";
    return CLK_INVALID_VALUE;
  }
  p.n = (int)n;
  const dim3 grid((uint32_t)bw, (uint32_t)bh, (uint32_t)bd);
  hipLaunchKernelGGL(nanzero_scan_kernel, grid, dim3(256), 0, st, p);
  hipLaunchKernelGGL(nanzero_fix_kernel, grid, dim3(256), 0, st, p);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    *err = std::string("kernel launch failed: ") + hipGetErrorString(e);
    return CLK_OUT_OF_RESOURCES;
  }
  return 0;
}

}  // namespace clk
