// Device-side helpers shared by the gfx950 kernels: pixel traits, the
// saturating float->pixel conversion (CONVERT_dst_PIXEL_TYPE), the widening
// read (READ_src_IMAGE(...).x cast to float) and a correctly rounded divide
// by a launch-uniform divisor.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clk {

enum Op : int { OP_GAUSS = 0, OP_MEAN = 1, OP_MIN = 2, OP_MAX = 3 };

template <class T> struct Pix;
template <> struct Pix<int8_t>   { static constexpr bool kInt = true;  static constexpr float lo = -128.f;   static constexpr float hi = 127.f; };
template <> struct Pix<uint8_t>  { static constexpr bool kInt = true;  static constexpr float lo = 0.f;      static constexpr float hi = 255.f; };
template <> struct Pix<int16_t>  { static constexpr bool kInt = true;  static constexpr float lo = -32768.f; static constexpr float hi = 32767.f; };
template <> struct Pix<uint16_t> { static constexpr bool kInt = true;  static constexpr float lo = 0.f;      static constexpr float hi = 65535.f; };
template <> struct Pix<float>    { static constexpr bool kInt = false; static constexpr float lo = 0.f;      static constexpr float hi = 0.f; };

// CONVERT_dst_PIXEL_TYPE (reference README.md:35-47): identity for float;
// for an integer type saturate to [min,max], truncate toward zero, NaN -> 0
// (OpenCL 1.2 convert_<T>_sat with its default round-toward-zero).
template <class T>
__device__ __forceinline__ T convert_sat(float f) {
  if constexpr (!Pix<T>::kInt) {
    return f;
  } else {
    f = (f != f) ? 0.f : f;
    f = fminf(fmaxf(f, Pix<T>::lo), Pix<T>::hi);
    return static_cast<T>(static_cast<int>(f));
  }
}

// Correctly rounded a / b for a launch-uniform divisor b >= 1 with
// r = RN(1/b) computed on the host: q = RN(a*r) is faithful, the residual
// a - b*q is exact under FMA, and RN(q + residual*r) is RN(a/b) (Markstein).
// Inputs whose residual could underflow, and non-finite values, take the
// IEEE divide.
__device__ __forceinline__ float div_uniform(float a, float b, float r) {
  const float q = a * r;
  const float e = __builtin_fmaf(-q, b, a);
  const float q1 = __builtin_fmaf(e, r, q);
  const float aa = __builtin_fabsf(a);
  if (__builtin_expect(!(aa >= 0x1p-96f && aa <= 0x1p+120f), 0)) {
    return a / b;
  }
  return q1;
}

// The same quotient without the range guard, for launches whose numerators
// are known to be 0 or within [2^-96, 2^120] in magnitude (sums of integer
// pixels with weights not below 2^-48).
__device__ __forceinline__ float div_uniform_fast(float a, float b, float r) {
  const float q = a * r;
  const float e = __builtin_fmaf(-q, b, a);
  return __builtin_fmaf(e, r, q);
}

// Image pixels live in global memory.  The batch engine reads its image
// pointers from a table in memory, where the compiler cannot tell them from
// LDS or scratch pointers and emits flat accesses, which are waited for at
// once (a flat access may complete out of order), so nothing overlaps a
// load's latency.  Every image access goes through a global-space pointer.
#define CLK_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const CLK_GLOBAL T* gmem(const T* p) { return (const CLK_GLOBAL T*)p; }
template <class T>
__device__ __forceinline__ CLK_GLOBAL T* gmem(T* p) { return (CLK_GLOBAL T*)p; }
// Back to a generic pointer, for the helpers below that take one (they
// access through gmem again).
template <class T>
__device__ __forceinline__ const T* gen(const CLK_GLOBAL T* p) { return (const T*)p; }
template <class T>
__device__ __forceinline__ T* gen(CLK_GLOBAL T* p) { return (T*)p; }

// Load COLS consecutive pixels of type T, widened to float.  `p` must be
// aligned to sizeof(T)*COLS when COLS > 1.
template <class T, int COLS> struct alignas(sizeof(T) * COLS) PixVec { T v[COLS]; };

// N pixels from / to global memory as one access (a vector of builtin type:
// a class object cannot be copied through a global-space reference); N that
// is not a power of two goes pixel by pixel.
template <class T, int N>
__device__ __forceinline__ PixVec<T, N> gload_vec(const T* p) {
  PixVec<T, N> v;
  if constexpr (N > 1 && (N & (N - 1)) == 0) {
    typedef T V __attribute__((ext_vector_type(N)));
    const V x = *reinterpret_cast<const CLK_GLOBAL V*>(gmem(p));
#pragma unroll
    for (int i = 0; i < N; ++i) v.v[i] = x[i];
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) v.v[i] = gmem(p)[i];
  }
  return v;
}

template <class T, int N>
__device__ __forceinline__ void gstore_vec(T* p, const PixVec<T, N>& v) {
  if constexpr (N > 1 && (N & (N - 1)) == 0) {
    typedef T V __attribute__((ext_vector_type(N)));
    V x;
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = v.v[i];
    *reinterpret_cast<CLK_GLOBAL V*>(gmem(p)) = x;
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) gmem(p)[i] = v.v[i];
  }
}

template <class T, int COLS>
__device__ __forceinline__ void load_widen(const T* __restrict__ p, float (&out)[COLS]) {
  if constexpr (COLS == 1) {
    out[0] = static_cast<float>(gmem(p)[0]);
  } else {
    const PixVec<T, COLS> v = gload_vec<T, COLS>(p);
#pragma unroll
    for (int i = 0; i < COLS; ++i) out[i] = static_cast<float>(v.v[i]);
  }
}

template <class T, int COLS>
__device__ __forceinline__ void store_vec(T* __restrict__ p, const T (&in)[COLS]) {
  if constexpr (COLS == 1) {
    gmem(p)[0] = in[0];
  } else {
    PixVec<T, COLS> v;
#pragma unroll
    for (int i = 0; i < COLS; ++i) v.v[i] = in[i];
    gstore_vec<T, COLS>(p, v);
  }
}

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Streaming (non-temporal) store of N pixels: outputs are not read again by
// the kernel, so they should not displace inputs from the caches.  `p` must
// be aligned to the vector size.
template <class T, int N>
__device__ __forceinline__ void store_nt(T* __restrict__ p, const T (&in)[N]) {
  constexpr int B = N * (int)sizeof(T);
  if constexpr (B % 4 != 0 || (B > 8 && B % 16 != 0)) {
    store_vec<T, N>(p, in);
  } else {
    PixVec<T, N> v;
#pragma unroll
    for (int i = 0; i < N; ++i) v.v[i] = in[i];
    uint32_t w[B / 4];
    __builtin_memcpy(w, &v, B);
    if constexpr (B == 4) {
      __builtin_nontemporal_store(w[0], reinterpret_cast<CLK_GLOBAL uint32_t*>(gmem(p)));
    } else if constexpr (B == 8) {
      const u32x2 t = {w[0], w[1]};
      __builtin_nontemporal_store(t, reinterpret_cast<CLK_GLOBAL u32x2*>(gmem(p)));
    } else {
#pragma unroll
      for (int k = 0; k < B / 16; ++k) {
        const u32x4 t = {w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]};
        __builtin_nontemporal_store(t, reinterpret_cast<CLK_GLOBAL u32x4*>(gmem(p)) + k);
      }
    }
  }
}

// Streaming load of N pixels (read once), same alignment rule.
template <class T, int N>
__device__ __forceinline__ PixVec<T, N> load_nt(const T* __restrict__ p) {
  constexpr int B = N * (int)sizeof(T);
  PixVec<T, N> v;
  if constexpr (B % 4 != 0 || (B > 8 && B % 16 != 0)) {
    v = gload_vec<T, N>(p);
  } else {
    uint32_t w[B / 4];
    if constexpr (B == 4) {
      w[0] = __builtin_nontemporal_load(reinterpret_cast<const CLK_GLOBAL uint32_t*>(gmem(p)));
    } else if constexpr (B == 8) {
      const u32x2 t = __builtin_nontemporal_load(reinterpret_cast<const CLK_GLOBAL u32x2*>(gmem(p)));
      w[0] = t.x; w[1] = t.y;
    } else {
#pragma unroll
      for (int k = 0; k < B / 16; ++k) {
        const u32x4 t = __builtin_nontemporal_load(reinterpret_cast<const CLK_GLOBAL u32x4*>(gmem(p)) + k);
        w[4 * k] = t.x; w[4 * k + 1] = t.y; w[4 * k + 2] = t.z; w[4 * k + 3] = t.w;
      }
    }
    __builtin_memcpy(&v, w, B);
  }
  return v;
}

// Streaming load of B bytes (8, 16 or 32; `p` aligned to min(B, 16)) as
// 32-bit words, without going through a byte array.
template <int B>
__device__ __forceinline__ void load_nt_words(const void* __restrict__ p, uint32_t (&w)[B / 4]) {
  static_assert(B == 8 || B % 16 == 0, "8 bytes or whole 16-byte vectors");
  if constexpr (B == 8) {
    const u32x2 t = __builtin_nontemporal_load(static_cast<const CLK_GLOBAL u32x2*>(gmem(p)));
    w[0] = t.x; w[1] = t.y;
  } else {
#pragma unroll
    for (int k = 0; k < B / 16; ++k) {
      const u32x4 t = __builtin_nontemporal_load(static_cast<const CLK_GLOBAL u32x4*>(gmem(p)) + k);
      w[4 * k] = t.x; w[4 * k + 1] = t.y; w[4 * k + 2] = t.z; w[4 * k + 3] = t.w;
    }
  }
}

// The same streaming loads and stores at any byte alignment: the GPU runs
// in unaligned access mode (the ROCm default for gfx9 targets, which the
// compiler also assumes for amdhsa), so one vector instruction moves a
// chunk of pixels wherever a row starts.  Used where rows of arbitrary
// length make chunk addresses unaligned.
typedef u32x4 u32x4_ua __attribute__((aligned(1)));
typedef u32x2 u32x2_ua __attribute__((aligned(1)));

// A buffer resource over the n bytes at p (n < 2^31), for raw buffer loads
// whose address is a uniform base (scalar registers) plus a 32-bit lane
// offset: no per-lane 64-bit address arithmetic.  Word 3 is the gfx9
// family's raw-buffer setting.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t byte_rsrc(const void* p, int64_t n) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)n, 0x00020000);
}

// Pixel at byte offset off of a buffer resource (offsets inside it only).
template <class T>
__device__ __forceinline__ T rsrc_px(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  if constexpr (sizeof(T) == 1) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0));
  else if constexpr (sizeof(T) == 2) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0));
  else return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

template <int B>
__device__ __forceinline__ void load_nt_words_ua(const void* __restrict__ p, uint32_t (&w)[B / 4]) {
  static_assert(B == 8 || B % 16 == 0, "8 bytes or whole 16-byte vectors");
  if constexpr (B == 8) {
    const u32x2 t = __builtin_nontemporal_load(static_cast<const CLK_GLOBAL u32x2_ua*>(gmem(p)));
    w[0] = t.x; w[1] = t.y;
  } else {
#pragma unroll
    for (int k = 0; k < B / 16; ++k) {
      const u32x4 t = __builtin_nontemporal_load(static_cast<const CLK_GLOBAL u32x4_ua*>(gmem(p)) + k);
      w[4 * k] = t.x; w[4 * k + 1] = t.y; w[4 * k + 2] = t.z; w[4 * k + 3] = t.w;
    }
  }
}

// Streaming store of N pixels (N * sizeof(T) = 8 or a multiple of 16 bytes)
// at any alignment.
template <class T, int N>
__device__ __forceinline__ void store_nt_ua(T* __restrict__ p, const T (&in)[N]) {
  constexpr int B = N * (int)sizeof(T);
  static_assert(B == 8 || B % 16 == 0, "8 bytes or whole 16-byte vectors");
  PixVec<T, N> v;
#pragma unroll
  for (int i = 0; i < N; ++i) v.v[i] = in[i];
  uint32_t w[B / 4];
  __builtin_memcpy(w, &v, B);
  if constexpr (B == 8) {
    const u32x2 t = {w[0], w[1]};
    __builtin_nontemporal_store(t, reinterpret_cast<CLK_GLOBAL u32x2_ua*>(gmem(p)));
  } else {
#pragma unroll
    for (int k = 0; k < B / 16; ++k) {
      const u32x4 t = {w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]};
      __builtin_nontemporal_store(t, reinterpret_cast<CLK_GLOBAL u32x4_ua*>(gmem(p)) + k);
    }
  }
}

__device__ __forceinline__ int64_t clamp64(int64_t v, int64_t lo, int64_t hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// Whole-wave shifts by one lane (DPP wave_shr:1 / wave_shl:1): lane l gets
// v of lane l-1 (from_left) or l+1 (from_right); the lane with no source
// (0, resp. 63) keeps old.
__device__ __forceinline__ uint32_t dpp_from_left(uint32_t old, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ uint32_t dpp_from_right(uint32_t old, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, 0x130, 0xf, 0xf, false);
}

}  // namespace clk
