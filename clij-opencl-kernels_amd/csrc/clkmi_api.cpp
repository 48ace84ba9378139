// C-ABI of libclkmi: argument validation, error reporting, by-name dispatch,
// the 3-pass host chain, and device-memory/stream helpers.  Kernel launches
// live in sep_launch.hip and median_kernels.hip.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "../../include/clkmi.h"
#include "clk_device.h"
#include "clk_internal.h"
#include "median_exact.h"

namespace clk {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_pool_mu;
hipMemPool_t g_pools[kMaxDevices] = {};
bool g_no_pool[kMaxDevices] = {};

// The library's pool on device `dev` (created on first use), or nullptr
// when the device cannot create one (the default pool then serves it).
hipMemPool_t device_pool(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> g(g_pool_mu);
  if (!g_pools[dev] && !g_no_pool[dev]) {
    hipMemPoolProps props;
    memset(&props, 0, sizeof(props));
    props.allocType = hipMemAllocationTypePinned;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = dev;
    hipMemPool_t np = nullptr;
    if (hipMemPoolCreate(&np, &props) == hipSuccess) {
      uint64_t keep = UINT64_MAX;
      (void)hipMemPoolSetAttribute(np, hipMemPoolAttrReleaseThreshold, &keep);
      g_pools[dev] = np;
    } else {
      (void)hipGetLastError();
      g_no_pool[dev] = true;
    }
  }
  return g_pools[dev];
}

// Kept blocks of one (device, stream): a lock held by the call that uses
// them, and one block per slot.
struct StreamBlocks {
  std::recursive_mutex mu;
  void* ptr[clk::kScratchSlots] = {nullptr, nullptr, nullptr};
  size_t bytes[clk::kScratchSlots] = {0, 0, 0};
  // side streams (CallScratch::aux): created on first use
  bool aux_tried = false, aux_ok = false;
  hipStream_t aux[clk::kAuxStreams] = {};
  hipEvent_t join[clk::kAuxStreams] = {};
  hipEvent_t fork = nullptr;
};

void destroy_aux(StreamBlocks& b) {
  for (int i = 0; i < clk::kAuxStreams; ++i) {
    if (b.aux[i]) (void)hipStreamDestroy(b.aux[i]);
    if (b.join[i]) (void)hipEventDestroy(b.join[i]);
    b.aux[i] = nullptr;
    b.join[i] = nullptr;
  }
  if (b.fork) (void)hipEventDestroy(b.fork);
  b.fork = nullptr;
  b.aux_tried = b.aux_ok = false;
}
std::mutex g_blocks_mu;  // guards the map only (never held across HIP calls)
std::map<std::pair<int, hipStream_t>, std::unique_ptr<StreamBlocks>> g_blocks;

StreamBlocks* blocks_of(int dev, hipStream_t st, bool create) {
  std::lock_guard<std::mutex> g(g_blocks_mu);
  auto it = g_blocks.find({dev, st});
  if (it != g_blocks.end()) return it->second.get();
  if (!create) return nullptr;
  auto& slot = g_blocks[{dev, st}];
  slot.reset(new StreamBlocks);
  return slot.get();
}
}  // namespace

hipError_t temp_alloc(void** p, size_t bytes, hipStream_t st) {
  int dev = 0;
  const hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  hipMemPool_t pool = device_pool(dev);
  return pool ? hipMallocFromPoolAsync(p, bytes, pool, st) : hipMallocAsync(p, bytes, st);
}

CallScratch::CallScratch(hipStream_t st) : st_(st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  const hipError_t e = hipStreamIsCapturing(st, &cs);
  if (e != hipSuccess) (void)hipGetLastError();
  int dev = 0;
  const hipError_t ed = hipGetDevice(&dev);
  transient_ = st == hipStreamPerThread || e != hipSuccess || cs != hipStreamCaptureStatusNone ||
               ed != hipSuccess;
  if (!transient_) {
    StreamBlocks* b = blocks_of(dev, st, true);
    lock_ = std::unique_lock<std::recursive_mutex>(b->mu);
    blocks_ = b;
  }
}

CallScratch::~CallScratch() {
  // transient blocks go back in stream order, after the launches that use them
  for (void*& q : owned_)
    if (q) {
      (void)hipFreeAsync(q, st_);
      q = nullptr;
    }
}

hipError_t CallScratch::get(int slot, size_t bytes, void** p) {
  if (slot < 0 || slot >= kScratchSlots) return hipErrorInvalidValue;
  if (transient_) {
    if (owned_[slot]) {
      (void)hipFreeAsync(owned_[slot], st_);
      owned_[slot] = nullptr;
    }
    // under capture the allocation becomes a node of the graph; the
    // default pool is the one capture supports
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    const bool cap = hipStreamIsCapturing(st_, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
    const hipError_t e = cap ? hipMallocAsync(p, bytes, st_) : temp_alloc(p, bytes, st_);
    if (e == hipSuccess) owned_[slot] = *p;
    return e;
  }
  StreamBlocks& b = *static_cast<StreamBlocks*>(blocks_);
  if (b.bytes[slot] < bytes) {
    if (b.ptr[slot]) (void)hipFreeAsync(b.ptr[slot], st_);
    b.ptr[slot] = nullptr;
    b.bytes[slot] = 0;
    void* q = nullptr;
    const hipError_t ea = temp_alloc(&q, bytes, st_);
    if (ea != hipSuccess) return ea;
    b.ptr[slot] = q;
    b.bytes[slot] = bytes;
  }
  *p = b.ptr[slot];
  return hipSuccess;
}

int CallScratch::aux(hipStream_t** streams, hipEvent_t** joins, hipEvent_t* fork) {
  if (transient_ || !blocks_) return 0;
  StreamBlocks& b = *static_cast<StreamBlocks*>(blocks_);
  if (!b.aux_tried) {
    b.aux_tried = true;
    bool good = hipEventCreateWithFlags(&b.fork, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; good && i < kAuxStreams; ++i)
      good = hipStreamCreateWithFlags(&b.aux[i], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&b.join[i], hipEventDisableTiming) == hipSuccess;
    if (!good) {
      (void)hipGetLastError();
      destroy_aux(b);
      b.aux_tried = true;
      return 0;
    }
    b.aux_ok = true;
  }
  if (!b.aux_ok) return 0;
  *streams = b.aux;
  *joins = b.join;
  *fork = b.fork;
  return kAuxStreams;
}

int64_t kept_bytes(hipStream_t st) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  StreamBlocks* b = blocks_of(dev, st, false);
  if (!b) return 0;
  std::lock_guard<std::recursive_mutex> g(b->mu);
  int64_t n = 0;
  for (size_t v : b->bytes) n += (int64_t)v;
  return n;
}

hipError_t release_kept(hipStream_t st) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  StreamBlocks* b = blocks_of(dev, st, false);
  if (b) {
    std::lock_guard<std::recursive_mutex> g(b->mu);
    bool any = false;
    for (int s = 0; s < kScratchSlots; ++s)
      if (b->ptr[s]) {
        (void)hipFreeAsync(b->ptr[s], st);
        b->ptr[s] = nullptr;
        b->bytes[s] = 0;
        any = true;
      }
    if (b->aux_tried) {
      // the side streams' work is joined into st; wait for it, then drop them
      any = true;
    }
    if (any) e = hipStreamSynchronize(st);
    if (b->aux_tried) destroy_aux(*b);
  }
  hipMemPool_t pool = device_pool(dev);
  if (pool) (void)hipMemPoolTrimTo(pool, 0);  // gives back what no block holds
  return e;
}

int64_t pool_reserved_bytes() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  hipMemPool_t pool = device_pool(dev);
  if (!pool) return 0;
  uint64_t v = 0;
  if (hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemCurrent, &v) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return (int64_t)v;
}

}  // namespace clk

namespace {

thread_local std::string g_last_error;

int32_t fail(int32_t code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int32_t ok() { return CLK_SUCCESS; }

int64_t voxels(const clk_image& im) { return im.width * im.height * im.depth; }

int32_t check_image(const clk_image* im, const char* what) {
  if (!im) return fail(CLK_INVALID_MEM_OBJECT, std::string(what) + ": image is NULL");
  if (clk_dtype_size(im->dtype) == 0)
    return fail(CLK_INVALID_VALUE, std::string(what) + ": invalid pixel type " + std::to_string(im->dtype));
  if (im->width < 0 || im->height < 0 || im->depth < 0)
    return fail(CLK_INVALID_VALUE, std::string(what) + ": negative size");
  if (im->reserved != 0) return fail(CLK_INVALID_VALUE, std::string(what) + ": reserved field must be 0");
  if (!im->ptr && voxels(*im) > 0) return fail(CLK_INVALID_MEM_OBJECT, std::string(what) + ": NULL device pointer");
  return CLK_SUCCESS;
}

// Both images valid, same size, and their byte ranges disjoint (the kernels
// read neighbours of src, so they cannot run in place).
int32_t check_pair(const char* kernel, const clk_image* dst, const clk_image* src) {
  int32_t rc = check_image(dst, "dst");
  if (rc) return rc;
  rc = check_image(src, "src");
  if (rc) return rc;
  if (dst->width != src->width || dst->height != src->height || dst->depth != src->depth)
    return fail(CLK_INVALID_VALUE, std::string(kernel) + ": src and dst sizes differ");
  const uintptr_t s0 = (uintptr_t)src->ptr, d0 = (uintptr_t)dst->ptr;
  const uintptr_t s1 = s0 + (uintptr_t)(voxels(*src) * clk_dtype_size(src->dtype));
  const uintptr_t d1 = d0 + (uintptr_t)(voxels(*dst) * clk_dtype_size(dst->dtype));
  if (voxels(*dst) > 0 && s0 < d1 && d0 < s1)
    return fail(CLK_MEM_COPY_OVERLAP, std::string(kernel) + ": src and dst overlap");
  return CLK_SUCCESS;
}

int32_t separable(int op, const char* name, void* stream, const clk_image* dst, const clk_image* src,
                  int32_t dim, int32_t N, float s) {
  int32_t rc = check_pair(name, dst, src);
  if (rc) return rc;
  std::string err;
  rc = clk::launch_separable(op, (hipStream_t)stream, *dst, *src, dim, N, s, &err);
  if (rc) return fail(rc, std::string(name) + ": " + err);
  return ok();
}

// The 2D kernels (…_2d_x.cl) index (i, j) only: an image of depth 1.
int32_t check_2d(const char* kernel, const clk_image* dst, const clk_image* src) {
  if ((dst && dst->depth != 1) || (src && src->depth != 1))
    return fail(CLK_INVALID_VALUE, std::string(kernel) + ": 2D kernels take images of depth 1");
  return CLK_SUCCESS;
}

// gaussian/mean/minimum/maximum_separable_2d: the 3D pass on a W x H x 1
// volume.  dir = (dim==0, dim==1) in 2D and the z pass of a one-plane volume
// reads the centre at every tap, so every dim gives the 2D kernel's result;
// the 2D minimum/maximum skip v = 0 (…_separable_2d_x.cl:21), which cannot
// change the result: res starts at the centre and only ever decreases
// (increases), so min(res, centre) keeps res.
int32_t separable_2d(int op, const char* name, void* stream, const clk_image* dst, const clk_image* src,
                     int32_t dim, int32_t N, float s) {
  int32_t rc = check_pair(name, dst, src);
  if (rc) return rc;
  rc = check_2d(name, dst, src);
  if (rc) return rc;
  std::string err;
  rc = clk::launch_separable(op, (hipStream_t)stream, *dst, *src, dim, N, s, &err);
  if (rc) return fail(rc, std::string(name) + ": " + err);
  return ok();
}

// ---- host chain -------------------------------------------------------------

// Kernel size of a Gaussian sigma: int(8*sigma), made odd.  The reference
// hosts choose N from sigma; this is the CLIJ convention.
int32_t sigma_to_n(float sigma) {
  int32_t n = (int32_t)(sigma * 8.0f);
  if (n % 2 == 0) n += 1;
  return n;
}

struct PassSpec {
  bool active;
  int32_t N;
  float s;
};

// The temporaries a chain needs: their pixel type, how many (0..2), and
// the bytes of one (256-byte aligned).
struct ChainTemps {
  int nact = 0, tdt = CLK_FLOAT, ntmp = 0;
  size_t one = 0;
};

ChainTemps chain_temps(int op, const clk_image& dst, const clk_image& src, const PassSpec (&ps)[3]) {
  ChainTemps t;
  for (int a = 0; a < 3; ++a) t.nact += ps[a].active;
  // float for Gaussian and mean; the source type for min/max, whose values
  // are exact in any type that holds the source
  t.tdt = (op == clk::OP_MIN || op == clk::OP_MAX) ? src.dtype : CLK_FLOAT;
  t.ntmp = t.nact - 1 < 2 ? (t.nact > 0 ? t.nact - 1 : 0) : 2;
  t.one = ((size_t)voxels(dst) * clk_dtype_size(t.tdt) + 255) & ~(size_t)255;
  return t;
}

// Runs the active passes in axis order x, y, z through the temporaries at
// `blk` (t.ntmp of t.one bytes); an inactive axis is skipped (the next pass
// reads the previous image directly, which gives the same values as a float
// copy); with no active axis the result is a converting copy.
int32_t run_chain(int op, hipStream_t st, const clk_image& dst, const clk_image& src, const PassSpec (&ps)[3],
                  const ChainTemps& t, void* blk, bool allow_tune, std::string* err) {
  if (t.nact == 0) return clk::launch_convert_copy(st, dst, src, err);
  void* tmp[2] = {blk, t.ntmp > 1 ? static_cast<char*>(blk) + t.one : nullptr};
  clk_image cur = src;
  int done = 0, ti = 0;
  int32_t rc = 0;
  for (int a = 0; a < 3 && rc == 0; ++a) {
    if (!ps[a].active) continue;
    ++done;
    clk_image out = dst;
    if (done < t.nact) {
      out.dtype = t.tdt;
      out.ptr = tmp[ti];
      ti ^= 1;
    }
    rc = clk::launch_separable(op, st, out, cur, a, ps[a].N, ps[a].s, err, allow_tune);
    cur = out;
  }
  return rc;
}

// The host chain with its temporaries in the stream's kept blocks
// (clk::CallScratch).
int32_t chain3(int op, const char* name, void* stream, const clk_image* dst, const clk_image* src,
               const PassSpec (&ps)[3]) {
  int32_t rc = check_pair(name, dst, src);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  std::string err;
  const ChainTemps t = chain_temps(op, *dst, *src, ps);
  clk::CallScratch scratch(st);  // held until the last pass is issued
  void* blk = nullptr;
  if (t.ntmp > 0 && voxels(*dst) > 0 && scratch.get(0, t.ntmp * t.one, &blk) != hipSuccess)
    return fail(CLK_MEM_OBJECT_ALLOCATION_FAILURE, std::string(name) + ": temporary allocation failed");
  rc = run_chain(op, st, *dst, *src, ps, t, blk, true, &err);
  return rc ? fail(rc, std::string(name) + ": " + err) : ok();
}

int32_t box_chain(int op, const char* name, void* stream, const clk_image* dst, const clk_image* src,
                  int32_t rx, int32_t ry, int32_t rz) {
  if (rx < 0 || ry < 0 || rz < 0) return fail(CLK_INVALID_VALUE, std::string(name) + ": negative radius");
  const PassSpec ps[3] = {{rx > 0, 2 * rx + 1, 0.f}, {ry > 0, 2 * ry + 1, 0.f}, {rz > 0, 2 * rz + 1, 0.f}};
  return chain3(op, name, stream, dst, src, ps);
}

// ---- by-name dispatch ---------------------------------------------------------

struct NamedArgs {
  const clk_image* dst = nullptr;
  const clk_image* src = nullptr;
  bool has[5] = {false, false, false, false, false};  // dim/Nx, N/Ny, s/Nz
  int32_t iv[3] = {0, 0, 0};
  float fv = 0.f;
};

bool strip_program_name(const char* in, std::string* out) {
  std::string k(in);
  const std::string suf = "_x.cl";
  if (k.size() > suf.size() && k.compare(k.size() - suf.size(), suf.size(), suf) == 0)
    k = k.substr(0, k.size() - suf.size());
  *out = k;
  return true;
}

}  // namespace

extern "C" {

int32_t clk_dtype_size(int32_t dtype) {
  switch (dtype) {
    case CLK_CHAR: case CLK_UCHAR: return 1;
    case CLK_SHORT: case CLK_USHORT: return 2;
    case CLK_FLOAT: return 4;
  }
  return 0;
}

const char* clk_last_error_string(void) { return g_last_error.c_str(); }

int32_t clk_version(void) { return 100; }

int32_t clk_gaussian_blur_separable_3d(void* stream, const clk_image* dst, const clk_image* src,
                                       int32_t dim, int32_t N, float s) {
  return separable(clk::OP_GAUSS, "gaussian_blur_separable_3d", stream, dst, src, dim, N, s);
}

int32_t clk_mean_separable_3d(void* stream, const clk_image* dst, const clk_image* src, int32_t dim,
                              int32_t N, float s) {
  return separable(clk::OP_MEAN, "mean_separable_3d", stream, dst, src, dim, N, s);
}

int32_t clk_minimum_separable_3d(void* stream, const clk_image* dst, const clk_image* src,
                                 int32_t dim, int32_t N, float s) {
  return separable(clk::OP_MIN, "minimum_separable_3d", stream, dst, src, dim, N, s);
}

int32_t clk_maximum_separable_3d(void* stream, const clk_image* dst, const clk_image* src,
                                 int32_t dim, int32_t N, float s) {
  return separable(clk::OP_MAX, "maximum_separable_3d", stream, dst, src, dim, N, s);
}

// Float -> float medians: the NaN / -0 scan and fix-up after the median
// launch (median_exact.h), its brick flags in the stream's kept blocks.
// (ex, ey, ez) are the half widths (N - 1) / 2 of the launch.
static int32_t median_nanzero(const char* name, hipStream_t st, const clk_image& dst, const clk_image& src, int ex,
                              int ey, int ez, bool sphere) {
  if (src.dtype != CLK_FLOAT || dst.dtype != CLK_FLOAT || voxels(dst) == 0) return ok();
  clk::CallScratch scratch(st);  // held until the fix-up is issued
  void* flags = nullptr;
  if (scratch.get(0, (size_t)clk::nanzero_flag_bytes(dst), &flags) != hipSuccess)
    return fail(CLK_MEM_OBJECT_ALLOCATION_FAILURE, std::string(name) + ": temporary allocation failed");
  std::string err;
  const int32_t rc = clk::launch_median_nanzero(st, dst, src, ex, ey, ez, sphere, flags, &err);
  return rc ? fail(rc, std::string(name) + ": " + err) : ok();
}

int32_t clk_median_box_3d(void* stream, const clk_image* dst, const clk_image* src, int32_t Nx,
                          int32_t Ny, int32_t Nz) {
  int32_t rc = check_pair("median_box_3d", dst, src);
  if (rc) return rc;
  std::string err;
  rc = clk::launch_median((hipStream_t)stream, *dst, *src, Nx, Ny, Nz, &err);
  if (rc) return fail(rc, err);
  return median_nanzero("median_box_3d", (hipStream_t)stream, *dst, *src, (Nx - 1) / 2, (Ny - 1) / 2, (Nz - 1) / 2,
                        false);
}

int32_t clk_gaussian_blur_separable_2d(void* stream, const clk_image* dst, const clk_image* src,
                                       int32_t dim, int32_t N, float s) {
  return separable_2d(clk::OP_GAUSS, "gaussian_blur_separable_2d", stream, dst, src, dim, N, s);
}

int32_t clk_mean_separable_2d(void* stream, const clk_image* dst, const clk_image* src, int32_t dim,
                              int32_t N, float s) {
  return separable_2d(clk::OP_MEAN, "mean_separable_2d", stream, dst, src, dim, N, s);
}

int32_t clk_minimum_separable_2d(void* stream, const clk_image* dst, const clk_image* src,
                                 int32_t dim, int32_t N, float s) {
  return separable_2d(clk::OP_MIN, "minimum_separable_2d", stream, dst, src, dim, N, s);
}

int32_t clk_maximum_separable_2d(void* stream, const clk_image* dst, const clk_image* src,
                                 int32_t dim, int32_t N, float s) {
  return separable_2d(clk::OP_MAX, "maximum_separable_2d", stream, dst, src, dim, N, s);
}

int32_t clk_median_box_2d(void* stream, const clk_image* dst, const clk_image* src, int32_t Nx,
                          int32_t Ny) {
  int32_t rc = check_pair("median_box_2d", dst, src);
  if (rc) return rc;
  rc = check_2d("median_box_2d", dst, src);
  if (rc) return rc;
  std::string err;
  rc = clk::launch_median((hipStream_t)stream, *dst, *src, Nx, Ny, 1, &err, true);
  if (rc) return fail(rc, err);
  return median_nanzero("median_box_2d", (hipStream_t)stream, *dst, *src, (Nx - 1) / 2, (Ny - 1) / 2, 0, false);
}

static int32_t sphere(int op, const char* name, void* stream, const clk_image* dst, const clk_image* src,
                      int32_t Nx, int32_t Ny, int32_t Nz) {
  int32_t rc = check_pair(name, dst, src);
  if (rc) return rc;
  std::string err;
  rc = clk::launch_sphere(op, (hipStream_t)stream, *dst, *src, Nx, Ny, Nz, &err);
  if (rc) return fail(rc, std::string(name) + ": " + err);
  if (op == 3)
    return median_nanzero(name, (hipStream_t)stream, *dst, *src, (Nx - 1) / 2, (Ny - 1) / 2, (Nz - 1) / 2, true);
  return ok();
}

int32_t clk_minimum_sphere_3d(void* stream, const clk_image* dst, const clk_image* src, int32_t Nx, int32_t Ny,
                              int32_t Nz) {
  return sphere(0, "minimum_sphere_3d", stream, dst, src, Nx, Ny, Nz);
}

int32_t clk_maximum_sphere_3d(void* stream, const clk_image* dst, const clk_image* src, int32_t Nx, int32_t Ny,
                              int32_t Nz) {
  return sphere(1, "maximum_sphere_3d", stream, dst, src, Nx, Ny, Nz);
}

int32_t clk_mean_sphere_3d(void* stream, const clk_image* dst, const clk_image* src, int32_t Nx, int32_t Ny,
                           int32_t Nz) {
  return sphere(2, "mean_sphere_3d", stream, dst, src, Nx, Ny, Nz);
}

int32_t clk_median_sphere_3d(void* stream, const clk_image* dst, const clk_image* src, int32_t Nx, int32_t Ny,
                             int32_t Nz) {
  return sphere(3, "median_sphere_3d", stream, dst, src, Nx, Ny, Nz);
}

int32_t clk_median_slice_by_slice_box_3d(void* stream, const clk_image* dst, const clk_image* src, int32_t Nx,
                                         int32_t Ny) {
  int32_t rc = check_pair("median_slice_by_slice_box_3d", dst, src);
  if (rc) return rc;
  std::string err;
  // the 2D box per z slice, values stored by the same C cast as median_box_2d
  // (median_slice_by_slice_box_3d_x.cl:46)
  rc = clk::launch_median((hipStream_t)stream, *dst, *src, Nx, Ny, 1, &err, true);
  if (rc) return fail(rc, "median_slice_by_slice_box_3d: " + err);
  return median_nanzero("median_slice_by_slice_box_3d", (hipStream_t)stream, *dst, *src, (Nx - 1) / 2, (Ny - 1) / 2,
                        0, false);
}

int32_t clk_copy_3d(void* stream, const clk_image* dst, const clk_image* src) {
  int32_t rc = check_pair("copy_3d", dst, src);
  if (rc) return rc;
  std::string err;
  rc = clk::launch_convert_copy((hipStream_t)stream, *dst, *src, &err);  // copy_3d_x.cl:13-14
  if (rc) return fail(rc, "copy_3d: " + err);
  return ok();
}

int32_t clk_detect_maxima_3d(void* stream, const clk_image* src, const clk_image* dst) {
  int32_t rc = check_pair("detect_maxima_3d", dst, src);
  if (rc) return rc;
  std::string err;
  rc = clk::launch_consumer(true, (hipStream_t)stream, *dst, *src, &err);
  if (rc) return fail(rc, "detect_maxima_3d: " + err);
  return ok();
}

int32_t clk_maximum_z_projection(void* stream, const clk_image* dst_max, const clk_image* src) {
  int32_t rc = check_image(dst_max, "dst_max");
  if (rc) return rc;
  rc = check_image(src, "src");
  if (rc) return rc;
  if (dst_max->width != src->width || dst_max->height != src->height || dst_max->depth != 1)
    return fail(CLK_INVALID_VALUE, "maximum_z_projection: dst_max must be width x height x 1 of src");
  const uintptr_t s0 = (uintptr_t)src->ptr, d0 = (uintptr_t)dst_max->ptr;
  const uintptr_t s1 = s0 + (uintptr_t)(voxels(*src) * clk_dtype_size(src->dtype));
  const uintptr_t d1 = d0 + (uintptr_t)(voxels(*dst_max) * clk_dtype_size(dst_max->dtype));
  if (voxels(*dst_max) > 0 && voxels(*src) > 0 && s0 < d1 && d0 < s1)
    return fail(CLK_MEM_COPY_OVERLAP, "maximum_z_projection: src and dst_max overlap");
  std::string err;
  rc = clk::launch_consumer(false, (hipStream_t)stream, *dst_max, *src, &err);
  if (rc) return fail(rc, "maximum_z_projection: " + err);
  return ok();
}

int32_t clk_execute(void* stream, const char* kernel_name, const clk_arg* args, int32_t nargs) {
  if (!kernel_name) return fail(CLK_INVALID_KERNEL_NAME, "kernel name is NULL");
  if (nargs < 0 || (nargs > 0 && !args)) return fail(CLK_INVALID_KERNEL_ARGS, "invalid argument list");
  std::string k;
  strip_program_name(kernel_name, &k);
  int op = -1;
  bool two_d = false;
  std::string base = k;
  if (k.size() > 3 && k.compare(k.size() - 3, 3, "_2d") == 0) {
    two_d = true;
    base = k.substr(0, k.size() - 3) + "_3d";
  }
  if (base == "gaussian_blur_separable_3d") op = clk::OP_GAUSS;
  else if (base == "mean_separable_3d") op = clk::OP_MEAN;
  else if (base == "minimum_separable_3d") op = clk::OP_MIN;
  else if (base == "maximum_separable_3d") op = clk::OP_MAX;
  else if (base == "median_box_3d") op = 4;
  else if (!two_d) {
    // kernels with (dst, src[, Nx, Ny[, Nz]]): spheres, slice median, copy,
    // and the consumers (minimum_sphere_3d_x.cl:5-9,
    // median_slice_by_slice_box_3d_x.cl:24-28, copy_3d_x.cl:1-4,
    // detect_maxima_3d_x.cl:2-5, maximum_z_projection_x.cl:2-5)
    int kind = -1, nsc = 3;
    if (k == "minimum_sphere_3d") kind = 0;
    else if (k == "maximum_sphere_3d") kind = 1;
    else if (k == "mean_sphere_3d") kind = 2;
    else if (k == "median_sphere_3d") kind = 3;
    else if (k == "median_slice_by_slice_box_3d") { kind = 4; nsc = 2; }
    else if (k == "copy_3d") { kind = 5; nsc = 0; }
    else if (k == "detect_maxima_3d") { kind = 6; nsc = 0; }
    else if (k == "maximum_z_projection") { kind = 7; nsc = 0; }
    if (kind >= 0) {
      static const char* nn[3] = {"Nx", "Ny", "Nz"};
      const clk_image *d = nullptr, *sr = nullptr;
      int32_t v[3] = {0, 0, 0};
      bool has[3] = {false, false, false};
      for (int32_t i = 0; i < nargs; ++i) {
        const clk_arg& a = args[i];
        if (!a.name) return fail(CLK_INVALID_KERNEL_ARGS, k + ": argument without a name");
        const std::string nm(a.name);
        // maximum_z_projection names its output dst_max (maximum_z_projection_x.cl:3)
        const std::string dname = kind == 7 ? "dst_max" : "dst";
        if (nm == dname || nm == "src") {
          if (a.kind != CLK_ARG_IMAGE) return fail(CLK_INVALID_KERNEL_ARGS, k + ": " + nm + " must be an image");
          (nm == dname ? d : sr) = a.image;
          continue;
        }
        int slot = -1;
        for (int j = 0; j < nsc; ++j)
          if (nm == nn[j]) slot = j;
        if (slot < 0) return fail(CLK_INVALID_KERNEL_ARGS, k + ": unknown parameter " + nm);
        if (a.kind == CLK_ARG_IMAGE) return fail(CLK_INVALID_KERNEL_ARGS, k + ": " + nm + " must be a number");
        v[slot] = (a.kind == CLK_ARG_INT) ? a.i : (int32_t)a.f;
        has[slot] = true;
      }
      if (!d || !sr) return fail(CLK_INVALID_KERNEL_ARGS, k + ": both images are required");
      for (int j = 0; j < nsc; ++j)
        if (!has[j]) return fail(CLK_INVALID_KERNEL_ARGS, k + ": missing parameter " + nn[j]);
      if (kind <= 3) return sphere(kind, k.c_str(), stream, d, sr, v[0], v[1], v[2]);
      if (kind == 4) return clk_median_slice_by_slice_box_3d(stream, d, sr, v[0], v[1]);
      if (kind == 6) return clk_detect_maxima_3d(stream, sr, d);
      if (kind == 7) return clk_maximum_z_projection(stream, d, sr);
      return clk_copy_3d(stream, d, sr);
    }
  }
  if (op < 0) return fail(CLK_INVALID_KERNEL_NAME, "unknown kernel: " + k);
  // parameter names in declaration order (gaussian_blur_separable_3d_x.cl:8-11,
  // median_box_3d_x.cl:25-29; the 2D kernels: gaussian_blur_separable_2d_x.cl:8-12,
  // median_box_2d_x.cl:24-27, no Nz)
  static const char* sep_names[3] = {"dim", "N", "s"};
  static const char* med_names[3] = {"Nx", "Ny", "Nz"};
  const char** names = (op == 4) ? med_names : sep_names;
  const int nscalar = (op == 4 && two_d) ? 2 : 3;
  NamedArgs na;
  for (int32_t i = 0; i < nargs; ++i) {
    const clk_arg& a = args[i];
    if (!a.name) return fail(CLK_INVALID_KERNEL_ARGS, k + ": argument without a name");
    const std::string nm(a.name);
    if (nm == "dst" || nm == "src") {
      if (a.kind != CLK_ARG_IMAGE) return fail(CLK_INVALID_KERNEL_ARGS, k + ": " + nm + " must be an image");
      (nm == "dst" ? na.dst : na.src) = a.image;
      continue;
    }
    int slot = -1;
    for (int j = 0; j < nscalar; ++j)
      if (nm == names[j]) slot = j;
    if (slot < 0) return fail(CLK_INVALID_KERNEL_ARGS, k + ": unknown parameter " + nm);
    if (a.kind == CLK_ARG_IMAGE) return fail(CLK_INVALID_KERNEL_ARGS, k + ": " + nm + " must be a number");
    const bool is_float_param = (op != 4 && slot == 2);
    if (is_float_param) na.fv = (a.kind == CLK_ARG_FLOAT) ? a.f : (float)a.i;
    else na.iv[slot] = (a.kind == CLK_ARG_INT) ? a.i : (int32_t)a.f;
    na.has[slot] = true;
  }
  if (!na.dst || !na.src) return fail(CLK_INVALID_KERNEL_ARGS, k + ": dst and src are required");
  if (op == 4) {
    for (int j = 0; j < nscalar; ++j)
      if (!na.has[j]) return fail(CLK_INVALID_KERNEL_ARGS, k + ": missing parameter " + names[j]);
    if (two_d) return clk_median_box_2d(stream, na.dst, na.src, na.iv[0], na.iv[1]);
    return clk_median_box_3d(stream, na.dst, na.src, na.iv[0], na.iv[1], na.iv[2]);
  }
  if (!na.has[0] || !na.has[1])
    return fail(CLK_INVALID_KERNEL_ARGS, k + ": missing parameter dim or N");
  if (op == clk::OP_GAUSS && !na.has[2]) return fail(CLK_INVALID_KERNEL_ARGS, k + ": missing parameter s");
  if (two_d) return separable_2d(op, k.c_str(), stream, na.dst, na.src, na.iv[0], na.iv[1], na.fv);
  return separable(op, k.c_str(), stream, na.dst, na.src, na.iv[0], na.iv[1], na.fv);
}

int32_t clk_gaussian_blur_3d(void* stream, const clk_image* dst, const clk_image* src, float sx,
                             float sy, float sz) {
  const PassSpec ps[3] = {{sx > 0, sigma_to_n(sx), sx}, {sy > 0, sigma_to_n(sy), sy},
                          {sz > 0, sigma_to_n(sz), sz}};
  return chain3(clk::OP_GAUSS, "gaussian_blur_3d", stream, dst, src, ps);
}

int32_t clk_mean_box_3d(void* stream, const clk_image* dst, const clk_image* src, int32_t rx,
                        int32_t ry, int32_t rz) {
  return box_chain(clk::OP_MEAN, "mean_box_3d", stream, dst, src, rx, ry, rz);
}

int32_t clk_minimum_box_3d(void* stream, const clk_image* dst, const clk_image* src, int32_t rx,
                           int32_t ry, int32_t rz) {
  return box_chain(clk::OP_MIN, "minimum_box_3d", stream, dst, src, rx, ry, rz);
}

int32_t clk_maximum_box_3d(void* stream, const clk_image* dst, const clk_image* src, int32_t rx,
                           int32_t ry, int32_t rz) {
  return box_chain(clk::OP_MAX, "maximum_box_3d", stream, dst, src, rx, ry, rz);
}

int32_t clk_gaussian_blur_2d(void* stream, const clk_image* dst, const clk_image* src, float sx, float sy) {
  const int32_t rc = check_2d("gaussian_blur_2d", dst, src);
  if (rc) return rc;
  const PassSpec ps[3] = {{sx > 0, sigma_to_n(sx), sx}, {sy > 0, sigma_to_n(sy), sy}, {false, 1, 0.f}};
  return chain3(clk::OP_GAUSS, "gaussian_blur_2d", stream, dst, src, ps);
}

static int32_t box_chain_2d(int op, const char* name, void* stream, const clk_image* dst, const clk_image* src,
                            int32_t rx, int32_t ry) {
  const int32_t rc = check_2d(name, dst, src);
  if (rc) return rc;
  return box_chain(op, name, stream, dst, src, rx, ry, 0);
}

int32_t clk_mean_box_2d(void* stream, const clk_image* dst, const clk_image* src, int32_t rx, int32_t ry) {
  return box_chain_2d(clk::OP_MEAN, "mean_box_2d", stream, dst, src, rx, ry);
}

int32_t clk_minimum_box_2d(void* stream, const clk_image* dst, const clk_image* src, int32_t rx, int32_t ry) {
  return box_chain_2d(clk::OP_MIN, "minimum_box_2d", stream, dst, src, rx, ry);
}

int32_t clk_maximum_box_2d(void* stream, const clk_image* dst, const clk_image* src, int32_t rx, int32_t ry) {
  return box_chain_2d(clk::OP_MAX, "maximum_box_2d", stream, dst, src, rx, ry);
}

int32_t clk_batch_filter_3d(void* stream, const clk_filter_job* jobs, int32_t njobs) {
  if (njobs < 0 || (njobs > 0 && !jobs)) return fail(CLK_INVALID_VALUE, "clk_batch_filter_3d: invalid job list");
  for (int32_t j = 0; j < njobs; ++j) {
    const clk_filter_job& jb = jobs[j];
    const std::string tag = "clk_batch_filter_3d: job " + std::to_string(j);
    if (jb.kind < CLK_JOB_GAUSSIAN_BLUR_3D || jb.kind > CLK_JOB_MEDIAN_BOX_3D || jb.reserved != 0)
      return fail(CLK_INVALID_VALUE, tag + ": invalid kind");
    int32_t rc = check_pair(tag.c_str(), &jb.dst, &jb.src);
    if (rc) return rc;
    if (jb.kind != CLK_JOB_GAUSSIAN_BLUR_3D)
      for (int a = 0; a < 3; ++a)
        if (jb.radius[a] < 0) return fail(CLK_INVALID_VALUE, tag + ": negative radius");
    if (jb.kind == CLK_JOB_GAUSSIAN_BLUR_3D)
      for (int a = 0; a < 3; ++a)
        if (jb.sigma[a] > 0 && sigma_to_n(jb.sigma[a]) > CLK_GAUSSIAN_MAX_TAPS)
          return fail(CLK_INVALID_VALUE, tag + ": sigma too large");
  }
  std::string err;
  const int32_t rc = clk::run_batch((hipStream_t)stream, jobs, njobs, &err);
  return rc ? fail(rc, err) : ok();
}

int32_t clk_filter_3d_noalloc(void* stream, const clk_filter_job* job) {
  static const char* kName = "clk_filter_3d_noalloc";
  if (!job) return fail(CLK_INVALID_VALUE, std::string(kName) + ": job is NULL");
  if (job->kind < CLK_JOB_GAUSSIAN_BLUR_3D || job->kind > CLK_JOB_MEDIAN_BOX_3D || job->reserved != 0)
    return fail(CLK_INVALID_VALUE, std::string(kName) + ": invalid kind");
  int32_t rc = check_pair(kName, &job->dst, &job->src);
  if (rc) return rc;
  const clk_image& dst = job->dst;
  const clk_image& src = job->src;
  hipStream_t st = (hipStream_t)stream;
  std::string err;
  if (job->kind == CLK_JOB_MEDIAN_BOX_3D) {
    for (int a = 0; a < 3; ++a)
      if (job->radius[a] < 0) return fail(CLK_INVALID_VALUE, std::string(kName) + ": negative radius");
    rc = clk::launch_median(st, dst, src, 2 * job->radius[0] + 1, 2 * job->radius[1] + 1,
                            2 * job->radius[2] + 1, &err);
    if (rc) return fail(rc, std::string(kName) + ": " + err);
    if (src.dtype != CLK_FLOAT || dst.dtype != CLK_FLOAT || voxels(dst) == 0) return ok();
    // float -> float: the NaN / -0 fix-up, its brick flags in the fixed array
    const size_t need = (size_t)clk::nanzero_flag_bytes(dst);
    if (need > clk::kNoAllocScratchBytes)
      return fail(CLK_INVALID_VALUE, std::string(kName) + ": the median needs " + std::to_string(need) +
                                         " bytes of brick flags, more than the " +
                                         std::to_string(clk::kNoAllocScratchBytes) + " bytes of its fixed array");
    void* flags = nullptr;
    const hipError_t e = clk::noalloc_scratch(&flags);
    if (e != hipSuccess)
      return fail(CLK_OUT_OF_RESOURCES, std::string(kName) + ": device array not found: " + hipGetErrorString(e));
    rc = clk::launch_median_nanzero(st, dst, src, job->radius[0], job->radius[1], job->radius[2], false, flags, &err);
    return rc ? fail(rc, std::string(kName) + ": " + err) : ok();
  }
  PassSpec ps[3];
  const int op = job->kind;  // the separable CLK_JOB_* values equal the op codes
  for (int a = 0; a < 3; ++a) {
    if (op == clk::OP_GAUSS) {
      const float sg = job->sigma[a];
      ps[a] = {sg > 0, sigma_to_n(sg), sg};
      if (sg > 0 && ps[a].N > CLK_GAUSSIAN_MAX_TAPS)
        return fail(CLK_INVALID_VALUE, std::string(kName) + ": sigma too large");
    } else {
      const int32_t r = job->radius[a];
      if (r < 0) return fail(CLK_INVALID_VALUE, std::string(kName) + ": negative radius");
      ps[a] = {r > 0, 2 * r + 1, 0.f};
    }
  }
  const ChainTemps t = chain_temps(op, dst, src, ps);
  const size_t need = (size_t)t.ntmp * t.one;
  if (need > clk::kNoAllocScratchBytes)
    return fail(CLK_INVALID_VALUE, std::string(kName) + ": the filter needs " + std::to_string(need) +
                                       " bytes of temporaries, more than the " +
                                       std::to_string(clk::kNoAllocScratchBytes) + " bytes of its fixed array");
  void* blk = nullptr;
  if (need > 0 && voxels(dst) > 0) {
    const hipError_t e = clk::noalloc_scratch(&blk);
    if (e != hipSuccess)
      return fail(CLK_OUT_OF_RESOURCES, std::string(kName) + ": device array not found: " + hipGetErrorString(e));
  }
  rc = run_chain(op, st, dst, src, ps, t, blk, false, &err);
  return rc ? fail(rc, std::string(kName) + ": " + err) : ok();
}

int32_t clk_release_temporaries(void* stream) {
  const hipError_t e = clk::release_kept((hipStream_t)stream);
  if (e != hipSuccess)
    return fail(CLK_OUT_OF_RESOURCES, std::string("clk_release_temporaries: ") + hipGetErrorString(e));
  return ok();
}

int32_t clk_temporaries_held(void* stream, int64_t* stream_bytes, int64_t* pool_bytes) {
  if (stream_bytes) *stream_bytes = clk::kept_bytes((hipStream_t)stream);
  if (pool_bytes) *pool_bytes = clk::pool_reserved_bytes();
  return ok();
}

int32_t clk_init(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(CLK_OUT_OF_RESOURCES, "no HIP device");
  return ok();
}

int32_t clk_malloc(void** ptr, int64_t bytes) {
  if (!ptr || bytes < 0) return fail(CLK_INVALID_VALUE, "clk_malloc: invalid arguments");
  if (hipMalloc(ptr, (size_t)(bytes > 0 ? bytes : 1)) != hipSuccess)
    return fail(CLK_MEM_OBJECT_ALLOCATION_FAILURE, "clk_malloc: hipMalloc failed");
  return ok();
}

int32_t clk_free(void* ptr) {
  if (ptr && hipFree(ptr) != hipSuccess) return fail(CLK_INVALID_MEM_OBJECT, "clk_free: hipFree failed");
  return ok();
}

int32_t clk_memcpy_h2d(void* dst_dev, const void* src_host, int64_t bytes, void* stream) {
  if (bytes < 0 || (bytes > 0 && (!dst_dev || !src_host))) return fail(CLK_INVALID_VALUE, "clk_memcpy_h2d: invalid arguments");
  if (bytes == 0) return ok();
  if (hipMemcpyAsync(dst_dev, src_host, (size_t)bytes, hipMemcpyHostToDevice, (hipStream_t)stream) != hipSuccess ||
      hipStreamSynchronize((hipStream_t)stream) != hipSuccess)
    return fail(CLK_OUT_OF_RESOURCES, "clk_memcpy_h2d failed");
  return ok();
}

int32_t clk_memcpy_d2h(void* dst_host, const void* src_dev, int64_t bytes, void* stream) {
  if (bytes < 0 || (bytes > 0 && (!dst_host || !src_dev))) return fail(CLK_INVALID_VALUE, "clk_memcpy_d2h: invalid arguments");
  if (bytes == 0) return ok();
  if (hipMemcpyAsync(dst_host, src_dev, (size_t)bytes, hipMemcpyDeviceToHost, (hipStream_t)stream) != hipSuccess ||
      hipStreamSynchronize((hipStream_t)stream) != hipSuccess)
    return fail(CLK_OUT_OF_RESOURCES, "clk_memcpy_d2h failed");
  return ok();
}

int32_t clk_stream_create(void** stream) {
  if (!stream) return fail(CLK_INVALID_VALUE, "clk_stream_create: NULL");
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
    return fail(CLK_OUT_OF_RESOURCES, "clk_stream_create failed");
  *stream = (void*)s;
  return ok();
}

int32_t clk_stream_destroy(void* stream) {
  // its kept temporaries must not pass to a later stream with the same handle
  if (stream) (void)clk::release_kept((hipStream_t)stream);
  if (stream && hipStreamDestroy((hipStream_t)stream) != hipSuccess)
    return fail(CLK_INVALID_VALUE, "clk_stream_destroy failed");
  return ok();
}

int32_t clk_autotune_save(const char* path) {
  if (!path) return fail(CLK_INVALID_VALUE, "clk_autotune_save: NULL path");
  if (clk::autotune_save(path)) return fail(CLK_INVALID_VALUE, std::string("clk_autotune_save: cannot write ") + path);
  return ok();
}

int32_t clk_autotune_load(const char* path) {
  if (!path) return fail(CLK_INVALID_VALUE, "clk_autotune_load: NULL path");
  if (clk::autotune_load(path))
    return fail(CLK_INVALID_VALUE, std::string("clk_autotune_load: cannot read or parse ") + path);
  return ok();
}

int32_t clk_stream_sync(void* stream) {
  const hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) return fail(CLK_OUT_OF_RESOURCES, std::string("clk_stream_sync: ") + hipGetErrorString(e));
  return ok();
}

}  // extern "C"
